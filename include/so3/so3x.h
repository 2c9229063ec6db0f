/*
 * so3x.h -- extensions of the SO(3) C API for the MI355X build.
 *
 * These entry points have no counterpart in the reference; they add what a
 * GPU caller needs beside the drop-in functions of so3_core.h:
 * device-pointer transforms on a caller stream, ahead-of-time preparation of
 * the Delta tables and DFT plans, and diagnostics.  All return 0 on success
 * and a negative code on failure, with the message in so3x_last_error().
 *
 * Conventions: f and flmn are device pointers (HBM) unless stated; f is
 * double* when parameters->reality is set and double complex* otherwise;
 * `stream` is a hipStream_t (NULL = default stream).  Calls are asynchronous
 * with respect to the host and ordered on `stream`.
 */
#ifndef SO3X_H
#define SO3X_H

#include "so3_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { SO3X_VIA_SSHT = 0, SO3X_DIRECT = 1 };

/* Message of the last failure on the calling thread. */
const char *so3x_last_error(void);

/* Inverse transform flmn -> f on device memory (so3_core_inverse_via_ssht /
 * _direct semantics, selected by `method`). */
int so3x_inverse(void *f, const void *flmn, const so3_parameters_t *parameters, int method,
                 void *stream);
/* Forward transform f -> flmn on device memory. */
int so3x_forward(void *flmn, const void *f, const so3_parameters_t *parameters, int method,
                 void *stream);

/* Build (or fetch) the Delta tables and DFT plans for these parameters. */
int so3x_prepare(const so3_parameters_t *parameters);

/* Device bytes of workspace one transform with these parameters needs. */
long long so3x_workspace_bytes(const so3_parameters_t *parameters);

/* Free every cached table, plan and workspace. */
int so3x_release_all(void);

/* Batched 1-D DFTs of `howmany` contiguous sequences of length n (device
 * memory, in place), dir = -1 forward, +1 backward, unnormalised.  Exposes
 * the library's DFT kernels for testing. */
int so3x_dft(void *x, int n, int howmany, int dir, void *stream);

/* Stage profiler: enable (1) / disable (0) HIP-event timing of every stage
 * launch; so3x_stage_times fills up to `cap` entries (stage name, total
 * milliseconds, launch count), synchronising on the recorded events, and
 * returns the number of stages; so3x_stage_reset clears the totals. */
int so3x_stage_timing(int on);
int so3x_stage_times(const char **names, double *ms, int *count, int cap);
int so3x_stage_reset(void);

/* One-line description of the device and build ("gfx950, 256 CUs, ..."). */
int so3x_device_info(char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif
