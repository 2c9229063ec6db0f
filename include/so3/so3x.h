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
 * _direct semantics, selected by `method`).  Single transforms share one
 * workspace per device; calls on different streams are ordered on it by
 * events (each waits for the previous call, whatever its stream), so they
 * are safe but do not overlap.  Use a batch (so3x_batch) for overlap. */
int so3x_inverse(void *f, const void *flmn, const so3_parameters_t *parameters, int method,
                 void *stream);
/* Forward transform f -> flmn on device memory. */
int so3x_forward(void *flmn, const void *f, const so3_parameters_t *parameters, int method,
                 void *stream);

/* Allocation-free inverse.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same result as
 * so3x_inverse, but it calls no hipMalloc, hipFree or other allocating
 * routine: its workspace is a static device array of 64 MiB per GPU (held
 * until the process ends), and the parameters must have been prepared on
 * this device before (so3x_prepare, or one so3x_inverse call with them).
 * Fails with a message when the transform's workspace exceeds 64 MiB (about
 * L = N = 80 complex), or when the plan or Delta table is not resident.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (different streams without ordering between them) race. */
int so3x_inverse_noalloc(void *f, const void *flmn, const so3_parameters_t *parameters, int method,
                         void *stream);

/* The MATLAB sample layout of the reference's mex interface
 * (src/matlab/so3_inverse_mex.c:310-339, so3_forward_mex.c:78-97):
 * f_m[a nG nB + b nG + g] (gamma fastest) with real and imaginary parts in
 * separate arrays, against the library's f[g nA nB + b nA + a].  Device
 * memory, asynchronous on `stream`; im is NULL for real signals
 * (parameters->reality = 1); from_mex takes im = NULL as zero imaginary
 * parts.  nA, nB, nG are so3_sampling_nalpha / nbeta / ngamma. */
int so3x_to_mex_layout(double *re, double *im, const void *f, const so3_parameters_t *parameters, void *stream);
int so3x_from_mex_layout(void *f, const double *re, const double *im, const so3_parameters_t *parameters,
                         void *stream);

/* Build (or fetch) the Delta tables and DFT plans for these parameters. */
int so3x_prepare(const so3_parameters_t *parameters);

/* Autotuner keyed on (L, N, reality, sampling).  so3x_tune times, on
 * `stream`, the choices the library can make for these parameters -- the
 * workgroup order of the inverse and forward contractions, the length Pw of
 * the forward transform's weight convolution (every 2^a or 3 * 2^a from
 * 4L - 3 up to twice the default, which decides the FFT engine and plan that
 * run it), and for gamma lengths 2N - 1 in [65, 127] the engine of the gamma
 * DFTs (1: the workgroup engine, 2: the 256-point register FFT; shorter
 * lengths take a direct DFT) -- and keeps
 * the fastest for every later transform with the same (L, N, reality,
 * sampling), single or batched.  It allocates and synchronises; call it
 * outside timed regions.  choice (4 ints, may be NULL) receives (inverse
 * order, forward order, Pw, gamma engine; 0 where the gamma engine is not a
 * choice), ms (4 doubles, may be NULL) the best inverse-contraction,
 * forward-contraction, forward-beta and inverse+forward gamma times (0 where
 * not timed).  The contraction order never changes a result; another Pw or
 * gamma engine changes only the rounding (the convolution is exact for every
 * candidate).  so3x_tune_get reports the choices in force; so3x_tune_save
 * writes the table as text ("L N reality mw_ss corder_inv corder_fwd pw
 * gamma" per line, 0 for the default rule) and so3x_tune_load reads one back
 * (lines without the gamma column keep its default rule; returns the number
 * of entries, merged into the table in force). */
int so3x_tune(const so3_parameters_t *parameters, int method, void *stream, int *choice, double *ms);
int so3x_tune_get(const so3_parameters_t *parameters, int method, int *choice);
int so3x_tune_save(const char *path);
int so3x_tune_load(const char *path);
/* Forget every tuned choice (back to the default rules). */
int so3x_tune_clear(void);

/* Device bytes of workspace one transform with these parameters needs. */
long long so3x_workspace_bytes(const so3_parameters_t *parameters);

/* Free every cached table, plan and workspace. */
int so3x_release_all(void);

/* Upper bound on the bytes of Delta tables kept resident (default 32 GiB);
 * building a table beyond it waits for the device and evicts the least
 * recently used tables. */
int so3x_cache_limit(long long bytes);

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

/* ------------------------------------------------------------------ batches
 * Many transforms of different sizes in one call.  The items run on a set
 * of `lanes` HIP streams (default priority) forked from `stream` and joined
 * back into it, each lane with its own workspace, larger items first
 * (longest-processing-time assignment on a cost model); the call returns
 * once everything is enqueued.  lanes <= 0 selects the lane count found best
 * for this batch's shape by so3x_batch_tune (4 when never tuned).  Items must
 * not alias each other's outputs.  Plans, tables and lane workspaces are
 * built during the call when missing (a prepared batch allocates nothing).
 */
enum {
  SO3X_ITEM_INVERSE = 0,        /* flmn_in -> f */
  SO3X_ITEM_FORWARD = 1,        /* f -> flmn_out */
  SO3X_ITEM_PAIR = 2,           /* flmn_in -> f -> flmn_out */
  SO3X_ITEM_S2TOSO3_INVERSE = 3 /* (flm, glm) -> flmn_out = h_lmn -> f */
};

typedef struct {
  so3_parameters_t parameters; /* of the SO(3) transform (h for s2toso3) */
  int kind;                    /* SO3X_ITEM_* */
  int method;                  /* SO3X_VIA_SSHT or SO3X_DIRECT */
  void *f;                     /* signal on the SO(3) grid */
  const void *flmn_in;         /* coefficients in (INVERSE, PAIR) or flm (S2TOSO3) */
  void *flmn_out;              /* coefficients out (FORWARD, PAIR) or h_lmn (S2TOSO3) */
  const void *glm;             /* S2TOSO3 only: the second S^2 signal's coefficients */
} so3x_item_t;

int so3x_batch(const so3x_item_t *items, int count, int lanes, void *stream);

/* Batch modes.  LANES: every item runs whole on one of the lanes (above).
 * PERSISTENT: the batch runs in three phases -- the inverse contractions of
 * all items in one persistent launch per signal type (complex / real), whose
 * workgroups pull (item, block) work from device queues (one per XCD, with
 * stealing), then each item's DFT stages on the lanes, then the forward
 * contractions of all items in one such launch.  Each item then has its own
 * workspace (summed over the items; batches beyond 64 GiB of workspaces run
 * in consecutive groups).  Results are bit-identical in both modes.  AUTO:
 * the mode and lane count so3x_batch_tune chose for this batch's shape when
 * lanes <= 0 (LANES with 4 lanes when never tuned), LANES when lanes > 0. */
enum { SO3X_BATCH_AUTO = 0, SO3X_BATCH_LANES = 1, SO3X_BATCH_PERSISTENT = 2 };
int so3x_batch_persistent(const so3x_item_t *items, int count, int lanes, void *stream);
/* The mode and lane count in force for this batch's shape (tuned or default). */
int so3x_batch_choice(const so3x_item_t *items, int count, int *mode, int *lanes);

/* Time this batch (already-prepared items) in both modes with 1, 2, 4 and 8
 * lanes on `stream` and remember the fastest mode and lane count for batches
 * of the same shape (the multiset of item kinds and parameter blocks).  Runs
 * the batch several times, so outputs are overwritten.  Returns the chosen
 * lane count (so3x_batch_choice reports the mode). */
int so3x_batch_tune(const so3x_item_t *items, int count, void *stream);

/* Capture a batch into a HIP graph (the items are prepared first, outside
 * the capture); launch it on any stream as often as wanted; destroy it.
 * The graph holds the device pointers of the items, which must stay valid
 * until it is destroyed.  It owns its workspace and keeps its Delta tables
 * resident (so3x_cache_limit never evicts them); launches of one graph run
 * one after another, whatever streams they are made on. */
int so3x_graph_create(const so3x_item_t *items, int count, int lanes, long long *graph);
/* As so3x_graph_create, in the given batch mode (SO3X_BATCH_*). */
int so3x_graph_create_mode(const so3x_item_t *items, int count, int lanes, int mode, long long *graph);
int so3x_graph_launch(long long graph, void *stream);
int so3x_graph_destroy(long long graph);

/* Completion tickets for asynchronous callers (the serving loop): a ticket
 * marks the current end of `stream`; so3x_ticket_done returns 1 once every
 * operation enqueued before it has finished, 0 before; so3x_ticket_wait
 * blocks until then and releases the ticket (done tickets must be released
 * with so3x_ticket_wait as well). */
int so3x_ticket(void *stream, long long *ticket);
int so3x_ticket_done(long long ticket);
int so3x_ticket_wait(long long ticket);

#ifdef __cplusplus
}
#endif
#endif
