/*
 * oracle.h -- CPU oracle for the so3 MI355X library.
 *
 * TEST INFRASTRUCTURE ONLY.  These functions restate the reference
 * algorithms (astro-informatics/so3 v1.3.7, src/c/so3_core.c,
 * src/c/so3_adjoint.c, src/c/so3_conv.c, and the SSHT spin transforms it
 * calls, as specified by the prototypes/ssht MATLAB files) in plain C so that the GPU
 * library can be checked against them.  Only the tests, the smoke check and
 * the CPU-baseline leg of the benchmark may load this code.
 */
#ifndef SO3_ORACLE_H
#define SO3_ORACLE_H

#include <complex.h>

typedef struct oracle_fft_plan {
  int n, sign;
  int nfac, fac[64];
  double complex *tw;
  int blue, M;
  struct oracle_fft_plan *sub, *subinv;
  double complex *chirp, *bhat;
  double complex *scratch;
} oracle_fft_plan;

oracle_fft_plan *oracle_fft_plan_create(int n, int sign);
void oracle_fft_plan_destroy(oracle_fft_plan *p);
void oracle_fft_exec(const oracle_fft_plan *p, double complex *x, long stride);
void oracle_dft(double complex *x, int n, int sign);

/* parameter block with the reference's field order (so3_types.h:130-201) */
typedef struct {
  int verbosity, reality, L0, L, N, sampling_scheme, n_order, storage, n_mode,
      dl_method, steerable;
} oracle_params;

#endif
