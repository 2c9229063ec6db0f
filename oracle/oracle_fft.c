/*
 * oracle_fft.c -- CPU DFT for the so3 oracle.
 *
 * TEST INFRASTRUCTURE ONLY: this file belongs to the CPU oracle that checks
 * the GPU library.  Nothing in the shipped library calls it.
 *
 * The reference performs its DFTs with FFTW (so3_core.c:137-158, 333-364,
 * 554-576, 779-812, 1168-1206, 1335-1398, 1411-1487).  FFTW is not part of
 * the reference tree, so this restates the DFT it computes:
 *   X[k] = sum_j x[j] exp(sign * 2 pi i j k / n),  sign = -1 forward, +1 backward,
 * unnormalised, as FFTW defines FFTW_FORWARD / FFTW_BACKWARD.
 * Mixed-radix decimation in time; prime factors above 64 use Bluestein's
 * chirp-z algorithm with a power-of-two inner length.  Twiddles are computed
 * in long double and rounded once.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void fill_roots(double complex *w, int n, int sign) {
  for (int k = 0; k < n; ++k) {
    long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
    w[k] = (double)cosl(a) + sign * (double)sinl(a) * I;
  }
}

oracle_fft_plan *oracle_fft_plan_create(int n, int sign) {
  oracle_fft_plan *p = (oracle_fft_plan *)calloc(1, sizeof(*p));
  p->n = n;
  p->sign = sign;
  int m = n, nf = 0;
  int big_prime = 0;
  static const int pref[] = {4, 2, 3, 5, 7};
  for (int i = 0; i < 5; ++i)
    while (m % pref[i] == 0) { p->fac[nf++] = pref[i]; m /= pref[i]; }
  for (int f = 11; m > 1 && f * f <= m; f += 2)
    while (m % f == 0) { p->fac[nf++] = f; m /= f; }
  if (m > 1) { if (m > 64) big_prime = 1; else p->fac[nf++] = m; }
  p->nfac = nf;
  if (big_prime || n > 0x7fffffff / 4) {
    /* Bluestein for the whole length */
    int M = 1;
    while (M < 2 * n - 1) M *= 2;
    p->blue = 1;
    p->M = M;
    p->sub = oracle_fft_plan_create(M, -1);
    p->subinv = oracle_fft_plan_create(M, +1);
    p->chirp = (double complex *)malloc(sizeof(double complex) * n);
    double complex *w2 = (double complex *)malloc(sizeof(double complex) * 2 * n);
    fill_roots(w2, 2 * n, sign); /* exp(sign * pi i k / n) */
    for (long j = 0; j < n; ++j) p->chirp[j] = w2[(j * j) % (2 * n)];
    free(w2);
    /* b_j = conj(chirp_j) for |j| < n, circularly; store its FFT_M */
    p->bhat = (double complex *)calloc(M, sizeof(double complex));
    for (int j = 0; j < n; ++j) {
      p->bhat[j] = conj(p->chirp[j]);
      if (j) p->bhat[M - j] = conj(p->chirp[j]);
    }
    oracle_fft_exec(p->sub, p->bhat, 1);
    p->tw = NULL;
  } else {
    p->tw = (double complex *)malloc(sizeof(double complex) * (n > 0 ? n : 1));
    fill_roots(p->tw, n, sign);
  }
  /* per-plan work area: a plan is used by one thread at a time */
  p->scratch = (double complex *)malloc(sizeof(double complex) * 2 * (p->blue ? p->M : n) + 16);
  return p;
}

void oracle_fft_plan_destroy(oracle_fft_plan *p) {
  if (!p) return;
  if (p->blue) {
    oracle_fft_plan_destroy(p->sub);
    oracle_fft_plan_destroy(p->subinv);
    free(p->chirp);
    free(p->bhat);
  }
  free(p->tw);
  free(p->scratch);
  free(p);
}

/* out[0..n) = DFT of in[0, s, 2s, ...] ; twiddle table tw has length N with
 * step N/n between the roots of this sub-transform. */
static void rec(const oracle_fft_plan *p, const double complex *in, long s,
                double complex *out, int n, int fi, int step) {
  if (n == 1) { out[0] = in[0]; return; }
  const int r = p->fac[fi], m = n / r;
  for (int q = 0; q < r; ++q) rec(p, in + q * s, s * r, out + q * m, m, fi + 1, step * r);
  double complex t[64], u[64];
  const int N = p->n;
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < r; ++q) t[q] = out[q * m + k] * p->tw[((long)q * k * step) % N];
    for (int pp = 0; pp < r; ++pp) {
      double complex acc = 0;
      for (int q = 0; q < r; ++q) acc += t[q] * p->tw[((long)((pp * q) % r) * (N / r))];
      u[pp] = acc;
    }
    for (int pp = 0; pp < r; ++pp) out[k + pp * m] = u[pp];
  }
}

void oracle_fft_exec(const oracle_fft_plan *p, double complex *x, long stride) {
  const int n = p->n;
  if (n <= 1) return;
  if (p->blue) {
    const int M = p->M;
    double complex *a = p->scratch;
    for (int j = 0; j < n; ++j) a[j] = x[j * stride] * p->chirp[j];
    for (int j = n; j < M; ++j) a[j] = 0;
    oracle_fft_exec(p->sub, a, 1);
    for (int j = 0; j < M; ++j) a[j] *= p->bhat[j];
    oracle_fft_exec(p->subinv, a, 1);
    for (int k = 0; k < n; ++k) x[k * stride] = a[k] * p->chirp[k] / (double)M;
    return;
  }
  double complex *in = p->scratch, *out = p->scratch + n;
  for (int j = 0; j < n; ++j) in[j] = x[j * stride];
  rec(p, in, 1, out, n, 0, 1);
  for (int j = 0; j < n; ++j) x[j * stride] = out[j];
}

/* ctypes-facing helper: one DFT of a contiguous array */
void oracle_dft(double complex *x, int n, int sign) {
  oracle_fft_plan *p = oracle_fft_plan_create(n, sign);
  oracle_fft_exec(p, x, 1);
  oracle_fft_plan_destroy(p);
}
