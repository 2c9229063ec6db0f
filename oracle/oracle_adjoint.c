/*
 * oracle_adjoint.c -- CPU restatement of the adjoint direct transforms of
 * astro-informatics/so3 v1.3.7 (src/c/so3_adjoint.c).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the GPU library never loads it.
 *
 * The adjoint of the direct inverse maps samples to coefficients,
 *   flmn = (2l+1)/(8 pi^2) sum_{a,b,g} exp(-i m alpha) d^l_mn(beta) exp(-i n gamma) f,
 * in the stages of so3_adjoint.c:34-400: forward DFT over alpha and gamma,
 * zero-extension of beta to 2L-1 samples and a forward DFT over beta, the
 * phase exp(-i m' pi/(2L-1)), and the Delta contraction with i^{m-n}.
 * The adjoint of the direct forward maps coefficients to samples
 * (so3_adjoint.c:402-937): Delta contraction with i^{n-m} for m' >= 0,
 * m' < 0 by (-1)^{m+n} symmetry, the weight convolution (length 4L-3,
 * scale 4 pi^2/(4L-3)), the phase exp(+i m' pi/(2L-1)), a backward DFT over
 * m' divided by 2L-1, the adjoint of the beta extension
 * (F(b) += (-1)^{m+n} F(2L-2-b) for b <= L-2) and a backward DFT over alpha
 * and gamma divided by (2L-1)(2N-1).
 *
 * Real signals (so3_adjoint.c:939-1312, 1313-1841) keep n >= 0: the r2c
 * transform of the first adjoint gives the n >= 0 half of the complex one;
 * the c2r transform of the second takes the real part of the n = 0 term and
 * twice the real part of every n > 0 term.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_int.h"

#define AIX(n, m, mp) ((((long)((n) + N - 1) * M + ((m) + L - 1)) * M) + ((mp) + L - 1))

static void adjoint_inverse_core(double complex *flmn, const double complex *fc, const double *fr,
                                 const oracle_params *p) {
  oracle_h_check(p);
  const int L = p->L, N = p->N, L0 = p->L0, real = fr != NULL;
  const int M = 2 * L - 1, NG = 2 * N - 1;
  const long stride = (long)L * M; /* one gamma sample: L rows of 2L-1 */
  const int nlo = real ? 0 : -N + 1;
  double *dl = oracle_h_dl(p);
  oracle_fft_plan *pn = oracle_fft_plan_create(NG, -1);
  oracle_fft_plan *pm = oracle_fft_plan_create(M, -1);
  double complex *seq = (double complex *)malloc(sizeof(double complex) * (M > NG ? M : NG));
  /* forward DFT over gamma, then over alpha: E[n slot][b][m slot] */
  double complex *E = (double complex *)calloc((size_t)NG * stride, sizeof(double complex));
  for (long i = 0; i < stride; ++i) {
    for (int g = 0; g < NG; ++g) seq[g] = real ? fr[g * stride + i] : fc[g * stride + i];
    oracle_fft_exec(pn, seq, 1);
    for (int g = 0; g < NG; ++g) E[g * stride + i] = seq[g];
  }
  for (int s = 0; s < NG; ++s)
    for (int b = 0; b < L; ++b) oracle_fft_exec(pm, E + (long)s * stride + (long)b * M, 1);
  /* beta: zero-extension, forward DFT, phase -> T(n, m, m') */
  double complex *T = (double complex *)calloc((size_t)NG * M * M, sizeof(double complex));
  for (int n = nlo; n <= N - 1; ++n) {
    if (!oracle_h_n_active(n, N, p->n_mode)) continue;
    const int nsl = (n + NG) % NG;
    for (int m = -(L - 1); m <= L - 1; ++m) {
      for (int b = 0; b < M; ++b) seq[b] = b < L ? E[(long)nsl * stride + (long)b * M + (m + M) % M] : 0.0;
      oracle_fft_exec(pm, seq, 1);
      for (int mp = -(L - 1); mp <= L - 1; ++mp)
        T[AIX(n, m, mp)] = seq[(mp + M) % M] * cexp(-I * mp * ORACLE_PI / (2.0 * L - 1.0));
    }
  }
  /* every stored coefficient with l >= |n| is written (so3_adjoint.c:234-243) */
  for (int n = nlo; n <= N - 1; ++n)
    for (int el = abs(n); el < L; ++el)
      for (int m = -el; m <= el; ++m) flmn[oracle_h_ind(el, m, n, p, real)] = 0.0;
  int *ns = (int *)malloc(sizeof(int) * (2 * N + 2));
  for (int el = L0; el < L; ++el) {
    const int cnt = oracle_h_l_nset(el, p, real, ns);
    const double elf = (2.0 * el + 1.0) / (8.0 * ORACLE_PI * ORACLE_PI);
    for (int k = 0; k < cnt; ++k) {
      const int n = ns[k];
      for (int m = -el; m <= el; ++m) {
        double complex acc = 0.0;
        for (int mp = -el; mp <= el; ++mp)
          acc += oracle_h_dlv(dl, el, mp, m) * oracle_h_dlv(dl, el, mp, n) * T[AIX(n, m, mp)];
        flmn[oracle_h_ind(el, m, n, p, real)] += oracle_h_ipow(m - n) * elf * acc;
      }
    }
  }
  free(ns); free(T); free(E); free(seq); free(dl);
  oracle_fft_plan_destroy(pn);
  oracle_fft_plan_destroy(pm);
}

static void adjoint_forward_core(double complex *fc, double *fr, const double complex *flmn,
                                 const oracle_params *p) {
  oracle_h_check(p);
  const int L = p->L, N = p->N, L0 = p->L0, real = fr != NULL;
  const int M = 2 * L - 1, NG = 2 * N - 1, Pw = 4 * L - 3;
  const long stride = (long)L * M;
  const int nlo = real ? 0 : -N + 1;
  double *dl = oracle_h_dl(p);
  /* G(n, m, m') for 0 <= m' <= l (so3_adjoint.c:589-668) */
  double complex *G = (double complex *)calloc((size_t)NG * M * M, sizeof(double complex));
  int *ns = (int *)malloc(sizeof(int) * (2 * N + 2));
  for (int el = L0; el < L; ++el) {
    const int cnt = oracle_h_l_nset(el, p, real, ns);
    for (int k = 0; k < cnt; ++k) {
      const int n = ns[k];
      for (int m = -el; m <= el; ++m) {
        const double complex c = flmn[oracle_h_ind(el, m, n, p, real)] * oracle_h_ipow(n - m);
        for (int mp = 0; mp <= el; ++mp)
          G[AIX(n, m, mp)] += c * oracle_h_dlv(dl, el, mp, n) * oracle_h_dlv(dl, el, mp, m);
      }
    }
  }
  free(ns);
  /* weight spectrum: backward DFT of w(q) at q mod Pw */
  oracle_fft_plan *pwb = oracle_fft_plan_create(Pw, +1);
  oracle_fft_plan *pwf = oracle_fft_plan_create(Pw, -1);
  oracle_fft_plan *pm = oracle_fft_plan_create(M, +1);
  double complex *wr = (double complex *)calloc(Pw, sizeof(double complex));
  for (int q = -2 * (L - 1); q <= 2 * (L - 1); ++q) wr[(q + Pw) % Pw] = oracle_h_weight(q);
  oracle_fft_exec(pwb, wr, 1);
  double complex *cv = (double complex *)malloc(sizeof(double complex) * Pw);
  double complex *seq = (double complex *)malloc(sizeof(double complex) * M);
  /* per (n, m): symmetry, convolution, phase, backward DFT over m', adjoint
   * of the extension -> E[n slot][b][m slot] */
  double complex *E = (double complex *)calloc((size_t)NG * stride, sizeof(double complex));
  for (int n = nlo; n <= N - 1; ++n) {
    if (!oracle_h_n_active(n, N, p->n_mode)) continue;
    const int nsl = (n + NG) % NG;
    for (int m = -(L - 1); m <= L - 1; ++m) {
      const double s = ((m + n) & 1) ? -1.0 : 1.0;
      for (int mp = 1; mp <= L - 1; ++mp) G[AIX(n, m, -mp)] = s * G[AIX(n, m, mp)];
      memset(cv, 0, sizeof(double complex) * Pw);
      for (int mp = -(L - 1); mp <= L - 1; ++mp) cv[(mp + Pw) % Pw] = G[AIX(n, m, mp)];
      oracle_fft_exec(pwb, cv, 1);
      for (int r = 0; r < Pw; ++r) cv[r] *= wr[r];
      oracle_fft_exec(pwf, cv, 1);
      for (int mp = -(L - 1); mp <= L - 1; ++mp) {
        const double complex F = cv[(mp + Pw) % Pw] * (4.0 * ORACLE_PI * ORACLE_PI / (4.0 * L - 3.0));
        seq[(mp + M) % M] = F * cexp(I * mp * ORACLE_PI / (2.0 * L - 1.0)) / (2.0 * L - 1.0);
      }
      oracle_fft_exec(pm, seq, 1);
      for (int b = 0; b <= L - 2; ++b) seq[b] += s * seq[2 * L - 2 - b];
      for (int b = 0; b < L; ++b) E[(long)nsl * stride + (long)b * M + (m + M) % M] = seq[b];
    }
  }
  /* backward DFT over alpha then gamma, / ((2L-1)(2N-1)) */
  const double norm = 1.0 / (2.0 * L - 1.0) / (2.0 * N - 1.0);
  for (int s = 0; s < NG; ++s)
    for (int b = 0; b < L; ++b) oracle_fft_exec(pm, E + (long)s * stride + (long)b * M, 1);
  oracle_fft_plan *pn = oracle_fft_plan_create(NG, +1);
  double complex *col = (double complex *)malloc(sizeof(double complex) * NG);
  for (long i = 0; i < stride; ++i) {
    if (!real) {
      for (int g = 0; g < NG; ++g) col[g] = E[g * stride + i];
    } else {
      /* c2r: the real part of the n = 0 term, n > 0 with its conjugate */
      col[0] = creal(E[i]);
      for (int n = 1; n < N; ++n) { col[n] = E[n * stride + i]; col[NG - n] = conj(col[n]); }
    }
    oracle_fft_exec(pn, col, 1);
    for (int g = 0; g < NG; ++g) {
      if (real) fr[g * stride + i] = creal(col[g]) * norm;
      else fc[g * stride + i] = col[g] * norm;
    }
  }
  free(col); free(E); free(seq); free(cv); free(wr); free(G); free(dl);
  oracle_fft_plan_destroy(pwb);
  oracle_fft_plan_destroy(pwf);
  oracle_fft_plan_destroy(pm);
  oracle_fft_plan_destroy(pn);
}

#undef AIX

void oracle_adjoint_inverse_direct(double complex *flmn, const double complex *f, const oracle_params *p) {
  adjoint_inverse_core(flmn, f, NULL, p);
}
void oracle_adjoint_inverse_direct_real(double complex *flmn, const double *f, const oracle_params *p) {
  adjoint_inverse_core(flmn, NULL, f, p);
}
void oracle_adjoint_forward_direct(double complex *f, const double complex *flmn, const oracle_params *p) {
  adjoint_forward_core(f, NULL, flmn, p);
}
void oracle_adjoint_forward_direct_real(double *f, const double complex *flmn, const oracle_params *p) {
  adjoint_forward_core(NULL, f, flmn, p);
}
