#define _POSIX_C_SOURCE 199309L
/*
 * oracle_so3.c -- CPU restatement of the so3 Wigner transforms.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the CPU
 * baseline of bench.py, never by the library.  Written from the reference's
 * algorithm descriptions, not copied:
 *   - index maps            : src/c/so3_sampling.c:316-650
 *   - Delta^l(pi/2) tables   : Risbo (1996) recursion and Trapani & Navaza (2006)
 *                             top-row + three-term recursion, as used by the
 *                             reference through SSHT (so3_core.c:983-1035)
 *   - spin SHT (SSHT) MW/MW_SS inverse/forward by separation of variables:
 *                             prototypes/ssht/inverse_sov_fft_space.m:28-52,
 *                             prototypes/ssht/forward_sov_for_fft_space.m:26-88
 *   - via-SSHT transforms    : src/c/so3_core.c:51-897
 *   - direct transforms      : src/c/so3_core.c:914-2385
 *   - adjoint transforms     : src/c/so3_adjoint.c:34-1841
 *   - convolutions           : src/c/so3_conv.c:44-90, 186-215
 * Conventions: Delta^l_{m'm} = d^l_{m'm}(pi/2) with
 *   d^l_{mn}(b) = i^{n-m} sum_{m'} Delta^l_{m'm} Delta^l_{m'n} exp(i m' b),
 * f(a,b,g) = sum (2l+1)/(8 pi^2) f_lmn exp(i m a) d^l_{mn}(b) exp(i n g).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define OPI 3.141592653589793238462643383279502884197

enum { MW = 0, MWSS = 1 };
enum { PADDED = 0, COMPACT = 1 };
enum { ZERO_FIRST = 0, NEG_FIRST = 1 };
enum { MODE_ALL = 0, MODE_EVEN, MODE_ODD, MODE_MAX, MODE_L };
enum { RISBO = 0, TRAPANI = 1 };

static void die(const char *msg) {
  printf("ERROR: %s.\n", msg);
  fflush(stdout);
  exit(1);
}

/* ------------------------------------------------------------------ */
/* index maps (so3_sampling.c:316-650)                                 */
/* ------------------------------------------------------------------ */
static int o_nalpha(const oracle_params *p) { return p->sampling_scheme == MW ? 2 * p->L - 1 : 2 * p->L; }
static int o_nbeta(const oracle_params *p) { return p->sampling_scheme == MW ? p->L : p->L + 1; }
static int o_ngamma(const oracle_params *p) { return p->steerable ? p->N : 2 * p->N - 1; }

int oracle_flmn_size(const oracle_params *p) {
  const int L = p->L, N = p->N;
  if (p->storage == PADDED) return p->reality ? N * L * L : (2 * N - 1) * L * L;
  return p->reality ? N * (6 * L * L - (N - 1) * (2 * N - 1)) / 6
                    : (2 * N - 1) * (3 * L * L - N * (N - 1)) / 3;
}
int oracle_f_size(const oracle_params *p) { return o_nalpha(p) * o_nbeta(p) * o_ngamma(p); }

/* start of the (l=0 or l=|n|) block of n in the storage, so that
 * index(l,m,n) = block_base(n) + l*l + l + m  for both storages */
static long o_block_base(int n, int order, const oracle_params *p) {
  const long L2 = (long)p->L * p->L;
  const int N = p->N, an = abs(n);
  if (p->storage == PADDED) {
    if (order == ZERO_FIRST) return ((n < 0) ? -2L * n - 1 : 2L * n) * L2;
    return (long)(N - 1 + n) * L2;
  }
  long off;
  if (order == ZERO_FIRST) {
    off = (long)(2 * an - 1) * (3 * L2 - an * (an - 1)) / 3;
    if (n >= 0) off += L2 - (long)n * n;
  } else {
    off = (long)(N - 1 + n) * L2 - (long)(2 * N - 1) * (N - 1) * N / 6;
    if (n <= 0) off += (long)an * (2 * an + 1) * (an + 1) / 6;
    else off -= (long)an * (2 * an - 1) * (an - 1) / 6;
  }
  return off - (long)n * n;
}
int oracle_elmn2ind(int el, int m, int n, const oracle_params *p) {
  if (p->storage == COMPACT && abs(n) > el) die("Tried to access component with n > l in compact storage");
  return (int)(o_block_base(n, p->n_order, p) + (long)el * el + el + m);
}
int oracle_elmn2ind_real(int el, int m, int n, const oracle_params *p) {
  if (p->storage == COMPACT && abs(n) > el) die("Tried to access component with n > l in compact storage");
  return (int)(o_block_base(n, NEG_FIRST, p) + (long)el * el + el + m - o_block_base(0, NEG_FIRST, p));
}
static int o_ind(int el, int m, int n, const oracle_params *p, int real) {
  return real ? oracle_elmn2ind_real(el, m, n, p) : oracle_elmn2ind(el, m, n, p);
}

/* is block n processed for this n_mode (via-SSHT skip rule, so3_core.c:166-170) */
static int o_n_active(int n, int N, int mode) {
  if (mode == MODE_EVEN && (n % 2)) return 0;
  if (mode == MODE_ODD && !(n % 2)) return 0;
  if (mode == MODE_MAX && abs(n) < N - 1) return 0;
  return 1;
}

/* ------------------------------------------------------------------ */
/* Delta^l(pi/2) quarter tables                                        */
/* layout: out[off(l) + mp*(l+1) + m], off(l) = sum_{k<l} (k+1)^2       */
/* ------------------------------------------------------------------ */
long oracle_dl_offset(int l) { return (long)l * (l + 1) * (2 * l + 1) / 6; }

static void dl_risbo(int L, double *out) {
  /* full (2l+1)^2 table; rows i = l + m', cols k = l + m.  Each half step
   * grows the table by one (Risbo 1996, spinor-product recursion) */
  const double p = sin(OPI / 4), q = cos(OPI / 4);
  const int maxn = 2 * L + 1;
  double *d = (double *)calloc((size_t)maxn * maxn, sizeof(double));
  double *dd = (double *)calloc((size_t)maxn * maxn, sizeof(double));
  d[0] = 1.0;
  int n = 1; /* current table size = 2l+1 */
  for (int l = 0; l < L; ++l) {
    if (l > 0) {
      for (int half = 0; half < 2; ++half) {
        const int n1 = n + 1;
        for (int R = 0; R < n1; ++R)
          for (int K = 0; K < n1; ++K) {
            double acc = 0.0;
            if (R < n && K < n) acc += sqrt((double)(n - R) * (n - K)) * q * d[R * n + K];
            if (R > 0 && K < n) acc -= sqrt((double)R * (n - K)) * p * d[(R - 1) * n + K];
            if (R < n && K > 0) acc += sqrt((double)(n - R) * K) * p * d[R * n + K - 1];
            if (R > 0 && K > 0) acc += sqrt((double)R * K) * q * d[(R - 1) * n + K - 1];
            dd[R * n1 + K] = acc / n;
          }
        double *t = d; d = dd; dd = t;
        n = n1;
      }
    }
    double *o = out + oracle_dl_offset(l);
    for (int mp = 0; mp <= l; ++mp)
      for (int m = 0; m <= l; ++m) o[mp * (l + 1) + m] = d[(l + mp) * n + (l + m)];
  }
  free(d);
  free(dd);
}

static void dl_trapani(int L, double *out) {
  /* top row Delta^l_{l,m} from Delta^{l-1}_{l-1,.} (Trapani & Navaza 2006,
   * eqs 9-10), then for each column m the three-term recursion in m'
   *   sqrt((l-m')(l+m'+1)) D_{m'+1,m} + sqrt((l+m')(l-m'+1)) D_{m'-1,m} = 2m D_{m',m}
   * run from m' = l down to m' = m (the eighth m' >= m), and the other
   * eighth from D_{m',m} = (-1)^{m'-m} D_{m,m'}. */
  double *top = (double *)calloc(L + 1, sizeof(double));
  double *prev = (double *)calloc(L + 1, sizeof(double));
  for (int l = 0; l < L; ++l) {
    if (l == 0) top[0] = 1.0;
    else {
      top[0] = -sqrt((2.0 * l - 1.0) / (2.0 * l)) * prev[0];
      for (int m = 1; m <= l; ++m)
        top[m] = sqrt(l * (2.0 * l - 1.0) / (2.0 * (l + m) * (l + m - 1.0))) * prev[m - 1];
    }
    double *o = out + oracle_dl_offset(l);
    const int s = l + 1;
    for (int m = 0; m <= l; ++m) {
      double nxt = 0.0, cur = top[m];
      o[l * s + m] = cur;
      for (int mp = l; mp > m; --mp) {
        double v = (2.0 * m * cur - sqrt((double)(l - mp) * (l + mp + 1)) * nxt) /
                   sqrt((double)(l + mp) * (l - mp + 1));
        o[(mp - 1) * s + m] = v;
        nxt = cur;
        cur = v;
      }
    }
    for (int m = 0; m <= l; ++m)
      for (int mp = 0; mp < m; ++mp) o[mp * s + m] = ((m - mp) & 1 ? -1.0 : 1.0) * o[m * s + mp];
    memcpy(prev, top, sizeof(double) * (L + 1));
  }
  free(top);
  free(prev);
}

void oracle_dl_quarter(int L, int method, double *out) {
  if (method == RISBO) dl_risbo(L, out);
  else if (method == TRAPANI) dl_trapani(L, out);
  else die("Invalid dl method");
}

/* Delta^l_{mp,m} for any signs from the quarter table:
 *   D_{m',-m} = (-1)^{l+m'} D_{m',m},  D_{-m',m} = (-1)^{l+m} D_{m',m} */
static inline double dlv(const double *dl, int l, int mp, int m) {
  double s = 1.0;
  if (mp < 0) { mp = -mp; if ((l + m) & 1) s = -s; }
  if (m < 0) { m = -m; if ((l + mp) & 1) s = -s; }
  return s * dl[oracle_dl_offset(l) + (long)mp * (l + 1) + m];
}

static inline double complex ipow(int k) { /* i^k */
  switch (((k % 4) + 4) % 4) {
  case 0: return 1.0;
  case 1: return I;
  case 2: return -1.0;
  default: return -I;
  }
}

/* ------------------------------------------------------------------ */
/* spin spherical-harmonic transforms (restating SSHT's sov algorithms) */
/* ------------------------------------------------------------------ */
typedef struct {
  int L, ss, nA, nB, P, Pw;
  oracle_fft_plan *pa_b, *pa_f, *pb_b, *pb_f, *pw_b, *pw_f;
  double complex *wr; /* IFFT of the quadrature weights, FFT order */
  double complex *F, *seq, *tmp;
} ssht_ws;

/* so3_sampling_weight (so3_sampling.c:32-51) */
static double complex o_weight(int p) {
  if (p == 1) return -I * (OPI / 2);
  if (p == -1) return I * (OPI / 2);
  if (p % 2 == 0) return 2.0 / (1.0 - (double)p * p);
  return 0.0;
}

static ssht_ws *ssht_ws_create(int L, int ss) {
  ssht_ws *w = (ssht_ws *)calloc(1, sizeof(*w));
  w->L = L; w->ss = ss;
  w->nA = ss ? 2 * L : 2 * L - 1;
  w->nB = ss ? L + 1 : L;
  w->P = ss ? 2 * L : 2 * L - 1;
  w->Pw = 4 * L - 3;
  w->pa_b = oracle_fft_plan_create(w->nA, +1);
  w->pa_f = oracle_fft_plan_create(w->nA, -1);
  w->pb_b = oracle_fft_plan_create(w->P, +1);
  w->pb_f = oracle_fft_plan_create(w->P, -1);
  w->pw_b = oracle_fft_plan_create(w->Pw, +1);
  w->pw_f = oracle_fft_plan_create(w->Pw, -1);
  w->wr = (double complex *)calloc(w->Pw, sizeof(double complex));
  /* weights w(p), p in [-2(L-1), 2(L-1)], placed at p mod Pw, then IFFT
   * (so3_core.c:1407-1436) */
  for (int p = -2 * (L - 1); p <= 2 * (L - 1); ++p) w->wr[(p + w->Pw) % w->Pw] = o_weight(p);
  oracle_fft_exec(w->pw_b, w->wr, 1);
  const int big = (w->Pw > w->nA ? w->Pw : w->nA) + 8;
  w->F = (double complex *)calloc((size_t)(2 * L - 1) * (2 * L), sizeof(double complex));
  w->seq = (double complex *)calloc(big, sizeof(double complex));
  w->tmp = (double complex *)calloc((size_t)(2 * L) * (L + 2), sizeof(double complex));
  return w;
}
static void ssht_ws_destroy(ssht_ws *w) {
  oracle_fft_plan_destroy(w->pa_b); oracle_fft_plan_destroy(w->pa_f);
  oracle_fft_plan_destroy(w->pb_b); oracle_fft_plan_destroy(w->pb_f);
  oracle_fft_plan_destroy(w->pw_b); oracle_fft_plan_destroy(w->pw_f);
  free(w->wr); free(w->F); free(w->seq); free(w->tmp); free(w);
}

/* f[b*nA + a] = sum_lm flm sY_lm(theta_b, phi_a), sY_lm = (-1)^s
 * sqrt((2l+1)/4pi) exp(i m phi) d^l_{m,-s}(theta); flm index l*l+l+m */
static void ssht_inverse(ssht_ws *w, double complex *f, const double complex *flm, int L0,
                         int spin, const double *dl) {
  const int L = w->L, nA = w->nA, nB = w->nB, P = w->P;
  const int M = 2 * L - 1;
  double complex *F = w->F; /* F[(m+L-1)*L + mp] */
  memset(F, 0, sizeof(double complex) * M * L);
  const int l0 = L0 > abs(spin) ? L0 : abs(spin);
  for (int l = l0; l < L; ++l) {
    const double sq = sqrt((2.0 * l + 1.0) / (4.0 * OPI));
    for (int m = -l; m <= l; ++m) {
      const double complex c = flm[l * l + l + m] * sq * ((spin & 1) ? -1.0 : 1.0) * ipow(-(m + spin));
      double complex *row = F + (long)(m + L - 1) * L;
      for (int mp = 0; mp <= l; ++mp) row[mp] += c * dlv(dl, l, mp, m) * dlv(dl, l, mp, -spin);
    }
  }
  for (int m = -(L - 1); m <= L - 1; ++m) {
    const double sgn = ((m + spin) & 1) ? -1.0 : 1.0; /* F(-m') = (-1)^{m+s} F(m') */
    double complex *seq = w->seq;
    memset(seq, 0, sizeof(double complex) * P);
    for (int mp = 0; mp < L; ++mp) {
      const double complex v = F[(long)(m + L - 1) * L + mp];
      const double complex ph = w->ss ? 1.0 : cexp(I * mp * OPI / (2.0 * L - 1.0));
      seq[mp] = v * ph;
      if (mp > 0) seq[P - mp] = sgn * v * conj(ph);
    }
    oracle_fft_exec(w->pb_b, seq, 1);
    const int slot = (m + nA) % nA;
    for (int b = 0; b < nB; ++b) w->tmp[(long)b * nA + slot] = seq[b];
  }
  for (int b = 0; b < nB; ++b) {
    double complex *row = w->tmp + (long)b * nA;
    if (w->ss) row[L] = 0.0; /* m = -L (Nyquist) is never populated */
    oracle_fft_exec(w->pa_b, row, 1);
    memcpy(f + (long)b * nA, row, sizeof(double complex) * nA);
  }
}

/* flm = int f sY*_lm dOmega by the MW / MW_SS quadrature: alpha DFT,
 * beta extension with (-1)^{m+s}, beta DFT (+ half-sample phase for MW),
 * convolution with the weights w(p) of length 4L-3, Delta contraction. */
static void ssht_forward(ssht_ws *w, double complex *flm, const double complex *f, int L0,
                         int spin, const double *dl) {
  const int L = w->L, nA = w->nA, nB = w->nB, P = w->P, Pw = w->Pw;
  const int M = 2 * L - 1;
  /* alpha DFT per b, /nA */
  for (int b = 0; b < nB; ++b) {
    double complex *row = w->tmp + (long)b * nA;
    memcpy(row, f + (long)b * nA, sizeof(double complex) * nA);
    oracle_fft_exec(w->pa_f, row, 1);
    for (int a = 0; a < nA; ++a) row[a] /= nA;
  }
  double complex *G = w->F; /* G[(m+L-1)*M + (mp+L-1)] */
  for (int m = -(L - 1); m <= L - 1; ++m) {
    const int slot = (m + nA) % nA;
    const double sgn = ((m + spin) & 1) ? -1.0 : 1.0;
    double complex *seq = w->seq;
    for (int b = 0; b < nB; ++b) seq[b] = w->tmp[(long)b * nA + slot];
    if (!w->ss)
      for (int b = L; b <= 2 * L - 2; ++b) seq[b] = sgn * seq[2 * L - 2 - b];
    else
      for (int b = L + 1; b <= 2 * L - 1; ++b) seq[b] = sgn * seq[2 * L - b];
    oracle_fft_exec(w->pb_f, seq, 1);
    /* Fmm'' for m'' in [-(L-1), L-1], then the weight convolution in
     * "real space" with FFTs of length 4L-3 */
    double complex *pad = w->seq + P; /* scratch behind seq is large enough? use separate */
    (void)pad;
    double complex *cv = (double complex *)calloc(Pw, sizeof(double complex));
    for (int mpp = -(L - 1); mpp <= L - 1; ++mpp) {
      double complex v = seq[(mpp + P) % P] / P;
      if (!w->ss) v *= cexp(-I * mpp * OPI / (2.0 * L - 1.0));
      cv[(mpp + Pw) % Pw] = v;
    }
    oracle_fft_exec(w->pw_b, cv, 1);
    for (int r = 0; r < Pw; ++r) cv[r] *= w->wr[r];
    oracle_fft_exec(w->pw_f, cv, 1);
    for (int mp = -(L - 1); mp <= L - 1; ++mp)
      G[(long)(m + L - 1) * M + (mp + L - 1)] = cv[(mp + Pw) % Pw] * (2.0 * OPI / Pw);
    free(cv);
  }
  const int l0 = L0 > abs(spin) ? L0 : abs(spin);
  for (int l = 0; l < L; ++l)
    for (int m = -l; m <= l; ++m) flm[l * l + l + m] = 0.0;
  for (int l = l0; l < L; ++l) {
    const double sq = sqrt((2.0 * l + 1.0) / (4.0 * OPI)) * ((spin & 1) ? -1.0 : 1.0);
    for (int m = -l; m <= l; ++m) {
      double complex acc = 0.0;
      const double complex *g = G + (long)(m + L - 1) * M + (L - 1);
      for (int mp = -l; mp <= l; ++mp) acc += dlv(dl, l, mp, m) * dlv(dl, l, mp, -spin) * g[mp];
      flm[l * l + l + m] = sq * ipow(m + spin) * acc;
    }
  }
}

/* ------------------------------------------------------------------ */
/* via-SSHT transforms (so3_core.c:51-897)                             */
/* ------------------------------------------------------------------ */
static double *dl_for(const oracle_params *p) {
  double *dl = (double *)malloc(sizeof(double) * (oracle_dl_offset(p->L) + 1));
  oracle_dl_quarter(p->L, p->dl_method, dl);
  return dl;
}

static void check_params(const oracle_params *p) {
  if (p->sampling_scheme != MW && p->sampling_scheme != MWSS) die("Invalid sampling scheme");
  if (p->storage != PADDED && p->storage != COMPACT) die("Invalid storage method");
  if (p->n_order != ZERO_FIRST && p->n_order != NEG_FIRST) die("Invalid n-order");
  if (p->n_mode < 0 || p->n_mode > MODE_L) die("Invalid n-mode");
  if (p->dl_method != RISBO && p->dl_method != TRAPANI) die("Invalid dl method");
}

/* copy block n of flmn into a full L*L flm array (zeros for l < |n| in
 * compact storage), so3_core.c:172-188 */
static void gather_block(double complex *flm, const double complex *flmn, int n,
                         const oracle_params *p, int real) {
  const int L = p->L;
  if (p->storage == PADDED) {
    const int base = o_ind(0, 0, n, p, real);
    memcpy(flm, flmn + base, sizeof(double complex) * L * L);
  } else {
    const int an = abs(n);
    const int base = o_ind(an, -an, n, p, real);
    for (int i = 0; i < n * n; ++i) flm[i] = 0.0;
    memcpy(flm + n * n, flmn + base, sizeof(double complex) * (L * L - n * n));
  }
}

static void inverse_via_ssht_core(double complex *fc, double *fr, const double complex *flmn,
                                  const oracle_params *p) {
  check_params(p);
  const int L = p->L, N = p->N, L0 = p->L0;
  const int real = fr != NULL;
  const int nA = o_nalpha(p), nB = o_nbeta(p);
  const long stride = (long)nA * nB;
  if (p->steerable && real) die("steerable real signals are not supported");
  const int ng = p->steerable ? 2 * N : 2 * N - 1; /* gamma DFT length */
  double *dl = dl_for(p);
  double complex *fn = (double complex *)calloc((size_t)ng * stride, sizeof(double complex));
  double complex *flm = (double complex *)malloc(sizeof(double complex) * L * L);
  ssht_ws *w = ssht_ws_create(L, p->sampling_scheme == MWSS);
  const int nlo = real ? 0 : -N + 1;
  for (int n = nlo; n <= N - 1; ++n) {
    if (!o_n_active(n, N, p->n_mode)) continue;
    const int L0e = L0 > abs(n) ? L0 : abs(n);
    gather_block(flm, flmn, n, p, real);
    for (int l = L0e; l < L; ++l) {
      const double fac = sqrt((double)(2 * l + 1) / (16. * pow(OPI, 3.)));
      for (int m = -l; m <= l; ++m) flm[l * l + l + m] *= fac;
    }
    const int slot = n < 0 ? n + ng : n;
    double complex *blk = fn + slot * stride;
    ssht_inverse(w, blk, flm, L0e, -n, dl);
    if (n % 2)
      for (long i = 0; i < stride; ++i) blk[i] = -blk[i];
  }
  if (!real) {
    /* backward DFT over the n slots at every (a, b) */
    oracle_fft_plan *pg = oracle_fft_plan_create(ng, +1);
    for (long i = 0; i < stride; ++i) oracle_fft_exec(pg, fn + i, stride);
    oracle_fft_plan_destroy(pg);
    const int nG = p->steerable ? N : 2 * N - 1;
    memcpy(fc, fn, sizeof(double complex) * nG * stride);
  } else {
    /* complex-to-real DFT of the half spectrum n = 0..N-1 (Im of n=0 ignored) */
    const int nG = 2 * N - 1;
    double complex *col = (double complex *)malloc(sizeof(double complex) * nG);
    oracle_fft_plan *pg = oracle_fft_plan_create(nG, +1);
    for (long i = 0; i < stride; ++i) {
      col[0] = creal(fn[i]);
      for (int n = 1; n < N; ++n) {
        col[n] = fn[n * stride + i];
        col[nG - n] = conj(col[n]);
      }
      oracle_fft_exec(pg, col, 1);
      for (int g = 0; g < nG; ++g) fr[g * stride + i] = creal(col[g]);
    }
    oracle_fft_plan_destroy(pg);
    free(col);
  }
  ssht_ws_destroy(w);
  free(flm);
  free(fn);
  free(dl);
}

static void forward_via_ssht_core(double complex *flmn, const double complex *fc, const double *fr,
                                  const oracle_params *p) {
  check_params(p);
  const int L = p->L, N = p->N, L0 = p->L0;
  const int real = fr != NULL;
  const int nA = o_nalpha(p), nB = o_nbeta(p);
  const long stride = (long)nA * nB;
  if (p->steerable && real) die("steerable real signals are not supported");
  double *dl = dl_for(p);
  const int nG = 2 * N - 1;
  double complex *fn = (double complex *)calloc((size_t)nG * stride, sizeof(double complex));
  if (!real) {
    if (p->steerable) {
      /* explicit DFT over the N gamma samples for n = -N+1, -N+3, ...
       * (so3_core.c:309-328) */
      for (int n = -N + 1; n < N; n += 2) {
        const int slot = n < 0 ? n + nG : n;
        for (int g = 0; g < N; ++g) {
          const double gamma = g * OPI / N;
          const double complex e = cexp(-I * n * gamma) * (2 * OPI / N);
          for (long i = 0; i < stride; ++i) fn[slot * stride + i] += e * fc[g * stride + i];
        }
      }
    } else {
      memcpy(fn, fc, sizeof(double complex) * nG * stride);
      oracle_fft_plan *pg = oracle_fft_plan_create(nG, -1);
      for (long i = 0; i < stride; ++i) oracle_fft_exec(pg, fn + i, stride);
      oracle_fft_plan_destroy(pg);
      const double fac = 2 * OPI / (double)nG;
      for (long i = 0; i < nG * stride; ++i) fn[i] *= fac;
    }
  } else {
    double complex *col = (double complex *)malloc(sizeof(double complex) * nG);
    oracle_fft_plan *pg = oracle_fft_plan_create(nG, -1);
    const double fac = 2 * OPI / (double)nG;
    for (long i = 0; i < stride; ++i) {
      for (int g = 0; g < nG; ++g) col[g] = fr[g * stride + i];
      oracle_fft_exec(pg, col, 1);
      for (int n = 0; n < N; ++n) fn[n * stride + i] = col[n] * fac;
    }
    oracle_fft_plan_destroy(pg);
    free(col);
  }
  double complex *flm = (double complex *)malloc(sizeof(double complex) * L * L);
  ssht_ws *w = ssht_ws_create(L, p->sampling_scheme == MWSS);
  const int nlo = real ? 0 : -N + 1;
  for (int n = nlo; n <= N - 1; ++n) {
    if (!o_n_active(n, N, p->n_mode)) continue;
    const int L0e = L0 > abs(n) ? L0 : abs(n);
    const int slot = n < 0 ? n + nG : n;
    ssht_forward(w, flm, fn + slot * stride, L0e, -n, dl);
    const double sgn = (n % 2) ? -1.0 : 1.0;
    for (int l = L0e; l < L; ++l) {
      const double fac = sgn * sqrt(4.0 * OPI / (double)(2 * l + 1));
      for (int m = -l; m <= l; ++m) flm[l * l + l + m] *= fac;
    }
    if (p->storage == PADDED) {
      memcpy(flmn + o_ind(0, 0, n, p, real), flm, sizeof(double complex) * L * L);
    } else {
      const int an = abs(n);
      memcpy(flmn + o_ind(an, -an, n, p, real), flm + n * n, sizeof(double complex) * (L * L - n * n));
    }
  }
  ssht_ws_destroy(w);
  free(flm);
  free(fn);
  free(dl);
}

void oracle_inverse_via_ssht(double complex *f, const double complex *flmn, const oracle_params *p) {
  inverse_via_ssht_core(f, NULL, flmn, p);
}
void oracle_inverse_via_ssht_real(double *f, const double complex *flmn, const oracle_params *p) {
  inverse_via_ssht_core(NULL, f, flmn, p);
}
void oracle_forward_via_ssht(double complex *flmn, const double complex *f, const oracle_params *p) {
  forward_via_ssht_core(flmn, f, NULL, p);
}
void oracle_forward_via_ssht_real(double complex *flmn, const double *f, const oracle_params *p) {
  forward_via_ssht_core(flmn, NULL, f, p);
}

/* ------------------------------------------------------------------ */
/* direct transforms (so3_core.c:914-2385)                             */
/* ------------------------------------------------------------------ */
/* n values of the per-l contraction loops (so3_core.c:1040-1079 complex,
 * :1762-1795 real); returns count written to ns[] */
static int direct_l_nset(int el, const oracle_params *p, int real, int *ns) {
  const int N = p->N;
  int s, e, d;
  if (!real) {
    switch (p->n_mode) {
    case MODE_ALL: s = -N + 1 > -el ? -N + 1 : -el; e = N - 1 < el ? N - 1 : el; d = 1; break;
    case MODE_EVEN:
      s = -N + 1 > -el ? -N + 1 : -el; s += (-s) % 2;
      e = N - 1 < el ? N - 1 : el; e -= e % 2; d = 2; break;
    case MODE_ODD:
      s = -N + 1 > -el ? -N + 1 : -el; s += 1 + s % 2;
      e = N - 1 < el ? N - 1 : el; e -= 1 - e % 2; d = 2; break;
    case MODE_MAX:
      if (el < N - 1) return 0;
      s = -N + 1; e = N - 1; d = 2 * N - 2 > 1 ? 2 * N - 2 : 1; break;
    default: /* MODE_L */
      if (el >= N) return 0;
      s = -el; e = el; d = 2 * el > 1 ? 2 * el : 1; break;
    }
  } else {
    switch (p->n_mode) {
    case MODE_ALL: s = 0; e = N - 1 < el ? N - 1 : el; d = 1; break;
    case MODE_EVEN: s = 0; e = N - 1 < el ? N - 1 : el; e -= e % 2; d = 2; break;
    case MODE_ODD: s = 1; e = N - 1 < el ? N - 1 : el; e -= 1 - e % 2; d = 2; break;
    case MODE_MAX: if (el < N - 1) return 0; s = N - 1; e = N - 1; d = 1; break;
    default: if (el >= N) return 0; s = el; e = el; d = 1; break;
    }
  }
  int c = 0;
  for (int n = s; n <= e; n += d) ns[c++] = n;
  return c;
}

/* Fmn m' contraction of the direct inverse (so3_core.c:1081-1114) into
 * T[(n+N-1)][(m+L-1)][(mp+L-1)], then symmetry in m', phase, and the 3-D
 * backward DFT over the torus; real: c2r semantics over n. */
static void inverse_direct_core(double complex *fc, double *fr, const double complex *flmn,
                                const oracle_params *p) {
  check_params(p);
  const int L = p->L, N = p->N, L0 = p->L0, real = fr != NULL;
  const int M = 2 * L - 1, NG = 2 * N - 1;
  double *dl = dl_for(p);
  double complex *T = (double complex *)calloc((size_t)NG * M * M, sizeof(double complex));
#define TIX(n, m, mp) ((((long)((n) + N - 1) * M + ((m) + L - 1)) * M) + ((mp) + L - 1))
  int *ns = (int *)malloc(sizeof(int) * (2 * N + 2));
  for (int el = L0; el < L; ++el) {
    const int cnt = direct_l_nset(el, p, real, ns);
    const double elf = (2.0 * el + 1.0) / (8.0 * OPI * OPI);
    for (int k = 0; k < cnt; ++k) {
      const int n = ns[k];
      for (int m = -el; m <= el; ++m) {
        const double complex c = elf * flmn[o_ind(el, m, n, p, real)] * ipow(n - m);
        for (int mp = 0; mp <= el; ++mp)
          T[TIX(n, m, mp)] += c * dlv(dl, el, mp, m) * dlv(dl, el, mp, n);
      }
    }
  }
  /* negative m' by symmetry, then the half-sample phase (so3_core.c:1148-1165) */
  const int nlo = real ? 0 : -N + 1;
  for (int n = nlo; n <= N - 1; ++n)
    for (int m = -(L - 1); m <= L - 1; ++m) {
      const double s = ((m + n) & 1) ? -1.0 : 1.0;
      for (int mp = 1; mp <= L - 1; ++mp) T[TIX(n, m, -mp)] = s * T[TIX(n, m, mp)];
      for (int mp = -(L - 1); mp <= L - 1; ++mp) T[TIX(n, m, mp)] *= cexp(I * mp * OPI / (2.0 * L - 1.0));
    }
  /* 3-D backward DFT: over m' (keep b < L), over m, over n */
  oracle_fft_plan *pm = oracle_fft_plan_create(M, +1);
  oracle_fft_plan *pn = oracle_fft_plan_create(NG, +1);
  double complex *seq = (double complex *)malloc(sizeof(double complex) * (M > NG ? M : NG));
  double complex *E = (double complex *)calloc((size_t)NG * L * M, sizeof(double complex)); /* [nslot][b][a] */
  for (int n = nlo; n <= N - 1; ++n) {
    const int ns_ = (n + NG) % NG;
    double complex *tmp = (double complex *)calloc((size_t)L * M, sizeof(double complex)); /* [b][mslot] */
    for (int m = -(L - 1); m <= L - 1; ++m) {
      for (int mp = -(L - 1); mp <= L - 1; ++mp) seq[(mp + M) % M] = T[TIX(n, m, mp)];
      oracle_fft_exec(pm, seq, 1);
      for (int b = 0; b < L; ++b) tmp[(long)b * M + (m + M) % M] = seq[b];
    }
    for (int b = 0; b < L; ++b) {
      oracle_fft_exec(pm, tmp + (long)b * M, 1);
      memcpy(E + ((long)ns_ * L + b) * M, tmp + (long)b * M, sizeof(double complex) * M);
    }
    free(tmp);
  }
  const long stride = (long)L * M;
  for (long i = 0; i < stride; ++i) {
    if (!real) {
      for (int g = 0; g < NG; ++g) seq[g] = E[g * stride + i];
    } else {
      seq[0] = creal(E[i]);
      for (int n = 1; n < N; ++n) { seq[n] = E[n * stride + i]; seq[NG - n] = conj(seq[n]); }
    }
    oracle_fft_exec(pn, seq, 1);
    for (int g = 0; g < NG; ++g) {
      if (real) fr[g * stride + i] = creal(seq[g]);
      else fc[g * stride + i] = seq[g];
    }
  }
#undef TIX
  oracle_fft_plan_destroy(pm);
  oracle_fft_plan_destroy(pn);
  free(seq); free(E); free(T); free(ns); free(dl);
}

/* forward direct (so3_core.c:1238-1640): 2-D DFT over (a, g) per b with
 * 1/((2L-1)(2N-1)), beta extension, beta DFT /(2L-1) and phase, weight
 * convolution (length 4L-3, 4 pi^2/(4L-3)), Delta contraction. */
static void forward_direct_core(double complex *flmn, const double complex *fc, const double *fr,
                                const oracle_params *p) {
  check_params(p);
  const int L = p->L, N = p->N, L0 = p->L0, real = fr != NULL;
  const int M = 2 * L - 1, NG = 2 * N - 1, Pw = 4 * L - 3;
  const long stride = (long)L * M;
  const int nlo = real ? 0 : -N + 1;
  double *dl = dl_for(p);
  /* gamma then alpha DFT, normalised */
  double complex *Fn = (double complex *)calloc((size_t)NG * stride, sizeof(double complex));
  oracle_fft_plan *pn = oracle_fft_plan_create(NG, -1);
  oracle_fft_plan *pm = oracle_fft_plan_create(M, -1);
  oracle_fft_plan *pwb = oracle_fft_plan_create(Pw, +1);
  oracle_fft_plan *pwf = oracle_fft_plan_create(Pw, -1);
  double complex *seq = (double complex *)malloc(sizeof(double complex) * (Pw + M + NG));
  for (long i = 0; i < stride; ++i) {
    for (int g = 0; g < NG; ++g) seq[g] = real ? fr[g * stride + i] : fc[g * stride + i];
    oracle_fft_exec(pn, seq, 1);
    for (int g = 0; g < NG; ++g) Fn[g * stride + i] = seq[g];
  }
  const double norm = 1.0 / (2.0 * L - 1.0) / (2.0 * N - 1.0);
  for (int g = 0; g < NG; ++g)
    for (int b = 0; b < L; ++b) {
      double complex *row = Fn + (long)g * stride + (long)b * M;
      oracle_fft_exec(pm, row, 1);
      for (int a = 0; a < M; ++a) row[a] *= norm;
    }
  /* weights */
  double complex *wr = (double complex *)calloc(Pw, sizeof(double complex));
  for (int q = -2 * (L - 1); q <= 2 * (L - 1); ++q) wr[(q + Pw) % Pw] = o_weight(q);
  oracle_fft_exec(pwb, wr, 1);
  /* G[(n)][(m)][(mp)] */
  double complex *G = (double complex *)calloc((size_t)NG * M * M, sizeof(double complex));
#define GIX(n, m, mp) ((((long)((n) + N - 1) * M + ((m) + L - 1)) * M) + ((mp) + L - 1))
  for (int n = nlo; n <= N - 1; ++n) {
    const int nsl = (n + NG) % NG;
    for (int m = -(L - 1); m <= L - 1; ++m) {
      const int msl = (m + M) % M;
      const double s = ((m + n) & 1) ? -1.0 : 1.0;
      for (int b = 0; b < L; ++b) seq[b] = Fn[(long)nsl * stride + (long)b * M + msl];
      for (int b = L; b < M; ++b) seq[b] = s * seq[2 * L - 2 - b];
      oracle_fft_exec(pm, seq, 1);
      double complex *cv = seq + M;
      memset(cv, 0, sizeof(double complex) * Pw);
      for (int mp = -(L - 1); mp <= L - 1; ++mp)
        cv[(mp + Pw) % Pw] = seq[(mp + M) % M] / (2.0 * L - 1.0) * cexp(-I * mp * OPI / (2.0 * L - 1.0));
      oracle_fft_exec(pwb, cv, 1);
      for (int r = 0; r < Pw; ++r) cv[r] *= wr[r];
      oracle_fft_exec(pwf, cv, 1);
      for (int mp = -(L - 1); mp <= L - 1; ++mp)
        G[GIX(n, m, mp)] = cv[(mp + Pw) % Pw] * 4.0 * OPI * OPI / (4.0 * L - 3.0);
    }
  }
  /* zero, then contract (so3_core.c:1548-1618) */
  for (int n = nlo; n <= N - 1; ++n)
    for (int el = abs(n); el < L; ++el)
      for (int m = -el; m <= el; ++m) flmn[o_ind(el, m, n, p, real)] = 0.0;
  int *ns = (int *)malloc(sizeof(int) * (2 * N + 2));
  for (int el = L0; el < L; ++el) {
    const int cnt = direct_l_nset(el, p, real, ns);
    for (int k = 0; k < cnt; ++k) {
      const int n = ns[k];
      for (int m = -el; m <= el; ++m) {
        double complex acc = 0.0;
        for (int mp = -el; mp <= el; ++mp) acc += dlv(dl, el, mp, m) * dlv(dl, el, mp, n) * G[GIX(n, m, mp)];
        flmn[o_ind(el, m, n, p, real)] += ipow(m - n) * acc;
      }
    }
  }
#undef GIX
  free(ns); free(G); free(wr); free(seq); free(Fn); free(dl);
  oracle_fft_plan_destroy(pn); oracle_fft_plan_destroy(pm);
  oracle_fft_plan_destroy(pwb); oracle_fft_plan_destroy(pwf);
}

void oracle_inverse_direct(double complex *f, const double complex *flmn, const oracle_params *p) {
  inverse_direct_core(f, NULL, flmn, p);
}
void oracle_inverse_direct_real(double *f, const double complex *flmn, const oracle_params *p) {
  inverse_direct_core(NULL, f, flmn, p);
}
void oracle_forward_direct(double complex *flmn, const double complex *f, const oracle_params *p) {
  forward_direct_core(flmn, f, NULL, p);
}
void oracle_forward_direct_real(double complex *flmn, const double *f, const oracle_params *p) {
  forward_direct_core(flmn, NULL, f, p);
}

/* ------------------------------------------------------------------ */
/* convolutions (so3_conv.c:44-90, 186-215)                            */
/* ------------------------------------------------------------------ */
void oracle_s2toso3(double complex *hlmn, const oracle_params *hp, const double complex *flm,
                    const double complex *glm) {
  const int real = hp->reality;
  /* every (l, m, n) slot of the storage: non-zero where |m|,|n| <= l */
  for (int i = 0; i < oracle_flmn_size(hp); ++i) hlmn[i] = 0.0;
  const int nlo = real ? 0 : -hp->N + 1;
  for (int n = nlo; n <= hp->N - 1; ++n)
    for (int el = abs(n); el < hp->L; ++el)
      for (int m = -el; m <= el; ++m)
        hlmn[o_ind(el, m, n, hp, real)] = flm[el * el + el + m] * conj(glm[el * el + el + n]) *
                                          (8 * OPI * OPI / (2 * el + 1));
}

/* ------------------------------------------------------------------ */
/* CPU-baseline timing on a bounded sample                              */
/* ------------------------------------------------------------------ */
#include <time.h>
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Times the via-SSHT work of the listed n slices of a complex transform
 * (per n: flm scaling, spin-(-n) SSHT inverse and forward, as in
 * so3_core.c:160-214 and :373-435), the Delta-table build, and one full
 * gamma DFT over every gamma_every-th (a, b) column, scaled to all
 * columns.  Returns per-n seconds in t_n[], and the table / gamma times in
 * out[0] / out[1]. */
void oracle_time_sample(const oracle_params *p, const int *ns, int count, double *t_n, double *out,
                        int gamma_every) {
  const int L = p->L, N = p->N;
  const int ss = p->sampling_scheme == MWSS;
  double t0 = now_s();
  double *dl = dl_for(p);
  out[0] = now_s() - t0;
  ssht_ws *w = ssht_ws_create(L, ss);
  const long stride = (long)w->nA * w->nB;
  double complex *flm = (double complex *)malloc(sizeof(double complex) * L * L);
  double complex *blk = (double complex *)malloc(sizeof(double complex) * stride);
  double complex *back = (double complex *)malloc(sizeof(double complex) * L * L);
  for (int i = 0; i < L * L; ++i) flm[i] = (double)((i * 7919) % 1000) / 1000.0 - 0.5 + I * 0.25;
  for (int k = 0; k < count; ++k) {
    const int n = ns[k];
    const int L0e = p->L0 > abs(n) ? p->L0 : abs(n);
    t0 = now_s();
    for (int l = L0e; l < L; ++l) {
      const double fac = sqrt((double)(2 * l + 1) / (16. * pow(OPI, 3.)));
      for (int m = -l; m <= l; ++m) flm[l * l + l + m] *= fac;
    }
    ssht_inverse(w, blk, flm, L0e, -n, dl);
    ssht_forward(w, back, blk, L0e, -n, dl);
    t_n[k] = now_s() - t0;
  }
  /* one gamma DFT over every (a, b) column */
  const int nG = 2 * N - 1;
  double complex *col = (double complex *)malloc(sizeof(double complex) * nG);
  oracle_fft_plan *pg = oracle_fft_plan_create(nG, -1);
  t0 = now_s();
  long done = 0;
  for (long i = 0; i < stride; i += gamma_every, ++done) {
    for (int g = 0; g < nG; ++g) col[g] = (double)((i + g) % 17);
    oracle_fft_exec(pg, col, 1);
  }
  out[1] = (now_s() - t0) * (double)stride / (double)(done ? done : 1);
  oracle_fft_plan_destroy(pg);
  free(col); free(flm); free(blk); free(back);
  ssht_ws_destroy(w);
  free(dl);
}

/* ------------------------------------------------------------------ */
/* helpers shared with oracle_adjoint.c (declared in oracle_int.h)      */
/* ------------------------------------------------------------------ */
int oracle_h_ind(int el, int m, int n, const oracle_params *p, int real) { return o_ind(el, m, n, p, real); }
int oracle_h_n_active(int n, int N, int mode) { return o_n_active(n, N, mode); }
double *oracle_h_dl(const oracle_params *p) { return dl_for(p); }
void oracle_h_check(const oracle_params *p) { check_params(p); }
int oracle_h_l_nset(int el, const oracle_params *p, int real, int *ns) { return direct_l_nset(el, p, real, ns); }
double complex oracle_h_weight(int p) { return o_weight(p); }
double oracle_h_dlv(const double *dl, int l, int mp, int m) { return dlv(dl, l, mp, m); }
double complex oracle_h_ipow(int k) { return ipow(k); }
