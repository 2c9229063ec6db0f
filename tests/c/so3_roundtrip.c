/*
 * so3_roundtrip.c -- a plain C99 caller of the drop-in API, written as a user
 * of the reference library would write it (cf. the reference harness
 * src/c/so3_test_csv.c:139-235): it includes <so3/so3.h>, fills
 * so3_parameters_t, allocates host buffers with so3_sampling_*_size, runs
 * inverse then forward and reports the maximum absolute round-trip error over
 * flmn_size entries.  It also calls the real-signal pair, the direct pair,
 * an adjoint, the S^2 -> SO(3) convolution and the index helpers.
 *
 * Build: gcc -std=c99 -Iinclude tests/c/so3_roundtrip.c -Lso3_amd -lso3_amd -lm
 * Run:   ./a.out L N      (prints one line per check; exit 1 on failure)
 */
#define _POSIX_C_SOURCE 199309L /* clock_gettime under -std=c99 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <so3/so3.h>

/* U[-1, 1) from a 64-bit linear congruential sequence */
static double uni(unsigned long long *s) {
  *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(*s >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

static double now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* random coefficients on the non-zero (l, m, n) of the parameters */
static void gen(double complex *flmn, const so3_parameters_t *p, unsigned long long seed) {
  int n0, n1, dn, l0, l1, dl, m0, m1, dm, ind;
  memset(flmn, 0, sizeof(double complex) * so3_sampling_flmn_size(p));
  so3_sampling_n_loop_values(&n0, &n1, &dn, p);
  for (int n = n0; n <= n1; n += dn) {
    so3_sampling_el_loop_values(&l0, &l1, &dl, n, p);
    for (int el = l0; el <= l1; el += dl) {
      so3_sampling_m_loop_values(&m0, &m1, &dm, el);
      for (int m = m0; m <= m1; m += dm) {
        if (p->reality) {
          so3_sampling_elmn2ind_real(&ind, el, m, n, p);
          if (n == 0 && m < 0) continue; /* filled from m > 0 below */
        } else {
          so3_sampling_elmn2ind(&ind, el, m, n, p);
        }
        flmn[ind] = uni(&seed) + I * uni(&seed);
        if (p->reality && n == 0) {
          int i2;
          if (m == 0) {
            flmn[ind] = creal(flmn[ind]);
          } else {
            so3_sampling_elmn2ind_real(&i2, el, -m, 0, p);
            flmn[i2] = ((m & 1) ? -1.0 : 1.0) * conj(flmn[ind]);
          }
        }
      }
    }
  }
}

static double maxdiff(const double complex *a, const double complex *b, int n) {
  double e = 0.0;
  for (int i = 0; i < n; ++i) e = fmax(e, cabs(a[i] - b[i]));
  return e;
}

int main(int argc, char **argv) {
  const int L = argc > 1 ? atoi(argv[1]) : 16, N = argc > 2 ? atoi(argv[2]) : 16;
  int fail = 0;
  so3_parameters_t p;
  memset(&p, 0, sizeof p); /* MW, zero-first, padded, all n, Risbo */
  p.L = L;
  p.N = N;
  const int nl = so3_sampling_flmn_size(&p), nf = so3_sampling_f_size(&p);
  double complex *flmn = malloc(sizeof(double complex) * nl);
  double complex *back = malloc(sizeof(double complex) * nl);
  double complex *f = malloc(sizeof(double complex) * nf);
  double complex *f2 = malloc(sizeof(double complex) * nf);

  /* complex via SSHT */
  gen(flmn, &p, 1);
  double t0 = now();
  so3_core_inverse_via_ssht(f, flmn, &p);
  so3_core_forward_via_ssht(back, f, &p);
  double t1 = now();
  double e = maxdiff(flmn, back, nl);
  printf("complex via_ssht L=%d N=%d: round-trip max abs error %.3e, %.3f s\n", L, N, e, t1 - t0);
  fail |= !(e < 1e-10);

  /* direct pair agrees with via SSHT */
  so3_core_inverse_direct(f2, flmn, &p);
  e = maxdiff(f, f2, nf);
  printf("direct vs via_ssht inverse: max abs difference %.3e\n", e);
  fail |= !(e < 1e-10);

  /* adjoint of the direct inverse: <A x, y> = <x, A^H y> with y = A x */
  so3_adjoint_inverse_direct(back, f2, &p);
  double complex lhs = 0, rhs = 0;
  for (int i = 0; i < nf; ++i) lhs += conj(f2[i]) * f2[i];
  for (int i = 0; i < nl; ++i) rhs += conj(flmn[i]) * back[i];
  e = cabs(lhs - rhs) / cabs(lhs);
  printf("adjoint inverse dot product: relative mismatch %.3e\n", e);
  fail |= !(e < 1e-10);

  /* real signals */
  so3_parameters_t pr = p;
  pr.reality = 1;
  const int nlr = so3_sampling_flmn_size(&pr);
  double *fr = malloc(sizeof(double) * so3_sampling_f_size(&pr));
  gen(flmn, &pr, 2);
  so3_core_inverse_via_ssht_real(fr, flmn, &pr);
  so3_core_forward_via_ssht_real(back, fr, &pr);
  e = maxdiff(flmn, back, nlr);
  printf("real via_ssht: round-trip max abs error %.3e\n", e);
  fail |= !(e < 1e-10);

  /* S^2 -> SO(3) convolution: h_lmn = f_lm conj(g_ln) 8 pi^2 / (2l+1) */
  double complex *flm = malloc(sizeof(double complex) * L * L), *glm = malloc(sizeof(double complex) * L * L);
  unsigned long long s = 3;
  for (int i = 0; i < L * L; ++i) flm[i] = uni(&s) + I * uni(&s), glm[i] = uni(&s) + I * uni(&s);
  so3_conv_s2toso3_harmonic_convolution(back, &p, flm, glm);
  int ind;
  const int el = L - 1, m = -(L / 2), n = (N - 1) / 2;
  so3_sampling_elmn2ind(&ind, el, m, n, &p);
  const double complex want = flm[el * el + el + m] * conj(glm[el * el + el + n]) * 8 * SO3_PI * SO3_PI / (2 * el + 1);
  e = cabs(back[ind] - want);
  printf("s2toso3 entry (%d, %d, %d): error %.3e\n", el, m, n, e);
  fail |= !(e < 1e-12);

  /* index helpers invert each other */
  int bad = 0;
  for (int i = 0; i < nl; ++i) {
    int l2, m2, n2, j;
    so3_sampling_ind2elmn(&l2, &m2, &n2, i, &p);
    so3_sampling_elmn2ind(&j, l2, m2, n2, &p);
    bad += j != i;
  }
  printf("index helpers: %d mismatches over %d indices\n", bad, nl);
  fail |= bad != 0;

  free(flmn); free(back); free(f); free(f2); free(fr); free(flm); free(glm);
  printf("%s\n", fail ? "FAILED" : "OK");
  return fail;
}
