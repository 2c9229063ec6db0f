/*
 * so3_sampling.c -- sizes, angles, weights and index maps of the SO(3) grid
 * and of the flmn coefficient storage.
 *
 * Behaviour follows the reference src/c/so3_sampling.c (cited per function).
 * Integer results are identical; floating-point results use the same
 * expression order, so they are identical bit for bit.  Invalid enum values
 * end the process through SO3_ERROR_GENERIC, as in the reference.
 */
#include <math.h>
#include <stdbool.h>
#include <stdlib.h>

#include "so3/so3_error.h"
#include "so3/so3_sampling.h"

/* pi/2 as the reference spells SSHT_PION2 / SO3_PION2. */
static const double kPiOn2 = SO3_PION2;

/* so3_sampling.c:32-51.  w(p) = int_0^pi exp(-i p b) sin(b) db. */
double complex so3_sampling_weight(const so3_parameters_t *parameters, int p) {
  if (parameters->sampling_scheme != SO3_SAMPLING_MW &&
      parameters->sampling_scheme != SO3_SAMPLING_MW_SS)
    SO3_ERROR_GENERIC("Invalid sampling scheme.");
  if (p == 1) return -I * kPiOn2;   /* real part is -0.0, as in C99 */
  if (p == -1) return I * kPiOn2;
  if (p % 2 == 0) return 2.0 / (1.0 - p * p);
  return 0.0;
}

/* so3_sampling.c:142-156 */
int so3_sampling_nalpha(const so3_parameters_t *parameters) {
  switch (parameters->sampling_scheme) {
  case SO3_SAMPLING_MW: return 2 * parameters->L - 1;
  case SO3_SAMPLING_MW_SS: return 2 * parameters->L;
  default: SO3_ERROR_GENERIC("Invalid sampling scheme.");
  }
}

/* so3_sampling.c:174-185 */
int so3_sampling_nbeta(const so3_parameters_t *parameters) {
  switch (parameters->sampling_scheme) {
  case SO3_SAMPLING_MW: return parameters->L;
  case SO3_SAMPLING_MW_SS: return parameters->L + 1;
  default: SO3_ERROR_GENERIC("Invalid sampling scheme.");
  }
}

/* so3_sampling.c:197-206 */
int so3_sampling_ngamma(const so3_parameters_t *parameters) {
  return parameters->steerable ? parameters->N : 2 * parameters->N - 1;
}

/* so3_sampling.c:80-89 */
int so3_sampling_f_size(const so3_parameters_t *parameters) {
  return so3_sampling_nalpha(parameters) * so3_sampling_nbeta(parameters) *
         so3_sampling_ngamma(parameters);
}

/* so3_sampling.c:109-124 (samples without pole degeneracy) */
int so3_sampling_n(const so3_parameters_t *parameters) {
  const int L = parameters->L;
  switch (parameters->sampling_scheme) {
  case SO3_SAMPLING_MW: return ((2 * L - 1) * (L - 1) + 1) * so3_sampling_ngamma(parameters);
  case SO3_SAMPLING_MW_SS: return ((2 * L) * (L - 1) + 2) * so3_sampling_ngamma(parameters);
  default: SO3_ERROR_GENERIC("Invalid sampling scheme.");
  }
}

/* so3_sampling.c:224-236 */
double so3_sampling_a2alpha(int a, const so3_parameters_t *parameters) {
  const int L = parameters->L;
  switch (parameters->sampling_scheme) {
  case SO3_SAMPLING_MW: return 2.0 * a * SO3_PI / (2.0 * L - 1.0);
  case SO3_SAMPLING_MW_SS: return 2.0 * a * SO3_PI / (2.0 * L);
  default: SO3_ERROR_GENERIC("Invalid sampling scheme.");
  }
}

/* so3_sampling.c:255-267 */
double so3_sampling_b2beta(int b, const so3_parameters_t *parameters) {
  const int L = parameters->L;
  switch (parameters->sampling_scheme) {
  case SO3_SAMPLING_MW: return (2.0 * b + 1.0) * SO3_PI / (2.0 * L - 1.0);
  case SO3_SAMPLING_MW_SS: return 2.0 * b * SO3_PI / (2.0 * L);
  default: SO3_ERROR_GENERIC("Invalid sampling scheme.");
  }
}

/* so3_sampling.c:286-297 */
double so3_sampling_g2gamma(int g, const so3_parameters_t *parameters) {
  const int N = parameters->N;
  if (parameters->steerable) return g * SO3_PI / N;
  return 2.0 * g * SO3_PI / (2.0 * N - 1.0);
}

/* so3_sampling.c:316-339.  Compact sizes use sum_{k<N} k^2 = (N-1)N(2N-1)/6. */
int so3_sampling_flmn_size(const so3_parameters_t *parameters) {
  const int L = parameters->L, N = parameters->N;
  switch (parameters->storage) {
  case SO3_STORAGE_PADDED:
    return parameters->reality ? N * L * L : (2 * N - 1) * L * L;
  case SO3_STORAGE_COMPACT:
    return parameters->reality ? N * (6 * L * L - (N - 1) * (2 * N - 1)) / 6
                               : (2 * N - 1) * (3 * L * L - N * (N - 1)) / 3;
  default: SO3_ERROR_GENERIC("Invalid storage method.");
  }
}

/* so3_sampling.c:369-426 */
void so3_sampling_elmn2ind(int *ind, int el, int m, int n,
                           const so3_parameters_t *parameters) {
  const int L = parameters->L, N = parameters->N;
  const int L2 = L * L;
  int absn, offset;
  switch (parameters->storage) {
  case SO3_STORAGE_PADDED:
    switch (parameters->n_order) {
    case SO3_N_ORDER_ZERO_FIRST:
      *ind = ((n < 0) ? -2 * n - 1 : 2 * n) * L2 + el * el + el + m;
      return;
    case SO3_N_ORDER_NEGATIVE_FIRST:
      *ind = (N - 1 + n) * L2 + el * el + el + m;
      return;
    default: SO3_ERROR_GENERIC("Invalid n-order.");
    }
  case SO3_STORAGE_COMPACT:
    absn = abs(n);
    if (parameters->n_order != SO3_N_ORDER_ZERO_FIRST &&
        parameters->n_order != SO3_N_ORDER_NEGATIVE_FIRST)
      SO3_ERROR_GENERIC("Invalid n-order.");
    if (absn > el)
      SO3_ERROR_GENERIC("Tried to access component with n > l in compact storage.");
    if (parameters->n_order == SO3_N_ORDER_ZERO_FIRST) {
      /* blocks 0, -1, 1, -2, 2, ...; block n holds L*L - n*n entries */
      offset = (2 * absn - 1) * (3 * L2 - absn * (absn - 1)) / 3;
      if (n >= 0) offset += L2 - n * n;
    } else {
      offset = (N - 1 + n) * L2 - (2 * N - 1) * (N - 1) * N / 6;
      if (n <= 0)
        offset += absn * (2 * absn + 1) * (absn + 1) / 6;
      else
        offset -= absn * (2 * absn - 1) * (absn - 1) / 6;
    }
    *ind = offset + el * el - n * n + el + m;
    return;
  default: SO3_ERROR_GENERIC("Invalid storage method.");
  }
}

/* so3_sampling.c:455-542 */
void so3_sampling_ind2elmn(int *el, int *m, int *n, int ind,
                           const so3_parameters_t *parameters) {
  const int L = parameters->L, N = parameters->N;
  const int L2 = L * L;
  int offset;
  switch (parameters->storage) {
  case SO3_STORAGE_PADDED:
    switch (parameters->n_order) {
    case SO3_N_ORDER_ZERO_FIRST: {
      const int blk = ind / L2;
      *n = (blk % 2) ? -(blk + 1) / 2 : blk / 2;
      break;
    }
    case SO3_N_ORDER_NEGATIVE_FIRST:
      *n = ind / L2 - (N - 1);
      break;
    default: SO3_ERROR_GENERIC("Invalid n-order.");
    }
    ind %= L2;
    *el = (int)sqrt((double)ind);
    *m = ind - (*el) * (*el) - *el;
    return;
  case SO3_STORAGE_COMPACT:
    switch (parameters->n_order) {
    case SO3_N_ORDER_ZERO_FIRST:
      /* walk the blocks 0, -1, 1, -2, 2, ... until ind falls inside one */
      offset = 0;
      *n = 0;
      while (ind + offset >= L2) {
        ind -= L2 - offset;
        if (*n >= 0) {
          *n = -(*n + 1);
          offset = (*n) * (*n);
        } else {
          *n = -(*n);
        }
      }
      break;
    case SO3_N_ORDER_NEGATIVE_FIRST:
      *n = -N + 1;
      offset = (*n) * (*n);
      while (ind + offset >= L2) {
        ind -= L2 - offset;
        (*n)++;
        offset = (*n) * (*n);
      }
      break;
    default: SO3_ERROR_GENERIC("Invalid n-order.");
    }
    ind += offset;
    *el = (int)sqrt((double)ind);
    *m = ind - (*el) * (*el) - *el;
    return;
  default: SO3_ERROR_GENERIC("Invalid storage method.");
  }
}

/* so3_sampling.c:571-597: real signals store n >= 0 only, in negative-first
 * order shifted so that (0,0,0) has index 0. */
void so3_sampling_elmn2ind_real(int *ind, int el, int m, int n,
                                const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  int base;
  if (q.storage != SO3_STORAGE_PADDED && q.storage != SO3_STORAGE_COMPACT)
    SO3_ERROR_GENERIC("Invalid storage method.");
  q.n_order = SO3_N_ORDER_NEGATIVE_FIRST;
  so3_sampling_elmn2ind(&base, 0, 0, 0, &q);
  so3_sampling_elmn2ind(ind, el, m, n, &q);
  *ind -= base;
}

/* so3_sampling.c:626-650 */
void so3_sampling_ind2elmn_real(int *el, int *m, int *n, int ind,
                                const so3_parameters_t *parameters) {
  so3_parameters_t q = *parameters;
  int base;
  if (q.storage != SO3_STORAGE_PADDED && q.storage != SO3_STORAGE_COMPACT)
    SO3_ERROR_GENERIC("Invalid storage method.");
  q.n_order = SO3_N_ORDER_NEGATIVE_FIRST;
  so3_sampling_elmn2ind(&base, 0, 0, 0, &q);
  so3_sampling_ind2elmn(el, m, n, base + ind, &q);
}

/* so3_sampling.c:677-758 */
void so3_sampling_n_loop_values(int *n_start, int *n_stop, int *n_inc,
                                const so3_parameters_t *parameters) {
  const int N = parameters->N;
  const int top_even = ((N - 1) % 2 == 0);
  if (parameters->reality) {
    switch (parameters->n_mode) {
    case SO3_N_MODE_ALL:
    case SO3_N_MODE_L:
      *n_start = 0; *n_stop = N - 1; *n_inc = 1; break;
    case SO3_N_MODE_EVEN:
      *n_start = 0; *n_stop = top_even ? N - 1 : N - 2; *n_inc = 2; break;
    case SO3_N_MODE_ODD:
      *n_start = 1; *n_stop = !top_even ? N - 1 : N - 2; *n_inc = 2; break;
    case SO3_N_MODE_MAXIMUM:
      *n_start = N - 1; *n_stop = N - 1; *n_inc = 1; break;
    default: SO3_ERROR_GENERIC("Invalid n-mode.");
    }
  } else {
    switch (parameters->n_mode) {
    case SO3_N_MODE_ALL:
    case SO3_N_MODE_L:
      *n_start = -N + 1; *n_stop = N - 1; *n_inc = 1; break;
    case SO3_N_MODE_EVEN:
      *n_start = top_even ? -N + 1 : -N + 2;
      *n_stop = top_even ? N - 1 : N - 2;
      *n_inc = 2;
      break;
    case SO3_N_MODE_ODD:
      *n_start = !top_even ? -N + 1 : -N + 2;
      *n_stop = !top_even ? N - 1 : N - 2;
      *n_inc = 2;
      break;
    case SO3_N_MODE_MAXIMUM:
      *n_start = -N + 1; *n_stop = N - 1; *n_inc = (N > 1) ? 2 * N - 2 : 1; break;
    default: SO3_ERROR_GENERIC("Invalid n-mode.");
    }
  }
}

/* so3_sampling.c:786-793 */
void so3_sampling_el_loop_values(int *el_start, int *el_stop, int *el_inc,
                                 const int n, const so3_parameters_t *parameters) {
  const int an = abs(n);
  *el_start = parameters->L0 > an ? parameters->L0 : an;
  *el_stop = (parameters->n_mode == SO3_N_MODE_L) ? an : parameters->L - 1;
  *el_inc = 1;
}

/* so3_sampling.c:812-817 */
void so3_sampling_m_loop_values(int *m_start, int *m_stop, int *m_inc, const int el) {
  *m_start = -el;
  *m_stop = el;
  *m_inc = 1;
}

/* so3_sampling.c:825-830 */
bool so3_sampling_is_i_in_loop_range(const int i, const int i_start, const int i_stop,
                                     const int i_inc) {
  if (i < i_start || i > i_stop) return false;
  return (i - i_start) % i_inc == 0;
}

/* so3_sampling.c:856-875 */
bool so3_sampling_is_elmn_non_zero(const int el, const int m, const int n,
                                   const so3_parameters_t *parameters) {
  int s, e, d;
  so3_sampling_n_loop_values(&s, &e, &d, parameters);
  if (!so3_sampling_is_i_in_loop_range(n, s, e, d)) return false;
  so3_sampling_el_loop_values(&s, &e, &d, n, parameters);
  if (!so3_sampling_is_i_in_loop_range(el, s, e, d)) return false;
  so3_sampling_m_loop_values(&s, &e, &d, el);
  return so3_sampling_is_i_in_loop_range(m, s, e, d);
}

/* so3_sampling.c:877-881 */
int so3_sampling_is_elmn_non_zero_return_int(const int el, const int m, const int n,
                                             const so3_parameters_t *parameters) {
  return (int)so3_sampling_is_elmn_non_zero(el, m, n, parameters);
}
