/*
 * so3_sampling.h -- grid sizes, sample angles, quadrature weights and flmn
 * index maps.  Declarations match the reference include/so3/so3_sampling.h:15-49
 * one for one; results are bit-identical to src/c/so3_sampling.c.
 */
#ifndef SO3_SAMPLING
#define SO3_SAMPLING

#include <stdbool.h>
#include "so3_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* so3_sampling.c:32-51 */
SO3_COMPLEX(double) so3_sampling_weight(const so3_parameters_t *parameters, int p);

/* so3_sampling.c:80-206 */
int so3_sampling_f_size(const so3_parameters_t *parameters);
int so3_sampling_n(const so3_parameters_t *parameters);
int so3_sampling_nalpha(const so3_parameters_t *parameters);
int so3_sampling_nbeta(const so3_parameters_t *parameters);
int so3_sampling_ngamma(const so3_parameters_t *parameters);

/* so3_sampling.c:224-297 */
double so3_sampling_a2alpha(int a, const so3_parameters_t *parameters);
double so3_sampling_b2beta(int b, const so3_parameters_t *parameters);
double so3_sampling_g2gamma(int g, const so3_parameters_t *parameters);

/* so3_sampling.c:316-650 */
int so3_sampling_flmn_size(const so3_parameters_t *parameters);
void so3_sampling_elmn2ind(int *ind, int el, int m, int n,
                           const so3_parameters_t *parameters);
void so3_sampling_ind2elmn(int *el, int *m, int *n, int ind,
                           const so3_parameters_t *parameters);
void so3_sampling_elmn2ind_real(int *ind, int el, int m, int n,
                                const so3_parameters_t *parameters);
void so3_sampling_ind2elmn_real(int *el, int *m, int *n, int ind,
                                const so3_parameters_t *parameters);

/* so3_sampling.c:677-882 */
void so3_sampling_n_loop_values(int *n_start, int *n_stop, int *n_inc,
                                const so3_parameters_t *parameters);
void so3_sampling_el_loop_values(int *el_start, int *el_stop, int *el_inc,
                                 const int n, const so3_parameters_t *parameters);
void so3_sampling_m_loop_values(int *m_start, int *m_stop, int *m_inc, const int el);
bool so3_sampling_is_elmn_non_zero(const int el, const int m, const int n,
                                   const so3_parameters_t *parameters);
int so3_sampling_is_elmn_non_zero_return_int(const int el, const int m, const int n,
                                             const so3_parameters_t *parameters);
/* Exported but undeclared in the reference (so3_sampling.c:825-830); declared
 * here so link-compatible callers resolve it. */
bool so3_sampling_is_i_in_loop_range(const int i, const int i_start,
                                     const int i_stop, const int i_inc);

#ifdef __cplusplus
}
#endif
#endif
