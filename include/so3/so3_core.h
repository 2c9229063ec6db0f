/*
 * so3_core.h -- forward and inverse Wigner transforms on SO(3).
 *
 * Each entry point replaces the reference function of the same name
 * (include/so3/so3_core.h:15-45; bodies src/c/so3_core.c:51-2385).  Pointers
 * are host pointers (device pointers are also accepted and used in place);
 * every call is synchronous.  All arithmetic runs in FP64 on the GPU.
 */
#ifndef SO3_CORE
#define SO3_CORE

#include "so3_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* so3_core.c:51 */
void so3_core_inverse_via_ssht(SO3_COMPLEX(double) * f, const SO3_COMPLEX(double) * flmn,
                               const so3_parameters_t *parameters);
/* so3_core.c:458 */
void so3_core_inverse_via_ssht_real(double *f, const SO3_COMPLEX(double) * flmn,
                                    const so3_parameters_t *parameters);
/* so3_core.c:246 */
void so3_core_forward_via_ssht(SO3_COMPLEX(double) * flmn, const SO3_COMPLEX(double) * f,
                               const so3_parameters_t *parameters);
/* so3_core.c:675 */
void so3_core_forward_via_ssht_real(SO3_COMPLEX(double) * flmn, const double *f,
                                    const so3_parameters_t *parameters);
/* so3_core.c:914 */
void so3_core_inverse_direct(SO3_COMPLEX(double) * f, const SO3_COMPLEX(double) * flmn,
                             const so3_parameters_t *parameters);
/* so3_core.c:1655 */
void so3_core_inverse_direct_real(double *f, const SO3_COMPLEX(double) * flmn,
                                  const so3_parameters_t *parameters);
/* so3_core.c:1238 */
void so3_core_forward_direct(SO3_COMPLEX(double) * flmn, const SO3_COMPLEX(double) * f,
                             const so3_parameters_t *parameters);
/* so3_core.c:1981 */
void so3_core_forward_direct_real(SO3_COMPLEX(double) * flmn, const double *f,
                                  const so3_parameters_t *parameters);

#ifdef __cplusplus
}
#endif
#endif
