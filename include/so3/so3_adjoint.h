/*
 * so3_adjoint.h -- exact adjoints of the direct transforms.
 * Replaces include/so3/so3_adjoint.h:35-49 (bodies src/c/so3_adjoint.c).
 */
#ifndef SO3_ADJOINT
#define SO3_ADJOINT

#include "so3_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* so3_adjoint.c:34  -- adjoint of so3_core_inverse_direct: f -> flmn */
void so3_adjoint_inverse_direct(SO3_COMPLEX(double) * flmn, const SO3_COMPLEX(double) * f,
                                const so3_parameters_t *parameters);
/* so3_adjoint.c:402 -- adjoint of so3_core_forward_direct: flmn -> f */
void so3_adjoint_forward_direct(SO3_COMPLEX(double) * f, const SO3_COMPLEX(double) * flmn,
                                const so3_parameters_t *parameters);
/* so3_adjoint.c:939 */
void so3_adjoint_inverse_direct_real(SO3_COMPLEX(double) * flmn, const double *f,
                                     const so3_parameters_t *parameters);
/* so3_adjoint.c:1313 */
void so3_adjoint_forward_direct_real(double *f, const SO3_COMPLEX(double) * flmn,
                                     const so3_parameters_t *parameters);

#ifdef __cplusplus
}
#endif
#endif
