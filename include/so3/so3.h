/* so3.h -- aggregate header (reference include/so3/so3.h:5-10). */
#ifndef SO3_H
#define SO3_H
#include "so3_types.h"
#include "so3_error.h"
#include "so3_sampling.h"
#include "so3_core.h"
#include "so3_adjoint.h"
#include "so3_conv.h"
#include "so3x.h"
#endif
