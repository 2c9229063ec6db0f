// conv.hpp -- harmonic convolutions on the GPU (device pointers, stream-ordered,
// synchronous on return).
#pragma once
#include "gpu_common.hpp"
#include "so3/so3_types.h"

namespace so3g {

void s2toso3(const so3_parameters_t *hp, const void *flm, const void *glm, void *hlmn, hipStream_t st);
void harmonic_convolution(const so3_parameters_t *hp, const so3_parameters_t *fp,
                          const so3_parameters_t *gp, const void *f, const void *g, void *h,
                          hipStream_t st);

}  // namespace so3g
