// conv.hpp -- harmonic convolutions on the GPU (device pointers, stream-ordered,
// synchronous on return).
#pragma once
#include "gpu_common.hpp"
#include "so3/so3_types.h"

namespace so3g {

// Enqueues h_lmn = s2toso3(flm, glm) on st (no allocation once the
// parameter block has been seen: cached_bases builds its index table).
void s2toso3(const so3_parameters_t *hp, const void *flm, const void *glm, void *hlmn, hipStream_t st);
const int64_t *cached_bases(const so3_parameters_t *hp);
void release_conv_caches();
void harmonic_convolution(const so3_parameters_t *hp, const so3_parameters_t *fp,
                          const so3_parameters_t *gp, const void *f, const void *g, void *h,
                          hipStream_t st);

}  // namespace so3g
