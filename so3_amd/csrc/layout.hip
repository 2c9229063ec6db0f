// layout.hip -- the MATLAB sample layout of the reference's mex interface.
//
// The library (like the reference C API) stores a signal on the SO(3) grid as
// f[g nA nB + b nA + a], alpha fastest.  The reference's MATLAB glue hands
// MATLAB the array f_m[a nG nB + b nG + g], gamma fastest, with the real and
// imaginary parts in separate arrays (src/matlab/so3_inverse_mex.c:310-339 for
// the output, so3_forward_mex.c:78-97 for the input).  This file converts
// between the two on the GPU: for every beta sample b the (g, a) plane is
// transposed through an LDS tile, loads coalesced along the source's fast
// axis, stores coalesced along the destination's.  HBM-bound: each sample is
// read once and written once.
#include "gpu_common.hpp"
#include "so3/so3_sampling.h"

namespace so3g {

namespace {

constexpr int kTA = 64;  // alpha per tile
constexpr int kTG = 16;  // gamma per tile

// to_mex: f (complex if im != null, else real) -> re, im in MATLAB order
// from_mex: the reverse.  Grid: (ceil(nA / kTA), ceil(nG / kTG), nB).
template <bool TO_MEX, bool CPLX>
__global__ __launch_bounds__(256) void k_mex_layout(int nA, int nB, int nG, double *__restrict__ f,
                                                    double *__restrict__ re, double *__restrict__ im) {
  __shared__ double tr[kTG][kTA + 1];
  __shared__ double ti[kTG][kTA + 1];
  const int a0 = blockIdx.x * kTA, g0 = blockIdx.y * kTG, b = blockIdx.z;
  const int tid = threadIdx.x;
  const int64_t plane = (int64_t)nA * nB;
  if (TO_MEX) {
    // load along alpha: thread (ta, gy), rows g0 + gy + 4 k
    const int ta = tid & 63, gy = tid >> 6;
    for (int k = 0; k < kTG / 4; ++k) {
      const int gl = gy + 4 * k, g = g0 + gl, a = a0 + ta;
      if (g < nG && a < nA) {
        const int64_t s = (int64_t)g * plane + (int64_t)b * nA + a;
        if (CPLX) {
          tr[gl][ta] = f[2 * s];
          ti[gl][ta] = f[2 * s + 1];
        } else {
          tr[gl][ta] = f[s];
        }
      }
    }
    __syncthreads();
    // store along gamma: thread (tg, ay), columns a0 + ay + 16 k
    const int tg = tid & 15, ay = tid >> 4;
    for (int k = 0; k < kTA / 16; ++k) {
      const int al = ay + 16 * k, a = a0 + al, g = g0 + tg;
      if (g < nG && a < nA) {
        const int64_t d = (int64_t)a * nG * nB + (int64_t)b * nG + g;
        re[d] = tr[tg][al];
        if (CPLX) im[d] = ti[tg][al];
      }
    }
  } else {
    // load along gamma from the MATLAB arrays
    const int tg = tid & 15, ay = tid >> 4;
    for (int k = 0; k < kTA / 16; ++k) {
      const int al = ay + 16 * k, a = a0 + al, g = g0 + tg;
      if (g < nG && a < nA) {
        const int64_t s = (int64_t)a * nG * nB + (int64_t)b * nG + g;
        tr[tg][al] = re[s];
        ti[tg][al] = CPLX && im ? im[s] : 0.0;
      }
    }
    __syncthreads();
    const int ta = tid & 63, gy = tid >> 6;
    for (int k = 0; k < kTG / 4; ++k) {
      const int gl = gy + 4 * k, g = g0 + gl, a = a0 + ta;
      if (g < nG && a < nA) {
        const int64_t d = (int64_t)g * plane + (int64_t)b * nA + a;
        if (CPLX) {
          f[2 * d] = tr[gl][ta];
          f[2 * d + 1] = ti[gl][ta];
        } else {
          f[d] = tr[gl][ta];
        }
      }
    }
  }
}

}  // namespace

void launch_mex_layout(bool to_mex, void *f, double *re, double *im, const so3_parameters_t *p, hipStream_t st) {
  if (!p || !f || !re) throw HipError("NULL pointer");
  if (p->L < 1 || p->N < 1) throw HipError("L and N must be positive");
  if (p->sampling_scheme != SO3_SAMPLING_MW && p->sampling_scheme != SO3_SAMPLING_MW_SS)
    throw HipError("Invalid sampling scheme.");
  const int nA = so3_sampling_nalpha(p), nB = so3_sampling_nbeta(p), nG = so3_sampling_ngamma(p);
  const bool cplx = !p->reality;
  if (cplx && to_mex && !im) throw HipError("a complex signal needs the imaginary array");
  const dim3 grid((unsigned)((nA + kTA - 1) / kTA), (unsigned)((nG + kTG - 1) / kTG), (unsigned)nB);
  double *fd = (double *)f;
  if (to_mex) {
    if (cplx) k_mex_layout<true, true><<<grid, 256, 0, st>>>(nA, nB, nG, fd, re, im);
    else k_mex_layout<true, false><<<grid, 256, 0, st>>>(nA, nB, nG, fd, re, nullptr);
  } else {
    if (cplx) k_mex_layout<false, true><<<grid, 256, 0, st>>>(nA, nB, nG, fd, re, im);
    else k_mex_layout<false, false><<<grid, 256, 0, st>>>(nA, nB, nG, fd, re, nullptr);
  }
  SO3_LAUNCH_CHECK();
}

}  // namespace so3g
