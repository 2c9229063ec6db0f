// dft_api.hip -- so3x_dft: the library's batched DFT kernel on contiguous
// sequences, exported so that its numerics can be tested in isolation.
#include <algorithm>
#include <string>

#include "fft_plan.hpp"
#include "kernels.hpp"
#include "so3/so3x.h"

namespace so3g {
namespace {

template <int DIR>
__global__ __launch_bounds__(256) void k_dft(FftDesc fd, int K, int ld, double2 *__restrict__ X,
                                             int howmany) {
  extern __shared__ double2 lds[];
  double2 *A = lds, *tw = lds + K * ld;
  fft_load_twiddles(tw, fd);
  const int n = fd.n;
  const int64_t r0 = (int64_t)blockIdx.x * K;
  for (int idx = threadIdx.x; idx < K * n; idx += blockDim.x) {
    const int s = idx / n, k = idx - s * n;
    A[s * ld + k] = (r0 + s < howmany) ? X[(r0 + s) * n + k] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  fft_run<DIR>(A, K, ld, fd, tw);
  for (int idx = threadIdx.x; idx < K * n; idx += blockDim.x) {
    const int s = idx / n, k = idx - s * n;
    if (r0 + s < howmany) X[(r0 + s) * n + k] = A[s * ld + k];
  }
}

}  // namespace
}  // namespace so3g


extern "C" int so3x_dft(void *x, int n, int howmany, int dir, void *stream) {
  using namespace so3g;
  try {
    if (n < 1 || howmany < 0 || (dir != 1 && dir != -1)) throw HipError("bad so3x_dft arguments");
    const FftDesc fd = fft_plan(n);
    const int ld = fft_ld(fd);
    const int cap = fft_capacity(fd);
    if (fd.len > cap) throw HipError("DFT length exceeds the register budget");
    const int tb = fd.len * (int)sizeof(double2);
    int K = std::max(1, std::min(8, cap / fd.len));
    while (K > 1 && (size_t)K * ld * sizeof(double2) + tb > (size_t)lds_limit()) --K;
    const size_t sh = (size_t)K * ld * sizeof(double2) + tb;
    if (sh > (size_t)lds_limit()) throw HipError("DFT length exceeds the LDS budget");
    const unsigned grid = (unsigned)((howmany + K - 1) / K);
    if (grid == 0) return 0;
    if (dir < 0) {
      if (sh > 65536) SO3_HIP(hipFuncSetAttribute((const void *)k_dft<-1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
      k_dft<-1><<<grid, 256, sh, (hipStream_t)stream>>>(fd, K, ld, (double2 *)x, howmany);
    } else {
      if (sh > 65536) SO3_HIP(hipFuncSetAttribute((const void *)k_dft<+1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
      k_dft<+1><<<grid, 256, sh, (hipStream_t)stream>>>(fd, K, ld, (double2 *)x, howmany);
    }
    SO3_LAUNCH_CHECK();
    return 0;
  } catch (const std::exception &e) {
    fprintf(stderr, "so3x_dft: %s\n", e.what());
    return -1;
  }
}
