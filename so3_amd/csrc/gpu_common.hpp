// gpu_common.hpp -- shared HIP helpers for the so3 MI355X kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace so3g {

// Every HIP failure becomes a C++ exception carrying the call site; the C
// entry points turn it into the reference's "ERROR: ... exit(1)" behaviour
// (drop-in API) or a status code (so3x_* API).
struct HipError : std::runtime_error {
  explicit HipError(const std::string &s) : std::runtime_error(s) {}
};

#define SO3_HIP(call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      throw ::so3g::HipError(std::string(#call) + ": " + hipGetErrorString(e_) +    \
                             " at " + __FILE__ + ":" + std::to_string(__LINE__));   \
  } while (0)

#define SO3_LAUNCH_CHECK() SO3_HIP(hipGetLastError())

__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by i^k
__host__ __device__ inline double2 cmul_ipow(double2 a, int k) {
  switch (k & 3) {
    case 0: return a;
    case 1: return make_double2(-a.y, a.x);
    case 2: return make_double2(-a.x, -a.y);
    default: return make_double2(a.y, -a.x);
  }
}

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace so3g
