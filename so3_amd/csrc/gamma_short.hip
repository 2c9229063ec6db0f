// gamma_short.hip -- the gamma DFT of short lengths as a direct DFT: one
// column per thread for nG = 2N - 1 <= 31 (N <= 16) and for real signals up
// to 63 (N <= 32), whose c2r / r2c forms keep only N - 1 complex (or
// 2 (N - 1) real) inputs per column in registers; complex signals with nG in
// [33, 63] take one column per pair of lanes (see k_gamma_dft2 below).
//
// Reference counterparts: inverse gamma, backward DFT over n
// (so3_core.c:137-158, 216; c2r :564-576); forward gamma, forward DFT over g
// times 2 pi / (2N-1) (so3_core.c:340-370; r2c :779-812).
//
// Column p = b nA + a of f (b < nB, a < nA) is column (b, a) of X2, whose n
// slot k sits at x2_row(k, b) ldA + a; sample g of column p of f sits at
// g S + p, S = nB nA.  Consecutive threads take consecutive p, so one access
// instruction moves 64 consecutive entries of one row (the workgroup engine
// instead staged K columns in LDS and ran its passes with a workgroup barrier
// each, at 2.5-9x the HBM time of the stage for these lengths).
//
// With H = (nG - 1) / 2, the sums u_j = x_j + x_{nG-j} and differences
// v_j = x_j - x_{nG-j} (1 <= j <= H) give each pair of outputs k, nG - k of
// the kernel exp(DIR 2 pi i jk / nG) (c, s: cosine and sine of 2 pi jk / nG):
//   A = x0.re + sum_j u_j.re c,   D = x0.im + sum_j u_j.im c,
//   B = sum_j v_j.im s,           C = sum_j v_j.re s,
//   X(k) = (A - DIR B) + i (D + DIR C),   X(nG-k) = (A + DIR B) + i (D - DIR C),
// 4 H^2 multiply-adds per column.  Real signals (n >= 0 stored, x_0 real):
//   inverse (c2r): f(g) = x0 + 2 sum_{n=1}^{H} (x_n.re c - x_n.im s), and
//                  f(nG-g) the same with +s;
//   forward (r2c): X(n) = f0 + sum_{g=1}^{H} (u_g c - i v_g s), n <= H.
// The cosines and sines come from a table indexed by jk mod nG, a constant of
// each unrolled term, so their addresses are uniform (scalar loads).
#include <map>
#include <mutex>
#include <vector>

#include "kernels.hpp"

namespace so3g {

namespace {

template <int NG, int REAL>
__global__ __launch_bounds__(256) void k_inv_gamma_dft(Geo g, const double2 *__restrict__ cs,
                                                       const double2 *__restrict__ X2, double2 *__restrict__ fc,
                                                       double *__restrict__ fr, double scale) {
  constexpr int H = (NG - 1) / 2;
  const int S = g.nB * g.nA;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= S) return;
  const int b = p / g.nA, a = p - b * g.nA;
  const double2 *col = X2 + x2_row(g, 0, b) * g.ldA + a;
  const int64_t xs = (int64_t)x2_w(g, b) * g.ldA;
  if (REAL) {
    double xr[H + 1], xi[H + 1];
#pragma unroll
    for (int n = 0; n <= H; ++n) {
      const double2 v = col[n * xs];
      xr[n] = v.x;
      xi[n] = v.y;
    }
    double s0 = 0.0;
#pragma unroll
    for (int n = 1; n <= H; ++n) s0 += xr[n];
    fr[p] = (xr[0] + 2.0 * s0) * scale;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      double P = 0.0, Q = 0.0;
#pragma unroll
      for (int n = 1; n <= H; ++n) {
        const double2 t = cs[(n * k) % NG];
        P = fma(xr[n], t.x, P);
        Q = fma(xi[n], t.y, Q);
      }
      fr[(int64_t)k * S + p] = (xr[0] + 2.0 * (P - Q)) * scale;
      fr[(int64_t)(NG - k) * S + p] = (xr[0] + 2.0 * (P + Q)) * scale;
    }
  } else {
    const double2 x0 = col[0];
    double2 u[H + 1], v[H + 1];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const double2 y = col[j * xs], z = col[(NG - j) * xs];
      u[j] = make_double2(y.x + z.x, y.y + z.y);
      v[j] = make_double2(y.x - z.x, y.y - z.y);
    }
    double2 s0 = x0;
#pragma unroll
    for (int j = 1; j <= H; ++j) s0 = make_double2(s0.x + u[j].x, s0.y + u[j].y);
    fc[p] = make_double2(s0.x * scale, s0.y * scale);
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      double A = x0.x, D = x0.y, B = 0.0, C = 0.0;
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const double2 t = cs[(j * k) % NG];
        A = fma(u[j].x, t.x, A);
        D = fma(u[j].y, t.x, D);
        B = fma(v[j].y, t.y, B);
        C = fma(v[j].x, t.y, C);
      }
      // DIR = +1
      fc[(int64_t)k * S + p] = make_double2((A - B) * scale, (D + C) * scale);
      fc[(int64_t)(NG - k) * S + p] = make_double2((A + B) * scale, (D - C) * scale);
    }
  }
}

template <int NG, int REAL>
__global__ __launch_bounds__(256) void k_fwd_gamma_dft(Geo g, const double2 *__restrict__ cs,
                                                       const double2 *__restrict__ fc,
                                                       const double *__restrict__ fr, double2 *__restrict__ X2,
                                                       double sc) {
  constexpr int H = (NG - 1) / 2;
  const int S = g.nB * g.nA;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= S) return;
  const int b = p / g.nA, a = p - b * g.nA;
  double2 *col = X2 + x2_row(g, 0, b) * g.ldA + a;
  const int64_t xs = (int64_t)x2_w(g, b) * g.ldA;
  if (REAL) {
    const double f0 = fr[p];
    double u[H + 1], v[H + 1];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const double y = fr[(int64_t)j * S + p], z = fr[(int64_t)(NG - j) * S + p];
      u[j] = y + z;
      v[j] = y - z;
    }
    double s0 = f0;
#pragma unroll
    for (int j = 1; j <= H; ++j) s0 += u[j];
    col[0] = make_double2(s0 * sc, 0.0);
#pragma unroll
    for (int n = 1; n <= H; ++n) {
      double re = f0, im = 0.0;
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const double2 t = cs[(j * n) % NG];
        re = fma(u[j], t.x, re);
        im = fma(v[j], t.y, im);
      }
      col[n * xs] = make_double2(re * sc, -im * sc);
    }
  } else {
    const double2 x0 = fc[p];
    double2 u[H + 1], v[H + 1];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const double2 y = fc[(int64_t)j * S + p], z = fc[(int64_t)(NG - j) * S + p];
      u[j] = make_double2(y.x + z.x, y.y + z.y);
      v[j] = make_double2(y.x - z.x, y.y - z.y);
    }
    double2 s0 = x0;
#pragma unroll
    for (int j = 1; j <= H; ++j) s0 = make_double2(s0.x + u[j].x, s0.y + u[j].y);
    col[0] = make_double2(s0.x * sc, s0.y * sc);
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      double A = x0.x, D = x0.y, B = 0.0, C = 0.0;
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const double2 t = cs[(j * k) % NG];
        A = fma(u[j].x, t.x, A);
        D = fma(u[j].y, t.x, D);
        B = fma(v[j].y, t.y, B);
        C = fma(v[j].x, t.y, C);
      }
      // DIR = -1
      col[k * xs] = make_double2((A + B) * sc, (D - C) * sc);
      col[(NG - k) * xs] = make_double2((A - B) * sc, (D + C) * sc);
    }
  }
}

// Complex gamma for nG in [33, 63] with one column per pair of lanes: the
// even lane of a pair carries the real parts of the column and the odd lane
// the imaginary parts.  Each lane forms u_j, v_j of its part and the two sums
//   even (re):  A = x0.re + sum_j u_j.re c,   C = sum_j v_j.re s
//   odd  (im):  D = x0.im + sum_j u_j.im c,   B = sum_j v_j.im s
// with the same twiddle indices in both lanes (uniform: scalar loads); one
// swap with the partner lane (B to the even lane, C to the odd one) gives
//   X(k).re = A - DIR B, X(nG-k).re = A + DIR B   (even lane)
//   X(k).im = D + DIR C, X(nG-k).im = D - DIR C   (odd lane)
// 2 H^2 multiply-adds and 2 H doubles of inputs per lane (H <= 31), and the
// two lanes' 8-byte accesses meet in each 16-byte entry.  DIR = +1: X2 -> f
// (inverse, times scale); -1: f -> X2 (forward, times scale).
__device__ __forceinline__ double swap_lane_pair(double x) {  // x of lane ^ 1
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int NG, int DIR>
__global__ __launch_bounds__(256) void k_gamma_dft2(Geo g, const double2 *__restrict__ cs,
                                                    const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                    double scale) {
  constexpr int H = (NG - 1) / 2;
  const int S = g.nB * g.nA;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = tid >> 1, part = tid & 1;
  // every lane computes (a pair past the last column repeats it) so that the
  // partner swaps see active lanes; only the stores are predicated
  const bool ok = c < S;
  const int cc = ok ? c : S - 1;
  const int b = cc / g.nA, a = cc - b * g.nA;
  const double2 *X2c = DIR > 0 ? src : dst;
  const double *xcol = (const double *)(X2c + x2_row(g, 0, b) * g.ldA + a) + part;
  const int64_t xs2 = (int64_t)x2_w(g, b) * g.ldA * 2;  // X2 slot step in doubles
  double u[H + 1], v[H + 1], x0;
  if (DIR > 0) {
    x0 = xcol[0];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const double y = xcol[j * xs2], z = xcol[(NG - j) * xs2];
      u[j] = y + z;
      v[j] = y - z;
    }
  } else {
    const double *fcol = (const double *)(src + cc) + part;
    const int64_t fs2 = (int64_t)S * 2;
    x0 = fcol[0];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const double y = fcol[j * fs2], z = fcol[(NG - j) * fs2];
      u[j] = y + z;
      v[j] = y - z;
    }
  }
  auto store = [&](int k, double val) {
    if (!ok) return;
    if (DIR > 0) ((double *)(dst + (int64_t)k * S + cc))[part] = val * scale;
    else ((double *)(dst + x2_row(g, 0, b) * g.ldA + a))[part + k * xs2] = val * scale;
  };
  double s0 = x0;
#pragma unroll
  for (int j = 1; j <= H; ++j) s0 += u[j];
  store(0, s0);
  const double sg = part == 0 ? -(double)DIR : (double)DIR;
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    double P = x0, Q = 0.0;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const double2 t = cs[(j * k) % NG];
      P = fma(u[j], t.x, P);
      Q = fma(v[j], t.y, Q);
    }
    const double o = swap_lane_pair(Q);
    store(k, P + sg * o);
    store(NG - k, P - sg * o);
  }
}

// cosines and sines of 2 pi m / nG (m < nG), rounded from long double; one
// table per (device, nG), kept for the lifetime of the process
const double2 *short_table(int nG) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, double2 *> cache;
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(dev, nG);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  std::vector<double2> h(nG);
  const long double tau = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < nG; ++m) {
    const long double t = tau * (long double)m / (long double)nG;
    h[m] = make_double2((double)cosl(t), (double)sinl(t));
  }
  double2 *d = nullptr;
  SO3_HIP(hipMalloc(&d, sizeof(double2) * nG));
  SO3_HIP(hipMemcpy(d, h.data(), sizeof(double2) * nG, hipMemcpyHostToDevice));
  cache[key] = d;
  return d;
}

template <int NG, int REAL>
void go_short(const Geo &g, const void *src, void *dst, int dir, double scale, hipStream_t st) {
  const double2 *cs = short_table(NG);
  const int64_t S = (int64_t)g.nB * g.nA;
  const unsigned grid = (unsigned)((S + 255) / 256);
  if (dir > 0) {
    if constexpr (REAL != 0)
      k_inv_gamma_dft<NG, 1><<<grid, 256, 0, st>>>(g, cs, (const double2 *)src, nullptr, (double *)dst, scale);
    else
      k_inv_gamma_dft<NG, 0><<<grid, 256, 0, st>>>(g, cs, (const double2 *)src, (double2 *)dst, nullptr, scale);
  } else {
    if constexpr (REAL != 0)
      k_fwd_gamma_dft<NG, 1><<<grid, 256, 0, st>>>(g, cs, nullptr, (const double *)src, (double2 *)dst, scale);
    else
      k_fwd_gamma_dft<NG, 0><<<grid, 256, 0, st>>>(g, cs, (const double2 *)src, nullptr, (double2 *)dst, scale);
  }
  SO3_LAUNCH_CHECK();
}

typedef void (*ShortFn)(const Geo &, const void *, void *, int, double, hipStream_t);
#define SO3_SHORT16(R, B)                                                                                    \
  go_short<B + 1, R>, go_short<B + 3, R>, go_short<B + 5, R>, go_short<B + 7, R>, go_short<B + 9, R>,       \
      go_short<B + 11, R>, go_short<B + 13, R>, go_short<B + 15, R>, go_short<B + 17, R>, go_short<B + 19, R>, \
      go_short<B + 21, R>, go_short<B + 23, R>, go_short<B + 25, R>, go_short<B + 27, R>, go_short<B + 29, R>, \
      go_short<B + 31, R>
const ShortFn kShortC[16] = {SO3_SHORT16(0, 0)};                      // nG = 1 .. 31
const ShortFn kShortR[32] = {SO3_SHORT16(1, 0), SO3_SHORT16(1, 32)};  // nG = 1 .. 63
#undef SO3_SHORT16

// complex nG in [33, 63]: one column per pair of lanes
template <int NG>
void go_short2(const Geo &g, const void *src, void *dst, int dir, double scale, hipStream_t st) {
  const double2 *cs = short_table(NG);
  const int64_t S = (int64_t)g.nB * g.nA;
  const unsigned grid = (unsigned)((2 * S + 255) / 256);
  if (dir > 0)
    k_gamma_dft2<NG, +1><<<grid, 256, 0, st>>>(g, cs, (const double2 *)src, (double2 *)dst, scale);
  else
    k_gamma_dft2<NG, -1><<<grid, 256, 0, st>>>(g, cs, (const double2 *)src, (double2 *)dst, scale);
  SO3_LAUNCH_CHECK();
}
const ShortFn kShortC2[16] = {go_short2<33>, go_short2<35>, go_short2<37>, go_short2<39>, go_short2<41>,
                              go_short2<43>, go_short2<45>, go_short2<47>, go_short2<49>, go_short2<51>,
                              go_short2<53>, go_short2<55>, go_short2<57>, go_short2<59>, go_short2<61>,
                              go_short2<63>};

}  // namespace

bool gamma_short_applies(const Geo &g) {
  // the kernels index columns with 32-bit threads (two per column above 31)
  const int64_t S = (int64_t)g.nB * g.nA;
  return g.N >= 1 && 2 * g.N - 1 <= kGammaShortMax && S < ((int64_t)1 << 30);
}

void gamma_short_prepare(const Geo &g) {
  if (gamma_short_applies(g)) (void)short_table(2 * g.N - 1);
}

bool launch_gamma_short(const Geo &g, const void *src, void *dst, int dir, double scale, hipStream_t st) {
  if (!gamma_short_applies(g)) return false;
  if (g.real) kShortR[g.N - 1](g, src, dst, dir, scale, st);
  else if (g.N <= 16) kShortC[g.N - 1](g, src, dst, dir, scale, st);
  else kShortC2[g.N - 17](g, src, dst, dir, scale, st);
  return true;
}

}  // namespace so3g
