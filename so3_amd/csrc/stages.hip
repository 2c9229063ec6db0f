// stages.hip -- DFT stages of the Wigner transform (beta, alpha, gamma).
//
// Every stage is one launch of batched DFTs.  Inner length 1024 (n = 1024,
// or a Bluestein n in [257, 512]) runs register-resident: one sequence per
// wavefront (fft_reg.hpp) or two (fft_reg32.hpp, the faster, used wherever
// the stage is written for it); other lengths run on the LDS-staged
// workgroup engine (fft_device.hpp), a workgroup owning K sequences.
// Reference counterparts:
//   inverse beta : phase exp(i m' pi/(2L-1)), m' < 0 by F(-m') = (-1)^{m+n} F(m'),
//                  DFT over m' -> b (so3_core.c:1148-1206; SSHT inverse)
//   alpha        : DFT over m <-> a (so3_core.c:1168-1206, 1335-1364)
//   inverse gamma: backward DFT over n (so3_core.c:137-158, 216; c2r :564-576)
//   forward gamma: forward DFT over g times 2 pi/(2N-1) (so3_core.c:340-370,
//                  r2c :779-812)
//   forward beta : extension F(2L-2-b) = (-1)^{m+n} F(b), DFT over b, /(2L-1),
//                  phase exp(-i m'' pi/(2L-1)), then the quadrature-weight
//                  convolution G(m') = 2 pi sum_k F(k) w(m'-k) as a product in
//                  the conjugate domain of length Pw >= 4L-3
//                  (so3_core.c:1366-1492), folded to m' >= 0.
#include <algorithm>
#include <type_traits>

#include "fft_plan.hpp"
#include "fft_reg.hpp"
#include "fft_reg16.hpp"
#include "fft_reg2k.hpp"
#include "fft_reg4k.hpp"
#include "fft_reg32.hpp"
#include "kernels.hpp"

namespace so3g {

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559005768;
constexpr int kThreads = kFftThreads;

__device__ inline int n_of_slot(const Geo &g, int slot) {
  if (g.real) return slot;
  return slot < g.N ? slot : slot - g.nG;
}
__device__ inline int m_of_slot(const Geo &g, int slot) { return slot <= g.L - 1 ? slot : slot - g.nA; }

// ---------------------------------------------------------------- inverse beta
__global__ __launch_bounds__(kThreads) void k_inv_beta(Geo g, FftDesc fd, int K, int ld,
                                                       const double2 *__restrict__ X1,
                                                       double2 *__restrict__ X2,
                                                       const int64_t *__restrict__ nbase,
                                                       const double2 *__restrict__ phase) {
  extern __shared__ double2 lds[];
  double2 *A = lds;
  const double2 *tw = fft_load_twiddles(lds + K * ld, fd);
  const int chunks = (g.nA + K - 1) / K;
  const int nslot = blockIdx.x / chunks;
  const int ms0 = (blockIdx.x % chunks) * K;
  const int n = n_of_slot(g, nslot);
  const bool active = nbase[n + g.N - 1] != kNoBlock;
  const int Pb = g.Pb, L = g.L;
  for (int idx = threadIdx.x; idx < K * Pb; idx += blockDim.x) {
    const int s = idx / Pb, k = idx - s * Pb;
    const int mslot = ms0 + s;
    double2 v = make_double2(0.0, 0.0);
    if (active && mslot < g.nA) {
      const int m = m_of_slot(g, mslot);
      const int mp = k <= L - 1 ? k : k - Pb;
      if (m >= -(L - 1) && mp >= -(L - 1)) {
        v = X1[((int64_t)nslot * g.nA + mslot) * L + (mp < 0 ? -mp : mp)];
        if (mp < 0 && ((m + n) & 1)) v = make_double2(-v.x, -v.y);
        v = cmul(v, phase[k]);
      }
    }
    A[s * ld + k] = v;
  }
  __syncthreads();
  fft_run<+1>(A, K, ld, fd, tw);
  for (int idx = threadIdx.x; idx < K * g.nB; idx += blockDim.x) {
    const int b = idx / K, s = idx - b * K;
    if (ms0 + s < g.nA) X2[x2_row(g, nslot, b) * g.ldA + ms0 + s] = A[s * ld + b];
  }
}

// ---------------------------------------------------------------- rows (alpha)
template <int DIR>
__global__ __launch_bounds__(kThreads) void k_rows(FftDesc fd, int K, int ld, double2 *__restrict__ X,
                                                   int64_t nrows, int n, int64_t rs, double scale) {
  extern __shared__ double2 lds[];
  double2 *A = lds;
  const double2 *tw = fft_load_twiddles(lds + K * ld, fd);
  const int64_t r0 = (int64_t)blockIdx.x * K;
  for (int idx = threadIdx.x; idx < K * n; idx += blockDim.x) {
    const int s = idx / n, k = idx - s * n;
    A[s * ld + k] = (r0 + s < nrows) ? X[(r0 + s) * rs + k] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  fft_run<DIR>(A, K, ld, fd, tw);
  for (int idx = threadIdx.x; idx < K * n; idx += blockDim.x) {
    const int s = idx / n, k = idx - s * n;
    if (r0 + s < nrows) X[(r0 + s) * rs + k] = cscale(A[s * ld + k], scale);
  }
}

// ---------------------------------------------------------------- register path
// Transforms whose inner length is 1024 (n = 1024, or a Bluestein n in
// [257, 512]): one sequence per wavefront in the register layouts of
// fft_reg.hpp.  The kernels are persistent: a block stages the twiddle
// tables in LDS once and then takes groups of sequences t = xcd_tile(block)
// + i * gridDim.x until none is left (the grid is what fits on the device at
// once, reg_grid).
constexpr int kRegWaves = kThreads / 64;
// wavefronts per workgroup of the rows kernel, whose wavefronts never
// synchronise: one workgroup per CU sharing one copy of the tables, three
// wavefronts per SIMD
constexpr int kRowsWaves = 12;
// wavefronts per workgroup of the tiled beta kernels (two workgroups per CU)
constexpr int kBetaW = 4;

// Workgroups go round-robin over the 8 XCDs, each with its own L2: block p
// runs on XCD p % 8.  Renumbering block p to (p % 8) q + p / 8, q = n / 8,
// lets the blocks that share an XCD take consecutive groups, so two groups
// that share 128-byte lines (W = 4 complex doubles is half a line) meet in
// one L2 instead of each XCD fetching the whole line.  The last n % 8 blocks
// keep their number.
__device__ __forceinline__ unsigned xcd_tile(unsigned p, unsigned n) {
  const unsigned q = n >> 3;
  return p < 8 * q ? (p & 7) * q + (p >> 3) : p;
}

// The lane index, recomputed in every pass of a persistent loop: an opaque
// value keeps the compiler from hoisting the many lane-dependent addresses
// of a sequence out of the loop, where they would stay live in registers.
__device__ __forceinline__ int loop_lane() {
  int l;
  asm volatile("v_and_b32 %0, 63, %1" : "=v"(l) : "v"((int)threadIdx.x));
  return l;
}

// LDS of a register-path block: `tabs` double2 of tables, then W transpose
// buffers, which the tiled kernels also use as their tile of `tile` double2
template <int W>
constexpr int tile_ld() { return W + 1; }
template <int W>
size_t reg_shared(int tabs, int64_t tile) {
  return sizeof(double2) * tabs +
         std::max<size_t>(sizeof(double2) * (size_t)tile, sizeof(double) * (size_t)W * kRegT);
}

// Row maps of the rows kernel: row r of the transform reads src and writes
// dst at
//   kRowsFlat : r rs in both (in place when src == dst)
//   kRowsX2F  : X2 row x2_row(k, b) -> f row r = k nB + b (rows of nA)
//   kRowsFX2  : f row r -> X2 row x2_row(k, b)
enum RowsMap { kRowsFlat = 0, kRowsX2F = 1, kRowsFX2 = 2 };

// Rows of length n = 1024 (no Bluestein): one row per wavefront in the
// layouts of fft_reg.hpp
template <int DIR>
__global__ __launch_bounds__(64 * kRowsWaves) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_rows_reg(
    FftDesc fd, const double2 *src, double2 *dst, int64_t nrows, int64_t rs, double scale, Geo g, int map) {
  extern __shared__ double2 lds[];
  const int wv = threadIdx.x >> 6;
  double2 *tw = lds;
  double *T = (double *)(lds + kRegTw) + wv * kRegT;
  reg_stage_tw(tw, fd);
  __syncthreads();
  const int64_t groups = (nrows + kRowsWaves - 1) / kRowsWaves;
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    const int lane = loop_lane();
    const int64_t r = t * kRowsWaves + wv;
    if (r >= nrows) break;  // the whole wavefront, last group
    int64_t ri = r * rs, ro = r * rs;
    if (map != kRowsFlat) {
      const int k = (int)(r / g.nB), b = (int)(r % g.nB);
      const int64_t xr = x2_row(g, k, b) * g.ldA, fr = r * g.nA;
      ri = map == kRowsX2F ? xr : fr;
      ro = map == kRowsX2F ? fr : xr;
    }
    const double2 *in = src + ri;
    double2 *row = dst + ro;
    double2 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = in[lane + 64 * q];
    fft_F<DIR>(v, T, tw, tw + 1024, lane);
#pragma unroll
    for (int c = 0; c < 16; ++c) row[o_index(lane, c)] = cscale(v[c], scale);
  }
}

// Rows of a Bluestein length n in [257, 512] on the two-sequence register
// FFT (fft_reg32.hpp): wavefront w of group t takes rows 2 (t W + w) and the
// next one, one per half-wavefront, so each access instruction moves two
// runs of 32 consecutive entries.  Two wavefronts per SIMD (a lane holds 32
// complex values), one workgroup per CU sharing one copy of the tables.
constexpr int kRows32Waves = 8;

template <int DIR, int ODD>
__global__ __launch_bounds__(64 * kRows32Waves) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_rows_r32(
    FftDesc fd, const double2 *src, double2 *dst, int64_t nrows, int n, int64_t rs, double scale, Geo g,
    int map) {
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar index math
  double2 *tab = lds;
  double *T = (double *)(lds + kR32BlueTab) + wv * kR32T;
  r32_stage_blue<DIR>(tab, fd);
  __syncthreads();
  const int64_t groups = (nrows + 2 * kRows32Waves - 1) / (2 * kRows32Waves);
  // offsets of this lane's row of group t in src (IN = 1) or dst; a missing
  // second row of the last group repeats the first one (both halves then
  // write the same values)
  auto row_at = [&](int64_t t, int lane, bool IN) {
    const int64_t r0 = (t * kRows32Waves + wv) * 2;
    const int h = r0 + 1 < nrows ? lane >> 5 : 0;
    if (map == kRowsFlat) return (r0 + h) * rs;
    int k = (int)(r0 / g.nB), b = (int)(r0 % g.nB) + h;
    if (b == g.nB) b = 0, ++k;
    const bool x2 = (map == kRowsX2F) == IN;
    return x2 ? x2_row(g, k, b) * g.ldA : (r0 + h) * g.nA;
  };
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    if ((t * kRows32Waves + wv) * 2 >= nrows) break;  // the whole wavefront, last group
    const int lane = loop_lane(), l = lane & 31;
    const double2 *in = src + row_at(t, lane, true);
    double2 v[32];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = l + 32 * q;
      v[q] = in[q < 8 ? k : min(k, n - 1)];  // past n: cleared by the chirp
    }
    blue32<DIR, ODD>(v, n, tab, T, lane);
    double2 *row = dst + row_at(t, lane, false);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = l + 32 * q;
      if (q < 8 || k < n) row[k] = cscale(v[q], scale);  // n > 256
    }
  }
}

// Rows of a Bluestein length n in [513, 1024] on the 2048-point register FFT
// (fft_reg2k.hpp): one row per wavefront, each access instruction moving 64
// consecutive entries (even samples in the lower half-wavefront, odd ones in
// the upper); row maps as in k_rows_reg.  One workgroup per CU shares the
// tables; two wavefronts per SIMD.
template <int DIR>
__global__ __launch_bounds__(64 * kRows32Waves) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_rows_r2k(
    FftDesc fd, const double2 *src, double2 *dst, int64_t nrows, int n, int64_t rs, double scale, Geo g,
    int map) {
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double2 *tab = lds;
  double *T = (double *)(lds + kR2kTab) + wv * kR32T;
  r2k_stage_blue(tab, fd);
  __syncthreads();
  const double2 *bh = DIR < 0 ? fd.bhat_f : fd.bhat_b;
  const int64_t groups = (nrows + kRows32Waves - 1) / kRows32Waves;
  // rows of the mapped layouts number nG nB < 2^31 (32-bit arithmetic)
  auto row_at = [&](int64_t r, bool IN) -> int64_t {
    if (map == kRowsFlat) return r * rs;
    const int ri = (int)r, k = ri / g.nB, b = ri - k * g.nB;
    const bool x2 = (map == kRowsX2F) == IN;
    return x2 ? x2_row(g, k, b) * g.ldA : r * g.nA;
  };
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    const int64_t r = t * kRows32Waves + wv;
    if (r >= nrows) break;  // the whole wavefront, last group
    const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
    const double2 *in = src + row_at(r, true);
    double2 v[32];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = in[min(2 * (l + 32 * q) + h, n - 1)];  // past n: cleared by the chirp
    // outputs stored as the last DFT-32 makes them, addresses from a fresh
    // opaque lane index (the compiler would otherwise keep the 16 load
    // addresses live across the transform)
    const int ls = loop_lane();
    double2 *row = dst + row_at(r, false);
    const double sc = 0.5 * scale;
    blue2k_out<DIR>(v, n, tab, bh, T, lane, [&](int q, double2 z) {
      const int j = 2 * ((ls & 31) + 32 * q) + (ls >> 5);
      if (j < n) row[j] = cscale(z, sc);
    });
  }
}

// wavefronts per workgroup of the 4096-point kernels (eight measured faster
// than two, whose barriers would couple only the pair, and four)
#ifndef SO3_R4K_WAVES
#define SO3_R4K_WAVES 8
#endif
constexpr int kR4kWaves = SO3_R4K_WAVES;

// Rows of a Bluestein length n in [1025, 2048] on the 4096-point register
// FFT (fft_reg4k.hpp): one row per pair of wavefronts (even samples in
// wavefront 0, odd ones in wavefront 1), two rows per workgroup of four
// wavefronts, one wavefront per SIMD; row maps as in k_rows_reg.  With the
// whole register file per wavefront (256 VGPRs and up to 256 AGPRs), each
// lane keeps its chirp factors and kernel-spectrum entries (the same for
// every row) in registers instead of reloading them for every row, and the
// next row's samples are loaded before this row's transform: 7-16% faster
// than eight wavefronts at two per SIMD reloading the tables.  Every
// wavefront runs every group (the transform synchronises the workgroup); a
// pair past the last row transforms a repeat of that row and stores nothing.
constexpr int kRows4kWaves = 4;

template <int DIR>
__global__ __launch_bounds__(64 * kRows4kWaves) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_rows_r4k(
    FftDesc fd, const double2 *src, double2 *dst, int64_t nrows, int n, int64_t rs, double scale, Geo g,
    int map) {
  constexpr int kPairs = kRows4kWaves / 2;
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), e = wv & 1;
  double2 *tab = lds;
  double *T = (double *)(lds + kR4kTab) + wv * kR32T;
  const double *Tp = (double *)(lds + kR4kTab) + (wv ^ 1) * kR32T;
  r4k_stage_blue(tab, fd);
  double2 cpre[16], spre[32];  // chirp of samples q < 16, kernel spectrum of outputs q < 32
  {
    const int lane = threadIdx.x & 63, l = lane & 31, h = lane >> 5;
    const double2 *__restrict__ bh = DIR < 0 ? fd.bhat_f : fd.bhat_b;
#pragma unroll
    for (int q = 0; q < 16; ++q) cpre[q] = r4k_chirp<DIR>(fd.chirp, n, 4 * (l + 32 * q) + 2 * h + e);
#pragma unroll
    for (int q = 0; q < 32; ++q) spre[q] = bh[l + 32 * q + 1024 * h + 2048 * e];
  }
  __syncthreads();
  const int64_t groups = (nrows + kPairs - 1) / kPairs;
  // rows of the mapped layouts number nG nB < 2^31 (32-bit arithmetic)
  auto row_at = [&](int64_t r, bool IN) -> int64_t {
    if (map == kRowsFlat) return r * rs;
    const int ri = (int)r, k = ri / g.nB, b = ri - k * g.nB;
    const bool x2 = (map == kRowsX2F) == IN;
    return x2 ? x2_row(g, k, b) * g.ldA : r * g.nA;
  };
  auto load_row = [&](int64_t t, double2 *x) {  // samples of group t, registers q < 16
    const int64_t r = t * kPairs + (wv >> 1);
    const int64_t rr = r < nrows ? r : nrows - 1;
    const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
    const double2 *in = src + row_at(rr, true);
#pragma unroll
    for (int q = 0; q < 16; ++q) x[q] = in[min(4 * (l + 32 * q) + 2 * h + e, n - 1)];  // past n: cleared by the chirp
  };
  double2 nx[16];
  int64_t t = xcd_tile(blockIdx.x, gridDim.x);
  if (t < groups) load_row(t, nx);
  for (; t < groups; t += gridDim.x) {
    const int64_t r = t * kPairs + (wv >> 1);
    const bool ok = r < nrows;  // a pair past the last row repeats it, unstored
    const int64_t rr = ok ? r : nrows - 1;
    const int lane = loop_lane();
    double2 v[32];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = nx[q];
    if (t + gridDim.x < groups) load_row(t + gridDim.x, nx);
    // outputs stored as the last DFT-32 makes them, addresses from a fresh
    // opaque lane index (see k_rows_r2k)
    const int ls = loop_lane();
    double2 *row = dst + row_at(rr, false);
    blue4k_pre_out<DIR>(v, tab, cpre, spre, T, Tp, e, lane, [&](int q, double2 z) {
      const int j = 4 * ((ls & 31) + 32 * q) + 2 * (ls >> 5) + e;
      if (ok && j < n) row[j] = cscale(z, scale);
    });
  }
}

// Gamma between X2 and f of a complex signal whose alpha stage does not run
// on the register FFT (the other cases: k_gamma_r32, k_gamma_real_r32),
// Bluestein length 2N-1 with inner length 1024: a group is
// kRegWaves consecutive columns p = b nA + a (one per wavefront), moved
// between HBM and the registers through an LDS tile of nG x kRegWaves
// entries, so that one access instruction covers 16 n slots x 4 columns: 16
// runs of 64 bytes (the neighbouring group on the same XCD, xcd_tile, uses
// the other half of each 128-byte line) instead of 64 scattered entries.
// The tile is unpadded; entry (k, c) sits at gtile(k, c), whose swizzle
// keeps both the row-wise and the column-wise accesses free of bank
// conflicts.  It shares LDS with the transpose buffers (barriers separate
// the uses).
__device__ __forceinline__ int gtile(int k, int c) { return 4 * k + (c ^ ((k >> 2) & 3)); }

// Gamma tile kernel modes: kGxF inverse, X2 -> f; kFxG forward, f -> X2
enum GammaMode { kGxF = 0, kFxG = 1 };

template <int MODE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gamma_tile(
    Geo g, FftDesc fd, const void *src, void *dst, double scale) {
  static_assert(kRegWaves == 4, "gamma tile of 4 columns");
  constexpr int DIR = MODE == kGxF ? +1 : -1;
  constexpr int kPF = 8;  // tile entries per thread: nG * 4 <= 2048
  extern __shared__ double2 lds[];
  const int wv = threadIdx.x >> 6;
  double2 *tab = lds, *tile = lds + kRegBlueTab;
  double *T = (double *)(lds + kRegBlueTab) + wv * kRegT;
  reg_stage_blue<DIR>(tab, fd);
  const int S = g.nB * g.nA;  // columns; f has S entries per g
  const int64_t groups = (S + kRegWaves - 1) / kRegWaves;
  const int nG = 2 * g.N - 1;
  const double2 *X2s = (const double2 *)src;
  double2 *X2d = (double2 *)dst;
  // A thread always handles column c = threadIdx.x % 4 of a group and the n
  // slots k = threadIdx.x / 4 + 64 i.  Column p of group t: X2 entry of slot
  // k at x0 + k xs (x2_row), f entry at p + k S.
  const int c = threadIdx.x & 3, k0 = threadIdx.x >> 2;
  auto column = [&](int64_t t, int &p, int64_t &x0, int64_t &xs) {
    p = (int)(t * kRegWaves) + c;
    const int b = p / g.nA, a = p - b * g.nA;
    x0 = p < S ? x2_row(g, 0, b) * g.ldA + a : 0;
    xs = (int64_t)x2_w(g, b < g.nB ? b : 0) * g.ldA;
  };
  auto fetch = [&](int64_t t, double2 (&P)[kPF]) {
    int p;
    int64_t x0, xs;
    column(t, p, x0, xs);
    const bool ok = t < groups && p < S;
#pragma unroll
    for (int i = 0; i < kPF; ++i) {
      const int k = k0 + 64 * i;
      double2 x = make_double2(0.0, 0.0);
      if (ok && k < nG) {
        if constexpr (MODE == kFxG) x = ((const double2 *)src)[(int64_t)k * S + p];
        else x = X2s[x0 + k * xs];
      }
      P[i] = x;
    }
  };
  double2 P[kPF];
  int64_t t = xcd_tile(blockIdx.x, gridDim.x);
  fetch(t, P);
  __syncthreads();  // tables staged
  for (; t < groups; t += gridDim.x) {
    const int lane = loop_lane();
#pragma unroll
    for (int i = 0; i < kPF; ++i) {
      const int k = k0 + 64 * i;
      if (k < nG) tile[gtile(k, c)] = P[i];
    }
    __syncthreads();
    double2 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = lane + 64 * q;
      v[q] = k < nG ? tile[gtile(k, wv)] : make_double2(0.0, 0.0);
    }
    __syncthreads();  // the tile is read before the transposes reuse it
    fetch(t + gridDim.x, P);  // the next group's loads run during the DFT
    blue_reg<DIR>(v, nG, tab, T, lane);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {  // nG <= 512
      const int k = lane + 64 * q;
      if (k < nG) tile[gtile(k, wv)] = v[q];
    }
    __syncthreads();
    {
      int p;
      int64_t x0, xs;
      column(t, p, x0, xs);
      if (p < S) {
#pragma unroll
        for (int i = 0; i < kPF; ++i) {
          const int k = k0 + 64 * i;
          if (k >= nG) continue;
          const double2 x = tile[gtile(k, c)];
          if constexpr (MODE == kGxF) ((double2 *)dst)[(int64_t)k * S + p] = cscale(x, scale);
          else X2d[x0 + k * xs] = cscale(x, scale);
        }
      }
    }
    __syncthreads();  // the tile is read out before the next group's entries
  }
}

// Gamma in place on X2 (complex) on the two-sequence register FFT: the
// columns p = b nA + a are numbered as in k_gamma_tile; wavefront w of group
// t takes columns 2 (t W + u) and the next one, u = quad(w) (a permutation
// of the wavefronts, set out at the group barrier), one per half-wavefront,
// and loads and stores their nG n slots straight from X2.  Each half of the
// workgroup starts together on W consecutive columns, whose entries of one
// slot share 128-byte lines, so most of each line is used from L1 instead of
// being staged through an LDS tile (no room beside eight transpose buffers).
template <int DIR>
__global__ __launch_bounds__(64 * kRows32Waves) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gamma_r32(
    Geo g, FftDesc fd, double2 *X2, double scale) {
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double2 *tab = lds;
  double *T = (double *)(lds + kR32BlueTab) + wv * kR32T;
  r32_stage_blue<DIR>(tab, fd);
  __syncthreads();
  const int S = g.nB * g.nA, nG = g.nG;
  const int64_t groups = (S + 2 * kRows32Waves - 1) / (2 * kRows32Waves);
  // One barrier per group keeps the wavefronts of a workgroup within part of
  // a group of each other.  Even wavefronts pass it before their loads, odd
  // ones after them, so the two wavefronts of a SIMD are apart by the time
  // the loads take (6% faster than passing it together).  The even
  // wavefronts take the first half of the group's columns and the odd ones
  // the second, so that the users of a 128-byte line pass the barrier
  // together and find the line in L2.  Every wavefront passes it once per
  // group.
  const bool late = (wv & 1) != 0;
  const int quad = (wv & 1) * (kRows32Waves / 2) + (wv >> 1);
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    if (!late) __syncthreads();
    const int p0 = (int)(t * kRows32Waves + quad) * 2;
    if (p0 >= S) {  // the whole wavefront, last group
      if (late) __syncthreads();
      break;
    }
    const int lane = loop_lane(), l = lane & 31;
    // a missing second column of the last group repeats the first one (both
    // halves then write the same values)
    const int p = p0 + (p0 + 1 < S ? lane >> 5 : 0);
    const int b = p / g.nA, a = p - b * g.nA;
    double2 *col = X2 + x2_row(g, 0, b) * g.ldA + a;
    const int64_t xs = (int64_t)x2_w(g, b) * g.ldA;
    double2 v[32];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = l + 32 * q;
      v[q] = col[(q < 8 ? k : min(k, nG - 1)) * xs];  // past nG: cleared by the chirp
    }
    if (late) __syncthreads();
    blue32<DIR, 1>(v, nG, tab, T, lane);  // nG = 2N - 1 is odd
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = l + 32 * q;
      if (q < 8 || k < nG) col[k * xs] = cscale(v[q], scale);  // nG > 256
    }
  }
}

// ---------------------------------------------------------------- gamma, 256-point register FFT
// Gamma of a complex signal between X2 and f for the Bluestein lengths
// nG = 2N - 1 in [33, 127] on the 256-point register FFT (fft_reg16.hpp):
// wavefront w of group t takes the four columns p = 4 (t W + w) + h, one
// per sequence h, and reads and writes their nG slots straight from X2 and
// f.  An access instruction then moves 16 runs of 64 bytes (four
// consecutive columns of 16 rows); the neighbouring wavefront uses the other
// half of each 128-byte line, and one workgroup barrier per group keeps the
// wavefronts of a workgroup in step so that the line is still in L2 for it.
// DIR = +1: inverse (X2 -> f), -1: forward (f -> X2).
constexpr int kR16Waves = 4;  // one wavefront per SIMD and workgroup, three workgroups per CU

template <int DIR>
__global__ __launch_bounds__(64 * kR16Waves) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_gamma_r16(
    Geo g, FftDesc fd, const double2 *__restrict__ src, double2 *__restrict__ dst, double scale) {
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double2 *tab = lds;
  double *T = (double *)(lds + kR16Tab) + wv * kR16T;
  r16_stage_blue<DIR>(tab, fd);
  __syncthreads();  // tables staged
  const int nG = 2 * g.N - 1;
  const int S = g.nB * g.nA;
  const int groups = (S + 4 * kR16Waves - 1) / (4 * kR16Waves);
  for (int t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    __syncthreads();  // every wavefront passes once per group
    const int lane = loop_lane(), l = lane & 15, h = lane >> 4;
    const int p = 4 * (t * kR16Waves + wv) + h;
    const bool ok = p < S;
    const int pc = ok ? p : S - 1;
    const int b = pc / g.nA, a = pc - b * g.nA;
    const int64_t xb = x2_row(g, 0, b) * g.ldA + a, xs = (int64_t)x2_w(g, b) * g.ldA;
    double2 v[16];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int j = min(l + 16 * q, nG - 1);  // entries j >= nG meet a zero chirp factor
      v[q] = DIR > 0 ? src[xb + j * xs] : src[(int64_t)j * S + pc];
    }
    double2 *const o = DIR > 0 ? dst + pc : dst + xb;
    const int64_t os = DIR > 0 ? (int64_t)S : xs;
    blue16(v, tab, T, lane, [&](int q, double2 z) {
      const int k = l + 16 * q;
      if (ok && k < nG) o[k * os] = cscale(z, scale);
    });
  }
}

// Gamma of a real signal between X2 (n slots 0..N-1) and f (nf = 2N - 1 real
// samples per column, nf in [33, 127]) on the 256-point register FFT:
// sequence h of wavefront w in group t takes the columns
// pA = 8 (t W + w) + 2 h and pB = pA + 1 as one complex DFT (the algebra of
// k_gamma_real_r32):
//   inverse (c2r): the Hermitian extension of X_A + i X_B (slot 0 real,
//     slots k >= N from conj(slot nf - k)); f_A, f_B are the real and
//     imaginary parts of the DFT;
//   forward (r2c): the DFT Y of f_A + i f_B separates into
//     X_A = (Y(k) + conj Y(-k)) / 2, X_B = (Y(k) - conj Y(-k)) / (2 i), k < N,
//     Y(-k) through the wavefront's transpose buffer.
template <int DIR>
__global__ __launch_bounds__(64 * kR16Waves) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_gamma_real_r16(
    Geo g, FftDesc fd, double2 *X2, double *f, double scale) {
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double2 *tab = lds;
  double *T = (double *)(lds + kR16Tab) + wv * kR16T;
  double2 *T2 = (double2 *)T;
  r16_stage_blue<DIR>(tab, fd);
  __syncthreads();  // tables staged
  const int S = g.nB * g.nA, N = g.N, nf = 2 * g.N - 1;
  const int groups = (S + 8 * kR16Waves - 1) / (8 * kR16Waves);
  for (int t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    __syncthreads();  // every wavefront passes once per group
    const int lane = loop_lane(), l = lane & 15, h = lane >> 4;
    const int pA = 8 * (t * kR16Waves + wv) + 2 * h, pB = pA + 1;
    const bool okA = pA < S, okB = pB < S;  // a column past the last one repeats it, unstored
    const int qA = okA ? pA : S - 1, qB = okB ? pB : S - 1;
    const int bA = qA / g.nA, aA = qA - bA * g.nA, bB = qB / g.nA, aB = qB - bB * g.nA;
    double2 *cA = X2 + x2_row(g, 0, bA) * g.ldA + aA, *cB = X2 + x2_row(g, 0, bB) * g.ldA + aB;
    const int64_t sA = (int64_t)x2_w(g, bA) * g.ldA, sB = (int64_t)x2_w(g, bB) * g.ldA;
    double2 v[16];
    if (DIR > 0) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k = min(l + 16 * q, nf - 1);  // past nf: cleared by the chirp
        const bool dir = k < N;
        const int kk = dir ? k : nf - k;
        double2 a = cA[kk * sA], b = cB[kk * sB];
        if (!dir) a.y = -a.y, b.y = -b.y;
        if (k == 0) a.y = 0.0, b.y = 0.0;  // slot 0 of a real signal is real
        v[q] = make_double2(a.x - b.y, a.y + b.x);
      }
      blue16(v, tab, T, lane, [&](int q, double2 z) {
        const int k = l + 16 * q;
        if (k >= nf) return;
        if (okA) f[(int64_t)k * S + pA] = z.x * scale;
        if (okB) f[(int64_t)k * S + pB] = z.y * scale;
      });
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k = min(l + 16 * q, nf - 1);
        v[q] = make_double2(f[(int64_t)k * S + qA], f[(int64_t)k * S + qB]);
      }
      blue16(v, tab, T, lane, [&](int q, double2 z) { v[q] = z; });
      // Y(-k) = Y(nf - k) (Y(0) for k = 0) through the buffer: entry c of
      // the sequence sits at (c >> 4) 64 + 16 h + (c & 15)
#pragma unroll
      for (int q = 0; q < 8; ++q) T2[q * 64 + lane] = v[q];
      __builtin_amdgcn_wave_barrier();
      r32_fence();
      const double hs = 0.5 * scale;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = l + 16 * j;
        const int c = k == 0 ? 0 : max(nf - k, 0);
        const double2 y = T2[(c >> 4) * 64 + 16 * h + (c & 15)];
        if (k >= N) continue;
        // X_A = (Y + conj Y(-k)) / 2, X_B = (Y - conj Y(-k)) / (2 i)
        const double2 xa = make_double2(hs * (v[j].x + y.x), hs * (v[j].y - y.y));
        const double2 xb = make_double2(hs * (v[j].y + y.y), -hs * (v[j].x - y.x));
        if (okA) cA[k * sA] = xa;
        if (okB) cB[k * sB] = xb;
      }
      __builtin_amdgcn_wave_barrier();
      r32_fence();  // the buffer is read before the next group's transposes
    }
  }
}

// ---------------------------------------------------------------- inverse gamma
__global__ __launch_bounds__(kThreads) void k_inv_gamma(Geo g, FftDesc fd, int K, int ld,
                                                        const double2 *__restrict__ X2,
                                                        double2 *__restrict__ fc,
                                                        double *__restrict__ fr, double scale) {
  extern __shared__ double2 lds[];
  double2 *A = lds;
  const double2 *tw = fft_load_twiddles(lds + K * ld, fd);
  // columns p = b nA + a of f; column (b, a) of X2 holds n slot k at
  // x2_row(k, b) ldA + a.  K divides the workgroup: thread t always handles
  // column s = t % K (one division per thread) and slots k = t / K + i (256 / K)
  const int64_t S = (int64_t)g.nB * g.nA;
  const int nG = 2 * g.N - 1;
  const int s = threadIdx.x % K, k0 = threadIdx.x / K, ks = blockDim.x / K;
  const int64_t p = (int64_t)blockIdx.x * K + s;
  const bool okp = p < S;
  const int b = okp ? (int)p / g.nA : 0, a = okp ? (int)p - b * g.nA : 0;
  const int64_t xb = x2_row(g, 0, b) * g.ldA + a, xs = (int64_t)x2_w(g, b) * g.ldA;
  for (int k = k0; k < nG; k += ks) {
    double2 v = make_double2(0.0, 0.0);
    if (okp) {
      if (!g.real) {
        v = X2[xb + k * xs];
      } else if (k == 0) {
        v = make_double2(X2[xb].x, 0.0);
      } else if (k < g.N) {
        v = X2[xb + k * xs];
      } else {
        v = cconj(X2[xb + (nG - k) * xs]);
      }
    }
    A[s * ld + k] = v;
  }
  __syncthreads();
  fft_run<+1>(A, K, ld, fd, tw);
  if (!okp) return;
  for (int k = k0; k < nG; k += ks) {
    if (g.real)
      fr[k * S + p] = A[s * ld + k].x * scale;
    else
      fc[k * S + p] = cscale(A[s * ld + k], scale);
  }
}

// ---------------------------------------------------------------- forward gamma
__global__ __launch_bounds__(kThreads) void k_fwd_gamma(Geo g, FftDesc fd, int K, int ld,
                                                        const double2 *__restrict__ fc,
                                                        const double *__restrict__ fr,
                                                        double2 *__restrict__ X2, double sc) {
  extern __shared__ double2 lds[];
  double2 *A = lds;
  const double2 *tw = fft_load_twiddles(lds + K * ld, fd);
  // columns p = b nA + a of f map to columns (b, a) of X2; K divides the
  // workgroup: thread t always handles column s = t % K, slots k = t / K + i (256 / K)
  const int64_t S = (int64_t)g.nB * g.nA;
  const int nG = 2 * g.N - 1;
  const int s = threadIdx.x % K, k0 = threadIdx.x / K, ks = blockDim.x / K;
  const int64_t p = (int64_t)blockIdx.x * K + s;
  const bool okp = p < S;
  for (int k = k0; k < nG; k += ks) {
    double2 v = make_double2(0.0, 0.0);
    if (okp) v = g.real ? make_double2(fr[k * S + p], 0.0) : fc[k * S + p];
    A[s * ld + k] = v;
  }
  __syncthreads();
  fft_run<-1>(A, K, ld, fd, tw);
  if (!okp) return;
  const int b = (int)p / g.nA, a = (int)p - b * g.nA;
  const int64_t xb = x2_row(g, 0, b) * g.ldA + a, xs = (int64_t)x2_w(g, b) * g.ldA;
  const int nout = g.real ? g.N : nG;
  for (int k = k0; k < nout; k += ks) X2[xb + k * xs] = cscale(A[s * ld + k], sc);
}

// ---------------------------------------------------------------- steerable gamma
// Steerable signals keep N gamma samples at pi g / N (so3_core.c:115-130,
// 219-222 inverse; :309-328 forward).
//   inverse: the n slots, n < 0 moved to n + 2N and entry N zero, go through
//     a backward DFT of length 2N; samples g < N are kept;
//   forward: the N samples, zero-extended to 2N, go through a forward DFT;
//     n of the parity of N - 1 take sc times entry n mod 2N, the other n
//     are zero (the reference computes only those n).
// A workgroup owns K columns p = b nA + a, as in k_inv_gamma / k_fwd_gamma.
__global__ __launch_bounds__(kThreads) void k_inv_gamma_steer(Geo g, FftDesc fd, int K, int ld,
                                                              const double2 *__restrict__ X2,
                                                              double2 *__restrict__ f, double scale) {
  extern __shared__ double2 lds[];
  double2 *A = lds;
  const double2 *tw = fft_load_twiddles(lds + K * ld, fd);
  const int64_t S = (int64_t)g.nB * g.nA;
  const int64_t p0 = (int64_t)blockIdx.x * K;
  const int N = g.N, M = 2 * g.N;
  for (int idx = threadIdx.x; idx < K * M; idx += blockDim.x) {
    const int j = idx / K, s = idx - j * K;
    const int64_t p = p0 + s;
    double2 v = make_double2(0.0, 0.0);
    if (p < S && j != N) {
      const int k = j < N ? j : j - 1;  // n = j - 2N sits at slot n + 2N - 1
      const int b = (int)(p / g.nA), a = (int)(p % g.nA);
      v = X2[x2_row(g, k, b) * g.ldA + a];
    }
    A[s * ld + j] = v;
  }
  __syncthreads();
  fft_run<+1>(A, K, ld, fd, tw);
  for (int idx = threadIdx.x; idx < K * N; idx += blockDim.x) {
    const int k = idx / K, s = idx - k * K;
    const int64_t p = p0 + s;
    if (p < S) f[k * S + p] = cscale(A[s * ld + k], scale);
  }
}

__global__ __launch_bounds__(kThreads) void k_fwd_gamma_steer(Geo g, FftDesc fd, int K, int ld,
                                                              const double2 *__restrict__ f,
                                                              double2 *__restrict__ X2, double sc) {
  extern __shared__ double2 lds[];
  double2 *A = lds;
  const double2 *tw = fft_load_twiddles(lds + K * ld, fd);
  const int64_t S = (int64_t)g.nB * g.nA;
  const int64_t p0 = (int64_t)blockIdx.x * K;
  const int N = g.N, M = 2 * g.N, nG = g.nG;
  for (int idx = threadIdx.x; idx < K * M; idx += blockDim.x) {
    const int j = idx / K, s = idx - j * K;
    const int64_t p = p0 + s;
    A[s * ld + j] = (p < S && j < N) ? f[j * S + p] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  fft_run<-1>(A, K, ld, fd, tw);
  for (int idx = threadIdx.x; idx < K * nG; idx += blockDim.x) {
    const int k = idx / K, s = idx - k * K;
    const int64_t p = p0 + s;
    if (p >= S) continue;
    const int n = k < N ? k : k - nG;
    const double2 v = ((n + N - 1) & 1) ? make_double2(0.0, 0.0) : cscale(A[s * ld + (n < 0 ? n + M : n)], sc);
    X2[x2_row(g, k, (int)(p / g.nA)) * g.ldA + p % g.nA] = v;
  }
}

// ---------------------------------------------------------------- forward beta
__global__ __launch_bounds__(kThreads) void k_fwd_beta(Geo g, FftDesc fb, FftDesc fw, int K, int ld,
                                                       const double2 *__restrict__ X2,
                                                       double2 *__restrict__ Gp,
                                                       const int64_t *__restrict__ nbase,
                                                       const double2 *__restrict__ phase,
                                                       const double2 *__restrict__ wr, int Pw,
                                                       int adjoint) {
  extern __shared__ double2 lds[];
  // one twiddle table serves both transforms when their lengths agree
  double2 *A = lds, *twb = lds + K * ld, *tww = fw.len == fb.len ? twb : twb + tw_count(fb);
  const int chunks = (g.nA + K - 1) / K;
  const int nslot = blockIdx.x / chunks;
  const int ms0 = (blockIdx.x % chunks) * K;
  const int n = n_of_slot(g, nslot);
  if (nbase[n + g.N - 1] == kNoBlock) return;  // block not transformed
  const double2 *twbp = fft_load_twiddles(twb, fb);
  const double2 *twwp = fw.len == fb.len ? twbp : fft_load_twiddles(tww, fw);
  const int Pb = g.Pb, L = g.L;
  for (int idx = threadIdx.x; idx < K * Pb; idx += blockDim.x) {
    const int k = idx / K, s = idx - k * K;
    const int mslot = ms0 + s;
    double2 v = make_double2(0.0, 0.0);
    if (mslot < g.nA) {
      const int m = m_of_slot(g, mslot);
      int b = k;
      bool neg = false, zero = false;
      if (!g.ss) {
        if (b >= L) { b = 2 * L - 2 - b; neg = ((m + n) & 1); zero = adjoint; }
      } else {
        if (b > L) { b = 2 * L - b; neg = ((m + n) & 1); zero = adjoint; }
      }
      v = zero ? make_double2(0.0, 0.0) : X2[x2_row(g, nslot, b) * g.ldA + mslot];
      if (neg) v = make_double2(-v.x, -v.y);
    }
    A[s * ld + k] = v;
  }
  __syncthreads();
  fft_run<-1>(A, K, ld, fb, twbp);
  // F~(m'') = DFT/Pb * phase, placed at m'' mod Pw (zero elsewhere); the
  // remap runs through registers because it widens the sequence in place
  {
    constexpr int kMaxPer = kFftElems * 2;
    double2 v[kMaxPer];
#pragma unroll
    for (int b = 0; b < kMaxPer; ++b) {
      const int idx = threadIdx.x + b * kThreads;
      v[b] = make_double2(0.0, 0.0);
      if (idx < K * Pw) {
        const int s = idx / Pw, r = idx - s * Pw;
        const int mpp = r <= L - 1 ? r : r - Pw;
        if (mpp >= -(L - 1)) v[b] = cmul(A[s * ld + (mpp < 0 ? mpp + Pb : mpp)], phase[mpp + L - 1]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kMaxPer; ++b) {
      const int idx = threadIdx.x + b * kThreads;
      if (idx < K * Pw) {
        const int s = idx / Pw, r = idx - s * Pw;
        A[s * ld + r] = v[b];
      }
    }
  }
  __syncthreads();
  double sc = 1.0;
  if (!adjoint) {
    fft_run<+1>(A, K, ld, fw, twwp);
    for (int idx = threadIdx.x; idx < K * Pw; idx += blockDim.x) {
      const int s = idx / Pw, r = idx - s * Pw;
      A[s * ld + r] = cmul(A[s * ld + r], wr[r]);
    }
    __syncthreads();
    fft_run<-1>(A, K, ld, fw, twwp);
    sc = kTwoPi / (double)Pw;
  }
  const double2 *R3 = A;
  const int C = 2 * g.NS;
  const int an = n < 0 ? -n : n;
  for (int idx = threadIdx.x; idx < K * L; idx += blockDim.x) {
    const int s = idx / L, mp = idx - s * L;
    const int mslot = ms0 + s;
    if (mslot >= g.nA) continue;
    const int m = m_of_slot(g, mslot);
    if (m < -(L - 1)) continue;  // MW_SS Nyquist column
    double2 G = cscale(R3[s * ld + mp], sc);
    if (mp > 0) {
      double2 Gn = cscale(R3[s * ld + (Pw - mp)], sc);
      if ((m + n) & 1) G = csub(G, Gn);
      else G = cadd(G, Gn);
    }
    const int am = m < 0 ? -m : m;
    const int sm = m < 0, sn = n < 0;
    double2 *dst = Gp + ((int64_t)an * L + am) * C * L + mp;
    // m = 0 or n = 0 serve both sign combinations
    for (int cm = 0; cm < 2; ++cm) {
      if (cm != sm && am != 0) continue;
      for (int cn = 0; cn < g.NS; ++cn) {
        if (cn != sn && an != 0) continue;
        // mixed-sign pairs carry D^l_{m',-|m|} = (-1)^{l+m'} D^l_{m',|m|}: the
        // (-1)^{m'} part is folded here, (-1)^l in the contraction
        dst[(int64_t)(cm + 2 * cn) * L] = ((cm != cn) && (mp & 1)) ? make_double2(-G.x, -G.y) : G;
      }
    }
  }
}

// forward beta, register path (Pb Bluestein with inner length 1024,
// Pw = 1024, so L <= 256; MW and MW_SS): sequence (n slot, m slot) per
// wavefront, W consecutive m slots of one n slot per group, the beta samples
// coming in through the LDS tile.
template <int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_fwd_beta_reg(
    Geo g, FftDesc fb, FftDesc fw, const double2 *__restrict__ X2, double2 *__restrict__ Gp,
    const int64_t *__restrict__ nbase, const double2 *__restrict__ phase, const double2 *__restrict__ wr,
    int adjoint) {
  extern __shared__ double2 lds[];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // one twiddle table serves the Bluestein and the convolution transforms
  // (both of length 1024)
  double2 *tw = lds, *tile = lds + kRegBlueTab;
  double *T = (double *)(lds + kRegBlueTab) + wv * kRegT;
  reg_stage_blue<-1>(tw, fb);
  __syncthreads();
  const int chunks = (g.nA + W - 1) / W;
  const int64_t groups = (int64_t)g.nG * chunks;
  const int Pb = g.Pb, L = g.L;
  const int C = 2 * g.NS;
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    const int lane = loop_lane();
    const int nslot = (int)(t / chunks);
    const int ms0 = (int)(t % chunks) * W;
    const int n = n_of_slot(g, nslot);
    if (nbase[n + g.N - 1] == kNoBlock) continue;  // block not transformed (whole workgroup)
    // beta samples of the W sequences: tile[b][s]
    for (int idx = threadIdx.x; idx < g.nB * W; idx += (64 * W)) {
      const int b = idx / W, c = idx - b * W;
      tile[b * tile_ld<W>() + c] =
          ms0 + c < g.nA ? X2[x2_row(g, nslot, b) * g.ldA + ms0 + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int mslot = ms0 + wv;
    const int m = m_of_slot(g, mslot);
    const bool flip = ((m + n) & 1) != 0;
    double2 v[16];
    // extension F(2L-2-b) = (-1)^{m+n} F(b) (zero for the adjoint)
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = lane + 64 * q;
      double2 x = make_double2(0.0, 0.0);
      if (k < Pb) {
        // MW: F(2L-2-b) for b >= L; MW_SS: F(2L-b) for b > L
        const int kr = g.ss ? 2 * L - k : 2 * L - 2 - k;
        if (g.ss ? k <= L : k < L) {
          x = tile[k * tile_ld<W>() + wv];
        } else if (!adjoint) {
          x = tile[kr * tile_ld<W>() + wv];
          if (flip) x = make_double2(-x.x, -x.y);
        }
      }
      v[q] = x;
    }
    __syncthreads();  // the tile is read before the transposes reuse it
    blue_reg<-1>(v, Pb, tw, T, lane);
    // F~(m'') times the phase (which carries 1/Pb) at m'' mod Pw, |m''| < L:
    // a gather through the transpose buffer, real parts then imaginary parts
    {
      double t[16];
#pragma unroll
      for (int q = 0; q < 8; ++q) T[lane + 64 * q] = v[q].x;
      __builtin_amdgcn_wave_barrier();
      phase_fence();
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int r = lane + 64 * q;
        const int mpp = r <= L - 1 ? r : r - 1024;
        t[q] = mpp >= -(L - 1) ? T[mpp < 0 ? mpp + Pb : mpp] : 0.0;
      }
      __builtin_amdgcn_wave_barrier();
      phase_fence();
#pragma unroll
      for (int q = 0; q < 8; ++q) T[lane + 64 * q] = v[q].y;
      __builtin_amdgcn_wave_barrier();
      phase_fence();
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int r = lane + 64 * q;
        const int mpp = r <= L - 1 ? r : r - 1024;
        double2 x = make_double2(0.0, 0.0);
        if (mpp >= -(L - 1)) x = cmul(make_double2(t[q], T[mpp < 0 ? mpp + Pb : mpp]), phase[mpp + L - 1]);
        v[q] = x;
      }
      __builtin_amdgcn_wave_barrier();
      phase_fence();
    }
    double sc = 1.0;
    if (!adjoint) {
      // quadrature-weight convolution as a product in the conjugate domain
      fft_F<+1>(v, T, tw, tw + 1024, lane);
#pragma unroll
      for (int c = 0; c < 16; ++c) v[c] = cmul(v[c], wr[o_index(lane, c)]);
      fft_G<-1>(v, T, tw, tw + 1024, lane);
      sc = kTwoPi / 1024.0;
    }
    // fold to m' >= 0: G'(m') = G(m') + (-1)^{m+n} G(-m'), m' = lane + 64 j
    constexpr int kFold = 4;  // L <= 256
    double g0r[kFold], gnr[kFold];
#pragma unroll
    for (int q = 0; q < 16; ++q) T[lane + 64 * q] = v[q].x;
    __builtin_amdgcn_wave_barrier();
    phase_fence();
#pragma unroll
    for (int j = 0; j < kFold; ++j) {
      const int mp = lane + 64 * j;
      g0r[j] = T[mp];
      gnr[j] = T[(1024 - mp) & 1023];
    }
    __builtin_amdgcn_wave_barrier();
    phase_fence();
#pragma unroll
    for (int q = 0; q < 16; ++q) T[lane + 64 * q] = v[q].y;
    __builtin_amdgcn_wave_barrier();
    phase_fence();
    const bool store = mslot < g.nA && m >= -(L - 1);  // not the MW_SS Nyquist column / tail
    const int an = n < 0 ? -n : n, am = m < 0 ? -m : m;
    const int sm = m < 0, sn = n < 0;
    double2 *dst = Gp + ((int64_t)an * L + am) * C * L;
#pragma unroll
    for (int j = 0; j < kFold; ++j) {
      const int mp = lane + 64 * j;
      double2 G0 = make_double2(g0r[j] * sc, T[mp] * sc);
      if (mp > 0) {
        const double2 Gn = make_double2(gnr[j] * sc, T[1024 - mp] * sc);
        G0 = flip ? csub(G0, Gn) : cadd(G0, Gn);
      }
      if (!store || mp >= L) continue;
      for (int cm = 0; cm < 2; ++cm) {
        if (cm != sm && am != 0) continue;
        for (int cn = 0; cn < g.NS; ++cn) {
          if (cn != sn && an != 0) continue;
          dst[(int64_t)(cm + 2 * cn) * L + mp] = ((cm != cn) && (mp & 1)) ? make_double2(-G0.x, -G0.y) : G0;
        }
      }
    }
    __syncthreads();  // transposes done before the next group's tile
  }
}

// ---------------------------------------------------------------- paired beta, register path
// A beta sequence (n, m) is symmetric under m' -> -m' with sign
// s = (-1)^{m+n}: X1(-m') = s X1(m') on the way in, and the extended samples
// obey ext(b') = s ext(b), b' = Pb - 1 - b (MW) or Pb - b (MW_SS), on the
// way out.  Its DFT then has the same symmetry: Z(b') = s Z(b) (inverse),
// F(-m'') = s F(m'') (forward).  Two sequences A, B of opposite signs share
// one DFT of A + i B and separate afterwards:
//   A = (Z(b) + s_A Z(b')) / 2,   B = (Z(b) - s_A Z(b')) / (2 i).
// Consecutive m have opposite signs, so a sequence of the register FFT
// takes the pair (m, m + 1); groups of consecutive m start at
// m_lo = -(L-1) - ss (the MW_SS Nyquist column, zero, first).
__device__ __forceinline__ int m_slot_of(const Geo &g, int m) { return m >= 0 ? m : m + g.nA; }

// Gamma of a real signal between X2 (n slots 0..N-1) and f (2N-1 real
// samples per column) on the two-sequence register FFT.  The samples are
// real, so one complex DFT serves two columns A, B: half-wavefront h of
// wavefront w takes columns p = 4 (t W + w) + 2 h and p + 1.
//   inverse (c2r): the Hermitian extension of A + i B (slot 0 real, slots
//     k >= N from conj(slot 2N-1-k)); f_A, f_B are the real and imaginary
//     parts of the DFT;
//   forward (r2c): the DFT Y of f_A + i f_B separates into
//     X_A = (Y(k) + conj Y(-k)) / 2, X_B = (Y(k) - conj Y(-k)) / (2 i),
//     Y(-k) through the wavefront's transpose buffer.
// The group barrier keeps the wavefronts that share lines in step.  (The
// staggered barrier of k_gamma_r32 measured slower here: the forward
// transform lost what the inverse gained.)
template <int DIR>
__global__ __launch_bounds__(64 * kRows32Waves) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gamma_real_r32(
    Geo g, FftDesc fd, double2 *X2, double *f, double scale) {
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double2 *tab = lds;
  double *T = (double *)(lds + kR32BlueTab) + wv * kR32T;
  r32_stage_blue<DIR>(tab, fd);
  __syncthreads();
  const int S = g.nB * g.nA, N = g.N, nf = 2 * g.N - 1;  // 257 <= nf <= 511
  const int64_t groups = (S + 4 * kRows32Waves - 1) / (4 * kRows32Waves);
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    __syncthreads();
    const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
    const int pA = (int)(t * kRows32Waves + wv) * 4 + 2 * h, pB = pA + 1;
    const bool okA = pA < S, okB = pB < S;  // a column past the last one repeats it, unstored
    const int qA = okA ? pA : S - 1, qB = okB ? pB : S - 1;
    const int bA = qA / g.nA, aA = qA - bA * g.nA, bB = qB / g.nA, aB = qB - bB * g.nA;
    double2 *cA = X2 + x2_row(g, 0, bA) * g.ldA + aA, *cB = X2 + x2_row(g, 0, bB) * g.ldA + aB;
    const int64_t sA = (int64_t)x2_w(g, bA) * g.ldA, sB = (int64_t)x2_w(g, bB) * g.ldA;
    double2 v[32];
    if (DIR > 0) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = min(l + 32 * q, nf - 1);  // past nf: cleared by the chirp
        const bool dir = k < N;
        const int kk = dir ? k : nf - k;
        double2 a = cA[kk * sA], b = cB[kk * sB];
        if (!dir) a.y = -a.y, b.y = -b.y;
        if (k == 0) a.y = 0.0, b.y = 0.0;  // slot 0 of a real signal is real
        v[q] = make_double2(a.x - b.y, a.y + b.x);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = min(l + 32 * q, nf - 1);
        v[q] = make_double2(f[(int64_t)k * S + qA], f[(int64_t)k * S + qB]);
      }
    }
    blue32<DIR, 1>(v, nf, tab, T, lane);  // nf odd
    if (DIR > 0) {
#pragma unroll
      for (int q = 0; q < 16; ++q) asm volatile("" : "+v"(v[q].x), "+v"(v[q].y));
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = l + 32 * q;
        if (q < 8 || k < nf) {  // nf > 256
          if (okA) f[(int64_t)k * S + pA] = v[q].x * scale;
          if (okB) f[(int64_t)k * S + pB] = v[q].y * scale;
        }
      }
    } else {
      // Y(-k) = Y(nf - k) for k < N through the buffer (row q holds register q
      // of the 64 lanes); entry j of a lane sits at (c >> 5 - j) 64 + 32 h + (c & 31),
      // c = nf - l; k = 0 reads Y(0)
      const int c = nf - l;
      const int zb = (c >> 5) * 64 + 32 * h + (c & 31);
      auto z_at = [&](int j) { return (j == 0 && l == 0) ? 32 * h : zb - 64 * j; };
      double zr[8];
#pragma unroll
      for (int q = 0; q < 16; ++q) T[q * 64 + lane] = v[q].x;
      __builtin_amdgcn_wave_barrier();
      r32_fence();
#pragma unroll
      for (int j = 0; j < 8; ++j) zr[j] = T[max(z_at(j), 0)];
      __builtin_amdgcn_wave_barrier();
      r32_fence();
#pragma unroll
      for (int q = 0; q < 16; ++q) T[q * 64 + lane] = v[q].y;
      __builtin_amdgcn_wave_barrier();
      r32_fence();
      const double hs = 0.5 * scale;
#pragma unroll
      for (int j = 0; j < 8; ++j) {  // k = l + 32 j < N <= 256
        const double zi = T[max(z_at(j), 0)];
        const int k = l + 32 * j;
        if (k >= N) continue;
        // X_A = (Y + conj Y(-k)) / 2, X_B = (Y - conj Y(-k)) / (2 i)
        const double2 xa = make_double2(hs * (v[j].x + zr[j]), hs * (v[j].y - zi));
        const double2 xb = make_double2(hs * (v[j].y + zi), -hs * (v[j].x - zr[j]));
        if (okA) cA[k * sA] = xa;
        if (okB) cB[k * sB] = xb;
      }
      __builtin_amdgcn_wave_barrier();
      r32_fence();  // the buffer is read before the next group's transposes
    }
  }
}

// Inverse beta, paired, on the two-sequence register FFT: half-wavefront h
// of wavefront w takes the pair (mA, mA + 1), mA = m0 + 4 w + 2 h, so a
// group covers 4 W consecutive m of one n slot.  X1 rows are read straight
// (the |m'| runs are contiguous), Z(b') comes through the wavefront's
// transpose buffer, and each lane stores its A and B entries straight into
// X2 (the 4 W consecutive m of a group fill 4 W consecutive entries of each
// X2 row).  ODD: Pb odd (MW).
template <int ODD>
__global__ __launch_bounds__(64 * kRows32Waves) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_inv_beta_r32(
    Geo g, FftDesc fd, const double2 *__restrict__ X1, double2 *__restrict__ X2,
    const int64_t *__restrict__ nbase, const double2 *__restrict__ phase) {
  constexpr int C = 4 * kRows32Waves;  // m per group
  constexpr int kQB = 9;               // b = l + 32 j < nB <= 257
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double2 *tab = lds;
  double *T = (double *)(lds + kR32BlueTab) + wv * kR32T;
  r32_stage_blue<+1>(tab, fd);
  __syncthreads();
  const int L = g.L, Pb = g.Pb;  // 129 <= L <= 256
  const int mlo = -(L - 1) - g.ss;
  const int chunks = (g.nA + C - 1) / C;
  const int64_t groups = (int64_t)g.nG * chunks;
  const int nBf = g.nB & ~(kX2Blk - 1);  // rows b < nBf lie in full X2 blocks
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    // one barrier per group keeps the wavefronts that share 128-byte lines
    // in step (their partial-line stores meet in L2); the staggered barrier
    // of k_fwd_beta_r32 measured 14% slower here
    __syncthreads();
    const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
    const int nslot = (int)(t / chunks);
    const int n = n_of_slot(g, nslot);
    const int m0 = mlo + (int)(t % chunks) * C + 4 * wv;
    const int mA = m0 + 2 * h, mB = mA + 1;
    const bool act = nbase[n + g.N - 1] != kNoBlock;
    // outputs: columns past m = L - 1 are not stored, the MW_SS Nyquist
    // column m = -L holds zeros
    const bool stA = mA <= L - 1, stB = mB <= L - 1;
    const bool dA = act && mA >= -(L - 1), dB = act && mB >= -(L - 1);
    const int cA = m_slot_of(g, stA ? mA : 0), cB = m_slot_of(g, stB ? mB : 0);
    // X2 row of b = l + 32 j
    const int64_t rb = (int64_t)((l >> 3) * kX2Blk * g.nG + nslot * kX2Blk + (l & 7));
    auto row_of = [&](int j) {
      const int b = l + 32 * j;
      return X2 + (b < nBf ? rb + (int64_t)j * 32 * g.nG : x2_row(g, nslot, b)) * g.ldA;
    };
    if (!act) {  // block n not in the transform: zeros
#pragma unroll
      for (int j = 0; j < kQB; ++j) {
        if (l + 32 * j >= g.nB) continue;
        double2 *row = row_of(j);
        if (stA) row[cA] = make_double2(0.0, 0.0);
        if (stB) row[cB] = make_double2(0.0, 0.0);
      }
      continue;
    }
    // A member outside [-(L-1), L-1] reads an in-range row of the same
    // parity: the pair keeps its symmetries, so the separation of the other
    // member stays exact, and the member's own outputs are not stored.
    auto fixm = [&](int m) { return m < -(L - 1) ? m + 2 : m > L - 1 ? (L - 1) - ((m - (L - 1)) & 1) : m; };
    const double2 *rA = X1 + ((int64_t)nslot * g.nA + m_slot_of(g, fixm(mA))) * L;
    const double2 *rB = X1 + ((int64_t)nslot * g.nA + m_slot_of(g, fixm(mB))) * L;
    const bool negA = ((m0 + n) & 1) != 0;  // s_A = -1 (then s_B = +1); wave-uniform
    double2 v[32];
    // entry k = l + 32 q holds m' = k (k <= L - 1) or k - Pb; k >= Pb only
    // needs to be finite (the chirp clears it).  q < 4: k < 128 < L; q >= 8:
    // k >= 256 >= L.
    auto prep = [&](auto NEG) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = q < 8 ? l + 32 * q : min(l + 32 * q, Pb - 1);
        const int mp = q < 4 ? k : q >= 8 ? k - Pb : (k <= L - 1 ? k : k - Pb);
        const int am = mp < 0 ? -mp : mp;
        double2 a = rA[am], b = rB[am];
        if (q >= 4 && mp < 0) {  // X1(-m') = s X1(m')
          if (NEG) a = make_double2(-a.x, -a.y);
          else b = make_double2(-b.x, -b.y);
        }
        if (q == 0 && l == 0) {  // the antisymmetric member vanishes at m' = 0
          if (NEG) a = make_double2(0.0, 0.0);
          else b = make_double2(0.0, 0.0);
        }
        if (!ODD && q >= 4 && q <= 8 && k == L) a = b = make_double2(0.0, 0.0);  // MW_SS: m' = -L
        v[q] = cmul(make_double2(a.x - b.y, a.y + b.x), phase[k]);
      }
    };
    if (negA) prep(std::true_type{});
    else prep(std::false_type{});
    blue32<+1, ODD>(v, Pb, tab, T, lane);
    // Z(b') for b < nB through the transpose buffer (row q holds register q
    // of the 64 lanes): b' = Pb - 1 - b (MW) or Pb - b (MW_SS, b > 0), so
    // entry j of a lane sits at (c >> 5 - j) 64 + 32 h + (c & 31),
    // c = b' + 32 j
    const int c = ODD ? Pb - 1 - l : Pb - l;
    const int zb = (c >> 5) * 64 + 32 * h + (c & 31);
    auto z_at = [&](int j) { return (!ODD && j == 0 && l == 0) ? 32 * h : zb - 64 * j; };
    double zr[kQB];
#pragma unroll
    for (int q = 0; q < 16; ++q) T[q * 64 + lane] = v[q].x;
    __builtin_amdgcn_wave_barrier();
    r32_fence();
#pragma unroll
    for (int j = 0; j < kQB; ++j) zr[j] = T[max(z_at(j), 0)];
    __builtin_amdgcn_wave_barrier();
    r32_fence();
#pragma unroll
    for (int q = 0; q < 16; ++q) T[q * 64 + lane] = v[q].y;
    __builtin_amdgcn_wave_barrier();
    r32_fence();
    const double sA = negA ? -1.0 : 1.0;
#pragma unroll
    for (int j = 0; j < kQB; ++j) {
      const double zi = T[max(z_at(j), 0)];
      if (l + 32 * j >= g.nB) continue;
      const double2 su = make_double2(v[j].x + sA * zr[j], v[j].y + sA * zi);
      const double2 df = make_double2(v[j].x - sA * zr[j], v[j].y - sA * zi);
      double2 *row = row_of(j);
      if (stA) row[cA] = dA ? make_double2(0.5 * su.x, 0.5 * su.y) : make_double2(0.0, 0.0);
      if (stB) row[cB] = dB ? make_double2(0.5 * df.y, -0.5 * df.x) : make_double2(0.0, 0.0);
    }
    __builtin_amdgcn_wave_barrier();
    r32_fence();  // the buffer is read before the next group's transposes
  }
}

// Inverse beta, paired, on the 2048-point register FFT (Pb in [513, 1024],
// L = 257 ... 512): wavefront w of group t takes the pair (mA, mA + 1),
// mA = m0 + 2 w, as one sequence A + i B, so a group covers 2 W consecutive
// m of one n slot.  Samples k of the sequence sit in layout E (k = 2 (l + 32
// q) + h); X1 rows are read straight (64 consecutive m' per instruction),
// Z(b') comes through the wavefront's transpose buffer, and each lane stores
// its A and B entries straight into X2.  ODD: Pb odd (MW).
template <int ODD>
__global__ __launch_bounds__(64 * kRows32Waves) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_inv_beta_r2k(
    Geo g, FftDesc fd, const double2 *__restrict__ X1, double2 *__restrict__ X2,
    const int64_t *__restrict__ nbase, const double2 *__restrict__ phase) {
  constexpr int C = 2 * kRows32Waves;  // m per group
  constexpr int kQB = 9;               // b = 2 (l + 32 j) + h < nB <= 513
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double2 *tab = lds;
  double *T = (double *)(lds + kR2kTab) + wv * kR32T;
  r2k_stage_blue(tab, fd);
  __syncthreads();
  const int L = g.L, Pb = g.Pb;  // 257 <= L <= 512
  const int mlo = -(L - 1) - g.ss;
  const int chunks = (g.nA + C - 1) / C;
  const int64_t groups = (int64_t)g.nG * chunks;
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    // one barrier per group keeps the wavefronts that share 128-byte lines
    // of X2 in step (their partial-line stores meet in L2)
    __syncthreads();
    const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
    const int nslot = (int)(t / chunks);
    const int n = n_of_slot(g, nslot);
    const int mA = mlo + (int)(t % chunks) * C + 2 * wv, mB = mA + 1;
    const bool act = nbase[n + g.N - 1] != kNoBlock;
    // outputs: columns past m = L - 1 are not stored, the MW_SS Nyquist
    // column m = -L holds zeros
    const bool stA = mA <= L - 1, stB = mB <= L - 1;
    const bool dA = act && mA >= -(L - 1), dB = act && mB >= -(L - 1);
    const int cA = m_slot_of(g, stA ? mA : 0), cB = m_slot_of(g, stB ? mB : 0);
    if (!act) {  // block n not in the transform: zeros
#pragma unroll
      for (int j = 0; j < kQB; ++j) {
        const int b = 2 * (l + 32 * j) + h;
        if (b >= g.nB) continue;
        double2 *row = X2 + x2_row(g, nslot, b) * g.ldA;
        if (stA) row[cA] = make_double2(0.0, 0.0);
        if (stB) row[cB] = make_double2(0.0, 0.0);
      }
      continue;
    }
    // A member outside [-(L-1), L-1] reads an in-range row of the same
    // parity: the pair keeps its symmetries, so the separation of the other
    // member stays exact, and the member's own outputs are not stored.
    auto fixm = [&](int m) { return m < -(L - 1) ? m + 2 : m > L - 1 ? (L - 1) - ((m - (L - 1)) & 1) : m; };
    const double2 *rA = X1 + ((int64_t)nslot * g.nA + m_slot_of(g, fixm(mA))) * L;
    const double2 *rB = X1 + ((int64_t)nslot * g.nA + m_slot_of(g, fixm(mB))) * L;
    const bool negA = ((mA + n) & 1) != 0;  // s_A = -1 (then s_B = +1); wave-uniform
    double2 v[32];
    // entry k holds m' = k (k <= L - 1) or k - Pb; k >= Pb only needs to be
    // finite (the chirp clears it)
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = min(2 * (l + 32 * q) + h, Pb - 1);
      const int mp = k <= L - 1 ? k : k - Pb;
      const int am = mp < 0 ? -mp : mp;
      double2 a = rA[am], b = rB[am];
      if (mp < 0) {  // X1(-m') = s X1(m')
        if (negA) a = make_double2(-a.x, -a.y);
        else b = make_double2(-b.x, -b.y);
      }
      if (k == 0) {  // the antisymmetric member vanishes at m' = 0
        if (negA) a = make_double2(0.0, 0.0);
        else b = make_double2(0.0, 0.0);
      }
      if (!ODD && k == L) a = b = make_double2(0.0, 0.0);  // MW_SS: m' = -L
      v[q] = cmul(make_double2(a.x - b.y, a.y + b.x), phase[k]);
    }
    blue2k<+1>(v, Pb, tab, fd.bhat_b, T, lane);  // twice the DFT
    // Z(b') for b < nB through the buffer (entry q * 64 + lane holds
    // register q of the lane): b' = Pb - 1 - b (MW) or Pb - b (MW_SS, b > 0)
    double2 *T2 = (double2 *)T;
#pragma unroll
    for (int q = 0; q < 16; ++q) T2[q * 64 + lane] = v[q];
    __builtin_amdgcn_wave_barrier();
    r32_fence();
    const double sA = negA ? -1.0 : 1.0;
#pragma unroll
    for (int j = 0; j < kQB; ++j) {
      const int b = 2 * (l + 32 * j) + h;
      const int bp = ODD ? Pb - 1 - b : (b == 0 ? 0 : Pb - b);
      const double2 zp = T2[max(bp, 0) >> 6 << 6 | (bp & 1) << 5 | ((max(bp, 0) >> 1) & 31)];
      if (b >= g.nB) continue;
      // A = (Z + s_A Z') / 2, B = (Z - s_A Z') / (2 i), with the factor 2 of blue2k
      const double2 su = make_double2(v[j].x + sA * zp.x, v[j].y + sA * zp.y);
      const double2 df = make_double2(v[j].x - sA * zp.x, v[j].y - sA * zp.y);
      double2 *row = X2 + x2_row(g, nslot, b) * g.ldA;
      if (stA) row[cA] = dA ? make_double2(0.25 * su.x, 0.25 * su.y) : make_double2(0.0, 0.0);
      if (stB) row[cB] = dB ? make_double2(0.25 * df.y, -0.25 * df.x) : make_double2(0.0, 0.0);
    }
    __builtin_amdgcn_wave_barrier();
    r32_fence();  // the buffer is read before the next group's transforms
  }
}

// Inverse beta, paired, on the 4096-point register FFT (Pb in [1025, 2048],
// L = 513 ... 1024): the pair (mA, mA + 1), mA = m0 + 2 p, of wavefront pair
// p as one sequence A + i B, sample k in wavefront k & 1 (layout E of
// fft_reg4k.hpp, k = 4 (l + 32 q) + 2 h + e), so a group covers W
// consecutive m of one n slot.  Z(b') has the parity of b (b' = Pb - 1 - b
// with Pb odd, or Pb - b with Pb even) and comes through the wavefront's own
// transpose buffer.  As in k_rows_r4k, four wavefronts per workgroup at one
// per SIMD keep the Bluestein chirp and kernel spectrum of their lanes in
// registers (16-22% faster than eight wavefronts reloading them).  ODD: Pb
// odd (MW).
constexpr int kInvBeta4kWaves = 4;

template <int ODD>
__global__ __launch_bounds__(64 * kInvBeta4kWaves) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_inv_beta_r4k(
    Geo g, FftDesc fd, const double2 *__restrict__ X1, double2 *__restrict__ X2,
    const int64_t *__restrict__ nbase, const double2 *__restrict__ phase) {
  constexpr int W = kInvBeta4kWaves;
  constexpr int C = W;        // m per group (a pair per two wavefronts)
  constexpr int kQB = 9;      // b = 4 (l + 32 j) + 2 h + e < nB <= 1025
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), e = wv & 1;
  double2 *tab = lds;
  double *T = (double *)(lds + kR4kTab) + wv * kR32T;
  const double *Tp = (double *)(lds + kR4kTab) + (wv ^ 1) * kR32T;
  r4k_stage_blue(tab, fd);
  double2 cpre[16], spre[32];  // chirp of samples q < 16, kernel spectrum of outputs q < 32
  {
    const int lane = threadIdx.x & 63, l = lane & 31, h = lane >> 5;
#pragma unroll
    for (int q = 0; q < 16; ++q) cpre[q] = r4k_chirp<+1>(fd.chirp, g.Pb, 4 * (l + 32 * q) + 2 * h + e);
#pragma unroll
    for (int q = 0; q < 32; ++q) spre[q] = fd.bhat_b[l + 32 * q + 1024 * h + 2048 * e];
  }
  __syncthreads();
  const int L = g.L, Pb = g.Pb;  // 513 <= L <= 1024
  const int mlo = -(L - 1) - g.ss;
  const int chunks = (g.nA + C - 1) / C;
  const int64_t groups = (int64_t)g.nG * chunks;
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    __syncthreads();  // one barrier per group keeps the workgroup's stores to X2 rows in step
    const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
    const int nslot = (int)(t / chunks);
    const int n = n_of_slot(g, nslot);
    const int mA = mlo + (int)(t % chunks) * C + 2 * (wv >> 1), mB = mA + 1;
    const bool act = nbase[n + g.N - 1] != kNoBlock;  // the whole workgroup
    const bool stA = mA <= L - 1, stB = mB <= L - 1;
    const bool dA = act && mA >= -(L - 1), dB = act && mB >= -(L - 1);
    const int cA = m_slot_of(g, stA ? mA : 0), cB = m_slot_of(g, stB ? mB : 0);
    if (!act) {  // block n not in the transform: zeros
#pragma unroll
      for (int j = 0; j < kQB; ++j) {
        const int b = 4 * (l + 32 * j) + 2 * h + e;
        if (b >= g.nB) continue;
        double2 *row = X2 + x2_row(g, nslot, b) * g.ldA;
        if (stA) row[cA] = make_double2(0.0, 0.0);
        if (stB) row[cB] = make_double2(0.0, 0.0);
      }
      continue;
    }
    auto fixm = [&](int m) { return m < -(L - 1) ? m + 2 : m > L - 1 ? (L - 1) - ((m - (L - 1)) & 1) : m; };
    const double2 *rA = X1 + ((int64_t)nslot * g.nA + m_slot_of(g, fixm(mA))) * L;
    const double2 *rB = X1 + ((int64_t)nslot * g.nA + m_slot_of(g, fixm(mB))) * L;
    const bool negA = ((mA + n) & 1) != 0;  // s_A = -1 (then s_B = +1); uniform over the pair
    double2 v[32];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = min(4 * (l + 32 * q) + 2 * h + e, Pb - 1);
      const int mp = k <= L - 1 ? k : k - Pb;
      const int am = mp < 0 ? -mp : mp;
      double2 a = rA[am], b = rB[am];
      if (mp < 0) {  // X1(-m') = s X1(m')
        if (negA) a = make_double2(-a.x, -a.y);
        else b = make_double2(-b.x, -b.y);
      }
      if (k == 0) {  // the antisymmetric member vanishes at m' = 0
        if (negA) a = make_double2(0.0, 0.0);
        else b = make_double2(0.0, 0.0);
      }
      if (!ODD && k == L) a = b = make_double2(0.0, 0.0);  // MW_SS: m' = -L
      v[q] = cmul(make_double2(a.x - b.y, a.y + b.x), phase[k]);
      if ((q & 3) == 3) r32_fence();  // loads issued four registers at a time
    }
    blue4k_pre<+1>(v, tab, cpre, spre, T, Tp, e, lane);
    // Z(b') through the buffer: entry q 64 + lane holds register q; sample
    // j = 2 i + e of this wavefront sits at (i >> 6) 64 + 32 (i & 1) + ((i >> 1) & 31)
    // (written from the registers after the transform: writing it as the last
    // DFT-32 makes it and reading Z(b) back measured 2-3% slower)
    double2 *T2 = (double2 *)T;
#pragma unroll
    for (int q = 0; q < 16; ++q) T2[q * 64 + lane] = v[q];
    __builtin_amdgcn_wave_barrier();
    r32_fence();
    const double sA = negA ? -1.0 : 1.0;
#pragma unroll
    for (int j = 0; j < kQB; ++j) {
      const int b = 4 * (l + 32 * j) + 2 * h + e;
      const int bp = max(ODD ? Pb - 1 - b : (b == 0 ? 0 : Pb - b), e);
      const int ip = bp >> 1;
      const double2 zp = T2[((ip >> 6) << 6) | ((ip & 1) << 5) | ((ip >> 1) & 31)];
      if (b >= g.nB) continue;
      // A = (Z + s_A Z') / 2, B = (Z - s_A Z') / (2 i)
      const double2 su = make_double2(v[j].x + sA * zp.x, v[j].y + sA * zp.y);
      const double2 df = make_double2(v[j].x - sA * zp.x, v[j].y - sA * zp.y);
      double2 *row = X2 + x2_row(g, nslot, b) * g.ldA;
      if (stA) row[cA] = dA ? make_double2(0.5 * su.x, 0.5 * su.y) : make_double2(0.0, 0.0);
      if (stB) row[cB] = dB ? make_double2(0.5 * df.y, -0.5 * df.x) : make_double2(0.0, 0.0);
    }
    __builtin_amdgcn_wave_barrier();
    r32_fence();  // the buffer is read before the next group's transforms
  }
}

// forward beta, paired (not the adjoint): the Bluestein DFT and the
// convolution with the even weight taps run once per pair.  The two odd
// taps w(+-1) = -+ i pi/2 cancel in the fold G'(m') = G(m') + s G(-m') for
// m' > 0 and give 2 pi i pi X~(1) at m' = 0 for the antisymmetric member.
// The pole samples (b = L-1 for MW, b = 0 and L for MW_SS) of the
// antisymmetric member break its antisymmetry; they are left out of the
// pair and added back through their folded responses (pole table).
// On the two-sequence register FFT, half-wavefront h of wavefront w takes the pair
// (mA, mA + 1), mA = m0 + 4 u + 2 h, u = quad(w) (a permutation of the
// wavefronts, set out at the group barrier).  In layout P the remap to the
// convolution index keeps m'' >= 0 in the lane's own registers 0..7 and
// gathers m'' < 0 (registers 24..31) through the transpose buffer; the fold
// G(m') + s G(-m') takes G(m') from the lane's own registers and G(-m')
// through the buffer.  The even weight spectrum is real and even, and is
// read from global memory (L1): LDS holds the Bluestein tables and eight
// transpose buffers.  ODD: Pb odd (MW).
template <int ODD>
__global__ __launch_bounds__(64 * kRows32Waves) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_fwd_beta_r32(
    Geo g, FftDesc fb, const double2 *__restrict__ X2, double2 *__restrict__ Gp,
    const int64_t *__restrict__ nbase, const double2 *__restrict__ phase, const double2 *__restrict__ wre,
    const double2 *__restrict__ pole) {
  constexpr int C = 4 * kRows32Waves;  // m per group
  constexpr double kPi = 3.141592653589793238462643383279502884;
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double2 *tab = lds;
  double *T = (double *)(lds + kR32BlueTab) + wv * kR32T;
  r32_stage_blue<-1>(tab, fb);
  __syncthreads();  // tables staged
  const int L = g.L, Pb = g.Pb;  // 129 <= L <= 256
  const int mlo = -(L - 1) - g.ss;
  const int chunks = (g.nA + C - 1) / C;
  const int64_t groups = (int64_t)g.nG * chunks;
  const int C2 = 2 * g.NS;
  // One barrier per group keeps the wavefronts of a workgroup within half a
  // group of each other.  Even wavefronts pass it before their loads, odd
  // ones after their Bluestein DFT: the two wavefronts of a SIMD are then
  // half a group apart, and one computes while the other waits for its
  // loads.  Every wavefront passes it once per group.  Neighbouring quads of
  // m share 128-byte lines, so the even wavefronts take the first half of
  // the group's quads and the odd ones the second: a line's users then pass
  // the barrier together, and the line is still in L2 when the last of
  // them reads it.
  const bool late = (wv & 1) != 0;
  const int quad = (wv & 1) * (kRows32Waves / 2) + (wv >> 1);
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    if (!late) __syncthreads();
    const int nslot = (int)(t / chunks);
    const int n = n_of_slot(g, nslot);
    if (nbase[n + g.N - 1] == kNoBlock) {  // block not transformed
      if (late) __syncthreads();
      continue;
    }
    const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
    const int m0 = mlo + (int)(t % chunks) * C + 4 * quad;
    const int mA = m0 + 2 * h, mB = mA + 1;
    const bool negA = ((m0 + n) & 1) != 0;  // s_A = -1 (then s_B = +1); wave-uniform
    const double sA = negA ? -1.0 : 1.0;
    // a member outside [-(L-1), L-1] reads an in-range column (the extension
    // imposes its symmetry; its outputs are not stored)
    auto fixm = [&](int m) { return m < -(L - 1) ? m + 2 : m > L - 1 ? (L - 1) - ((m - (L - 1)) & 1) : m; };
    const double2 *colA = X2 + m_slot_of(g, fixm(mA)), *colB = X2 + m_slot_of(g, fixm(mB));
    // pole samples of the antisymmetric member (one address per half-wavefront)
    const double2 *colX = negA ? colA : colB;
    const double2 p0 = colX[x2_row(g, nslot, ODD ? L - 1 : 0) * g.ldA];
    const double2 p1 = ODD ? make_double2(0.0, 0.0) : colX[x2_row(g, nslot, L) * g.ldA];
    double2 v[32];
    // extended samples: MW F(2L-2-b) for b >= L, MW_SS F(2L-b) for b > L;
    // entries k >= Pb only need to be finite (the chirp clears them)
    auto prep = [&](auto NEG) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = q < 8 ? l + 32 * q : min(l + 32 * q, Pb - 1);
        const bool direct = q < 4 || (ODD ? k < L : k <= L);
        const int kr = direct ? k : (ODD ? 2 * L - 2 - k : 2 * L - k);
        const int64_t off = x2_row(g, nslot, kr) * g.ldA;
        double2 a = colA[off], b = colB[off];
        if (!direct) {
          if (NEG) a = make_double2(-a.x, -a.y);
          else b = make_double2(-b.x, -b.y);
        }
        const bool pk = ODD ? (q >= 4 && k == L - 1) : ((q == 0 && l == 0) || (q >= 4 && k == L));
        if (pk) {  // the antisymmetric member's pole samples go through the pole table
          if (NEG) a = make_double2(0.0, 0.0);
          else b = make_double2(0.0, 0.0);
        }
        v[q] = make_double2(a.x - b.y, a.y + b.x);
      }
    };
    if (negA) prep(std::true_type{});
    else prep(std::false_type{});
    blue32<-1, ODD>(v, Pb, tab, T, lane);
    if (late) __syncthreads();
    // remap: u(r) = F~(m'') phase(m''), m'' = r (r <= L-1) or r - 1024
    // (r >= 1024 - (L-1)), zero otherwise; F~(+-1) for the m' = 0 fold.
    // Entry s of F~ sits at (s >> 5) 64 + 32 h + (s & 31) of the buffer.
    double2 z1, zm1;
    // phase(l) of this lane and the 16 lane-uniform block factors after the
    // table: one load per lane instead of 16
    const double2 pl = phase[l + L - 1];
    const double2 *pblk = phase + 2 * L - 1;
    {
      const int d = l + Pb - 1024;  // s = d + 32 q for r = l + 32 q >= 768
      const int gb = (d >> 5) * 64 + 32 * h + (d & 31);
      const int i1 = 32 * h + 1, im1 = ((Pb - 1) >> 5) * 64 + 32 * h + ((Pb - 1) & 31);
      double gr[8];
#pragma unroll
      for (int q = 0; q < 16; ++q) T[q * 64 + lane] = v[q].x;
      __builtin_amdgcn_wave_barrier();
      r32_fence();
#pragma unroll
      for (int q = 24; q < 32; ++q) gr[q - 24] = T[gb + 64 * q];
      const double r1 = T[i1], rm1 = T[im1];
      __builtin_amdgcn_wave_barrier();
      r32_fence();
#pragma unroll
      for (int q = 0; q < 16; ++q) T[q * 64 + lane] = v[q].y;
      __builtin_amdgcn_wave_barrier();
      r32_fence();
#pragma unroll
      for (int q = 24; q < 32; ++q) {
        const int r = l + 32 * q;
        const bool ok = r >= 1024 - (L - 1);
        const double2 x = make_double2(gr[q - 24], T[gb + 64 * q]);
        // phase(m'' = l + 32 (q - 32)) = phase(l) exp(-i 32 (q - 32) pi/(2L-1))
        v[q] = ok ? cmul(x, cmul(pl, pblk[q - 24])) : make_double2(0.0, 0.0);
      }
      z1 = cmul(make_double2(r1, T[i1]), phase[L]);
      zm1 = cmul(make_double2(rm1, T[im1]), phase[L - 2]);
      __builtin_amdgcn_wave_barrier();
      r32_fence();
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int r = l + 32 * q;
        v[q] = r <= L - 1 ? cmul(v[q], cmul(pl, pblk[q + 8])) : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int q = 8; q < 24; ++q) v[q] = make_double2(0.0, 0.0);
    }
    // convolution with the even weight taps (real spectrum, natural order)
    // (the first of the four output groups of the last DFT-32, k % 4 == 0,
    // is loaded before the transform, whose first stages hide the latency)
    double wk[8];
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) wk[k2] = wre[l + 128 * k2].x;
    fft32<+1>(v, T, tab, lane, [&](int k, double2 &z) {
      z = cscale(z, (k & 3) == 0 ? wk[k >> 2] : wre[l + 32 * k].x);
    });
    fft32<-1>(v, T, tab, lane, NoPost());
    // separate and fold, m' = l + 32 j < L: G(m') is register j of the lane,
    // G(-m') = G(1024 - m') comes through the buffer (rows 0 and 24..31)
    constexpr double sc = 2.0 * kPi / 1024.0;
    const int zb = 31 * 64 + 32 * h + 32 - l;  // entry of 1024 - m' for l > 0, minus 64 (31 - j)
    auto zm_at = [&](int j) { return l == 0 ? (j == 0 ? 32 * h : (32 - j) * 64 + 32 * h) : zb - 64 * j; };
    double zmr[8];
    T[lane] = v[0].x;
#pragma unroll
    for (int q = 24; q < 32; ++q) T[q * 64 + lane] = v[q].x;
    __builtin_amdgcn_wave_barrier();
    r32_fence();
#pragma unroll
    for (int j = 0; j < 8; ++j) zmr[j] = T[zm_at(j)];
    __builtin_amdgcn_wave_barrier();
    r32_fence();
    T[lane] = v[0].y;
#pragma unroll
    for (int q = 24; q < 32; ++q) T[q * 64 + lane] = v[q].y;
    __builtin_amdgcn_wave_barrier();
    r32_fence();
    const bool okA = mA >= -(L - 1) && mA <= L - 1, okB = mB >= -(L - 1) && mB <= L - 1;
    const int an = n < 0 ? -n : n, sn = n < 0;
    const int cn0 = an == 0 ? 0 : sn, cn1 = an == 0 ? g.NS - 1 : sn;
    int cm0[2], cm1[2];
    double2 *row[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int m = e ? mB : mA;
      const int am = m < 0 ? -m : m, sm = m < 0;
      cm0[e] = am == 0 ? 0 : sm;
      cm1[e] = am == 0 ? 1 : sm;
      row[e] = Gp + ((int64_t)an * L + am) * C2 * L;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int mp = l + 32 * j;
      const double zmi = T[zm_at(j)];
      if (mp >= L) continue;
      const double2 zp = make_double2(v[j].x * sc, v[j].y * sc);
      const double2 zm = make_double2(zmr[j] * sc, zmi * sc);
      double2 GA, GB;
      if (mp > 0) {
        GA = make_double2(zp.x + sA * zm.x, zp.y + sA * zm.y);
        GB = make_double2(zp.y - sA * zm.y, -(zp.x - sA * zm.x));
      } else {
        const double pp = kPi * kPi;
        const double2 dz = make_double2(z1.x - zm1.x, z1.y - zm1.y);
        if (!negA) {
          GA = zp;
          GB = make_double2(pp * dz.x, pp * dz.y);
        } else {
          GA = make_double2(-pp * dz.y, pp * dz.x);
          GB = make_double2(zp.y, -zp.x);
        }
      }
      double2 corr = cmul(p0, pole[mp]);
      if (!ODD) corr = cadd(corr, cmul(p1, pole[L + mp]));
      if (negA) GA = cadd(GA, corr);
      else GB = cadd(GB, corr);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        if (!(e ? okB : okA)) continue;
        const double2 G0 = e ? GB : GA;
#pragma nounroll
        for (int cn = cn0; cn <= cn1; ++cn) {
#pragma nounroll
          for (int cm = cm0[e]; cm <= cm1[e]; ++cm) {
            const double sg = ((cm != cn) && (mp & 1)) ? -1.0 : 1.0;
            row[e][(int64_t)(cm + 2 * cn) * L + mp] = make_double2(sg * G0.x, sg * G0.y);
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    r32_fence();  // the buffer is read before the next group's transposes
  }
}

// Forward beta, paired, on the 2048-point register FFT (Pb in [513, 1024],
// L = 257 ... 512, convolution length Pw = 2048): the algorithm of
// k_fwd_beta_r32 with one pair (mA, mA + 1), mA = m0 + 2 w, per wavefront,
// samples k in layout E (k = 2 (l + 32 q) + h).  After the Bluestein DFT the
// remap to the convolution index r keeps m'' = r >= 0 in the lane's own
// registers 0..7 and gathers m'' = r - 2048 < 0 (registers 24..31) through
// the transpose buffer; the convolution is a forward 2048-point FFT, the
// product with the real weight spectrum (global memory) and the inverse;
// the fold takes G(m') from registers 0..7 and G(-m') = G(2048 - m') from
// registers 24..31 through the buffer.  ODD: Pb odd (MW).
template <int ODD>
__global__ __launch_bounds__(64 * kRows32Waves) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_fwd_beta_r2k(
    Geo g, FftDesc fb, const double2 *__restrict__ X2, double2 *__restrict__ Gp,
    const int64_t *__restrict__ nbase, const double2 *__restrict__ phase, const double2 *__restrict__ wre,
    const double2 *__restrict__ pole) {
  constexpr int C = 2 * kRows32Waves;  // m per group
  constexpr double kPi = 3.141592653589793238462643383279502884;
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double2 *tab = lds;
  double *T = (double *)(lds + kR2kTab) + wv * kR32T;
  double2 *T2 = (double2 *)T;
  r2k_stage_blue(tab, fb);
  __syncthreads();  // tables staged
  const int L = g.L, Pb = g.Pb;  // 257 <= L <= 512
  const int mlo = -(L - 1) - g.ss;
  const int chunks = (g.nA + C - 1) / C;
  const int64_t groups = (int64_t)g.nG * chunks;
  const int C2 = 2 * g.NS;
  auto fixm = [&](int m) { return m < -(L - 1) ? m + 2 : m > L - 1 ? (L - 1) - ((m - (L - 1)) & 1) : m; };
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    // one barrier per group keeps the wavefronts that share 128-byte lines
    // of X2 in step
    __syncthreads();
    const int nslot = (int)(t / chunks);
    const int n = n_of_slot(g, nslot);
    if (nbase[n + g.N - 1] == kNoBlock) continue;  // block not transformed (the whole workgroup)
    const int mA = mlo + (int)(t % chunks) * C + 2 * wv, mB = mA + 1;
    const bool negA = ((mA + n) & 1) != 0;  // s_A = -1 (then s_B = +1); wave-uniform
    const double sA = negA ? -1.0 : 1.0;
    double2 v[32];
    {
      // extended samples: MW F(2L-2-b) for b >= L, MW_SS F(2L-b) for b > L;
      // entries k >= Pb only need to be finite (the chirp clears them); a
      // member outside [-(L-1), L-1] reads an in-range column (the extension
      // imposes its symmetry; its outputs are not stored)
      const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
      const double2 *colA = X2 + m_slot_of(g, fixm(mA)), *colB = X2 + m_slot_of(g, fixm(mB));
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = min(2 * (l + 32 * q) + h, Pb - 1);
        const bool direct = ODD ? k < L : k <= L;
        const int kr = direct ? k : (ODD ? 2 * L - 2 - k : 2 * L - k);
        const int64_t off = x2_row(g, nslot, kr) * g.ldA;
        double2 a = colA[off], b = colB[off];
        if (!direct) {
          if (negA) a = make_double2(-a.x, -a.y);
          else b = make_double2(-b.x, -b.y);
        }
        const bool pk = ODD ? k == L - 1 : (k == 0 || k == L);
        if (pk) {  // the antisymmetric member's pole samples go through the pole table
          if (negA) a = make_double2(0.0, 0.0);
          else b = make_double2(0.0, 0.0);
        }
        v[q] = make_double2(a.x - b.y, a.y + b.x);
      }
    }
    blue2k<-1>(v, Pb, tab, fb.bhat_f, T, loop_lane());  // twice the DFT, registers 0..15
    double2 z1, zm1;
    {
      // remap: u(r) = F~(m'') phase(m''), m'' = r (r <= L-1) or r - 2048
      // (r >= 2048 - (L-1)), zero otherwise; F~(+-1) for the m' = 0 fold.
      // Entry s of F~ sits at (s >> 6) 64 + 32 (s & 1) + ((s >> 1) & 31).
      const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
      auto at = [](int s) { return ((s >> 6) << 6) | ((s & 1) << 5) | ((s >> 1) & 31); };
#pragma unroll
      for (int q = 0; q < 16; ++q) T2[q * 64 + lane] = v[q];
      __builtin_amdgcn_wave_barrier();
      r32_fence();
      // phase(m'' = 2 l + h + 64 j) = phase(2 l + h) times block factor j
      const double2 pl = phase[2 * l + h + L - 1];
      const double2 *pblk = phase + 2 * L - 1 + 16;
#pragma unroll
      for (int q = 24; q < 32; ++q) {
        const int r = 2 * (l + 32 * q) + h;
        const bool ok = r >= 2048 - (L - 1);
        const double2 x = T2[at(r - 2048 + Pb)];  // r - 2048 + Pb >= Pb - 512 >= 1
        v[q] = ok ? cmul(x, cmul(pl, pblk[q - 24])) : make_double2(0.0, 0.0);
      }
      z1 = cscale(cmul(T2[at(1)], phase[L]), 0.5);
      zm1 = cscale(cmul(T2[at(Pb - 1)], phase[L - 2]), 0.5);
      __builtin_amdgcn_wave_barrier();
      r32_fence();
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int r = 2 * (l + 32 * q) + h;
        v[q] = r <= L - 1 ? cmul(v[q], cmul(pl, pblk[q + 8])) : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int q = 8; q < 24; ++q) v[q] = make_double2(0.0, 0.0);
    }
    {
      // convolution with the even weight taps: real spectrum, natural order
      const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
      fft2k_fwd(v, T, tab, lane, [&](int q, double2 &z) { z = cscale(z, wre[l + 32 * q + 1024 * h].x); });
      fft2k_inv<false>(v, T, tab, lane, [](int, double2 &) {});
    }
    // separate and fold, m' = 2 (l + 32 j) + h < L: G(m') is register j of
    // the lane, G(-m') = G(2048 - m') comes through the buffer (registers
    // 24..31 at (q - 24) 64 + lane, register 0 at 512 + lane)
    const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
    // pole samples of the antisymmetric member (one address per wavefront)
    const double2 *colX = X2 + m_slot_of(g, fixm(negA ? mA : mB));
    const double2 p0 = colX[x2_row(g, nslot, ODD ? L - 1 : 0) * g.ldA];
    const double2 p1 = ODD ? make_double2(0.0, 0.0) : colX[x2_row(g, nslot, L) * g.ldA];
    // 2 pi / 2048 with the factor 2 of the Bluestein DFT
    constexpr double sc = 2.0 * kPi / 4096.0;
#pragma unroll
    for (int q = 24; q < 32; ++q) T2[(q - 24) * 64 + lane] = v[q];
    T2[512 + lane] = v[0];
    __builtin_amdgcn_wave_barrier();
    r32_fence();
    const bool okA = mA >= -(L - 1) && mA <= L - 1, okB = mB >= -(L - 1) && mB <= L - 1;
    const int an = n < 0 ? -n : n, sn = n < 0;
    // G' rows of each member: combinations (cm, cn), cm in [cm0, cm1], cn in
    // [cn0, cn1] -- a single one unless m = 0 or n = 0 (both signs of that
    // index hold the same row)
    const int cn0 = an == 0 ? 0 : sn, cn1 = an == 0 ? g.NS - 1 : sn;
    int cm0[2], cm1[2];
    double2 *row[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int m = e ? mB : mA;
      const int am = m < 0 ? -m : m, sm = m < 0;
      cm0[e] = am == 0 ? 0 : sm;
      cm1[e] = am == 0 ? 1 : sm;
      row[e] = Gp + ((int64_t)an * L + am) * C2 * L;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int mp = 2 * (l + 32 * j) + h;
      const int rp = 2048 - mp;
      const double2 zmv = T2[(((rp >> 6) - 24) << 6) + ((rp & 1) << 5) + ((rp >> 1) & 31)];
      if (mp >= L) continue;
      const double2 zp = make_double2(v[j].x * sc, v[j].y * sc);
      const double2 zm = make_double2(zmv.x * sc, zmv.y * sc);
      double2 GA, GB;
      if (mp > 0) {
        GA = make_double2(zp.x + sA * zm.x, zp.y + sA * zm.y);
        GB = make_double2(zp.y - sA * zm.y, -(zp.x - sA * zm.x));
      } else {
        const double pp = kPi * kPi;
        const double2 dz = make_double2(z1.x - zm1.x, z1.y - zm1.y);
        if (!negA) {
          GA = zp;
          GB = make_double2(pp * dz.x, pp * dz.y);
        } else {
          GA = make_double2(-pp * dz.y, pp * dz.x);
          GB = make_double2(zp.y, -zp.x);
        }
      }
      double2 corr = cmul(p0, pole[mp]);
      if (!ODD) corr = cadd(corr, cmul(p1, pole[L + mp]));
      if (negA) GA = cadd(GA, corr);
      else GB = cadd(GB, corr);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        if (!(e ? okB : okA)) continue;
        const double2 G0 = e ? GB : GA;
#pragma nounroll
        for (int cn = cn0; cn <= cn1; ++cn) {
#pragma nounroll
          for (int cm = cm0[e]; cm <= cm1[e]; ++cm) {
            const double sg = ((cm != cn) && (mp & 1)) ? -1.0 : 1.0;
            row[e][(int64_t)(cm + 2 * cn) * L + mp] = make_double2(sg * G0.x, sg * G0.y);
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    r32_fence();  // the buffer is read before the next group's transforms
  }
}

// Forward beta, paired, on the 4096-point register FFT (Pb in [1025, 2048],
// L = 513 ... 1024, convolution length Pw = 4096): the algorithm of
// k_fwd_beta_r2k with one pair (mA, mA + 1), mA = m0 + 2 p, per pair of
// wavefronts p, sample k in wavefront k & 1 (layout E of fft_reg4k.hpp).
// The remap keeps m'' = r >= 0 in the wavefront's own registers 0..7 and
// gathers m'' = r - 4096 < 0 (registers 24..31) from entry s = r - 4096 + Pb
// of the Bluestein output, in the partner wavefront when Pb is odd (through
// its transpose buffer, between workgroup barriers); the fold takes G(-m') =
// G(4096 - m'), of the parity of m', from the wavefront's own registers
// 24..31.  ODD: Pb odd (MW).
template <int ODD>
__global__ __launch_bounds__(64 * kR4kWaves) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_fwd_beta_r4k(
    Geo g, FftDesc fb, const double2 *__restrict__ X2, double2 *__restrict__ Gp,
    const int64_t *__restrict__ nbase, const double2 *__restrict__ phase, const double2 *__restrict__ wre,
    const double2 *__restrict__ pole) {
  constexpr int W = kR4kWaves;
  constexpr int C = W;  // m per group
  constexpr double kPi = 3.141592653589793238462643383279502884;
  extern __shared__ double2 lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), e = wv & 1;
  double2 *tab = lds;
  double *T = (double *)(lds + kR4kTab) + wv * kR32T;
  const double *Tp = (double *)(lds + kR4kTab) + (wv ^ 1) * kR32T;
  double2 *T2 = (double2 *)T;
  r4k_stage_blue(tab, fb);
  __syncthreads();  // tables staged
  const int L = g.L, Pb = g.Pb;  // 513 <= L <= 1024
  const int mlo = -(L - 1) - g.ss;
  const int chunks = (g.nA + C - 1) / C;
  const int64_t groups = (int64_t)g.nG * chunks;
  const int C2 = 2 * g.NS;
  auto fixm = [&](int m) { return m < -(L - 1) ? m + 2 : m > L - 1 ? (L - 1) - ((m - (L - 1)) & 1) : m; };
  // entry i of a wavefront's samples (j = 2 i + e) in its buffer
  auto at = [](int i) { return ((i >> 6) << 6) | ((i & 1) << 5) | ((i >> 1) & 31); };
  for (int64_t t = xcd_tile(blockIdx.x, gridDim.x); t < groups; t += gridDim.x) {
    __syncthreads();  // one barrier per group keeps the workgroup's X2 reads in step
    const int nslot = (int)(t / chunks);
    const int n = n_of_slot(g, nslot);
    if (nbase[n + g.N - 1] == kNoBlock) continue;  // block not transformed (the whole workgroup)
    const int mA = mlo + (int)(t % chunks) * C + 2 * (wv >> 1), mB = mA + 1;
    const bool negA = ((mA + n) & 1) != 0;  // s_A = -1 (then s_B = +1); uniform over the pair
    const double sA = negA ? -1.0 : 1.0;
    double2 v[32];
    {
      // extended samples: MW F(2L-2-b) for b >= L, MW_SS F(2L-b) for b > L
      const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
      const double2 *colA = X2 + m_slot_of(g, fixm(mA)), *colB = X2 + m_slot_of(g, fixm(mB));
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int k = min(4 * (l + 32 * q) + 2 * h + e, Pb - 1);
        const bool direct = ODD ? k < L : k <= L;
        const int kr = direct ? k : (ODD ? 2 * L - 2 - k : 2 * L - k);
        const int64_t off = x2_row(g, nslot, kr) * g.ldA;
        double2 a = colA[off], b = colB[off];
        if (!direct) {
          if (negA) a = make_double2(-a.x, -a.y);
          else b = make_double2(-b.x, -b.y);
        }
        const bool pk = ODD ? k == L - 1 : (k == 0 || k == L);
        if (pk) {  // the antisymmetric member's pole samples go through the pole table
          if (negA) a = make_double2(0.0, 0.0);
          else b = make_double2(0.0, 0.0);
        }
        v[q] = make_double2(a.x - b.y, a.y + b.x);
        if ((q & 3) == 3) r32_fence();  // loads issued four registers at a time
      }
    }
    {
      // the DFT into the buffer (entry q 64 + lane holds register q), written
      // as the last DFT-32 makes it
      const int lane = loop_lane();
      blue4k_out<-1>(v, Pb, tab, fb, T, Tp, e, lane, [&](int q, double2 z) { T2[q * 64 + lane] = z; });
    }
    double2 z1, zm1;
    {
      // remap: u(r) = F~(m'') phase(m''), m'' = r (r <= L-1) or r - 4096
      // (r >= 4096 - (L-1)), zero otherwise; F~(+-1) for the m' = 0 fold
      const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
      __syncthreads();
      // entry s = r - 4096 + Pb has the parity e ^ (Pb & 1)
      const double2 *S2 = (const double2 *)(ODD ? Tp : T);
      // phase(m'' = 4 l + 2 h + e + 128 j) = phase(4 l + 2 h + e) times block factor j
      const double2 pl = phase[4 * l + 2 * h + e + L - 1];
      const double2 *pblk = phase + 2 * L - 1 + 32;
#pragma unroll
      for (int q = 24; q < 32; ++q) {
        const int r = 4 * (l + 32 * q) + 2 * h + e;
        const bool ok = r >= 4096 - (L - 1);
        const double2 x = S2[at((r - 4096 + Pb) >> 1)];  // r - 4096 + Pb >= Pb - 1024 >= 1
        v[q] = ok ? cmul(x, cmul(pl, pblk[q - 24])) : make_double2(0.0, 0.0);
      }
      // F~(1) in wavefront 1 at i = 0; F~(-1) = F~(Pb - 1) in wavefront (Pb - 1) & 1
      const double2 *B1 = (const double2 *)(e ? T : Tp), *Bm1 = (const double2 *)(((Pb - 1) & 1) == e ? T : Tp);
      z1 = cmul(B1[at(0)], phase[L]);
      zm1 = cmul(Bm1[at((Pb - 1) >> 1)], phase[L - 2]);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int r = 4 * (l + 32 * q) + 2 * h + e;
        v[q] = r <= L - 1 ? cmul(T2[q * 64 + lane], cmul(pl, pblk[q + 8])) : make_double2(0.0, 0.0);
      }
      __syncthreads();  // both wavefronts have read the buffers before the transforms reuse them
#pragma unroll
      for (int q = 8; q < 24; ++q) v[q] = make_double2(0.0, 0.0);
    }
    {
      // convolution with the even weight taps: real spectrum, natural order
      const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
      fft4k_fwd(v, T, Tp, tab, e, lane,
                [&](int q, double2 &z) { z = cscale(z, wre[l + 32 * q + 1024 * h + 2048 * e].x); });
      fft4k_inv<false>(v, T, Tp, tab, e, lane, [](int, double2 &) {});
    }
    // separate and fold, m' = 4 (l + 32 j) + 2 h + e < L: G(m') is register
    // j, G(-m') = G(4096 - m') sits at i' = (4096 - m') >> 1 of this
    // wavefront, registers 24..31 (at (q - 24) 64 + ...) or register 0 (512 + ...)
    const int lane = loop_lane(), l = lane & 31, h = lane >> 5;
    const double2 *colX = X2 + m_slot_of(g, fixm(negA ? mA : mB));
    const double2 p0 = colX[x2_row(g, nslot, ODD ? L - 1 : 0) * g.ldA];
    const double2 p1 = ODD ? make_double2(0.0, 0.0) : colX[x2_row(g, nslot, L) * g.ldA];
    constexpr double sc = 2.0 * kPi / 4096.0;
#pragma unroll
    for (int q = 24; q < 32; ++q) T2[(q - 24) * 64 + lane] = v[q];
    T2[512 + lane] = v[0];
    __builtin_amdgcn_wave_barrier();
    r32_fence();
    const bool okA = mA >= -(L - 1) && mA <= L - 1, okB = mB >= -(L - 1) && mB <= L - 1;
    const int an = n < 0 ? -n : n, sn = n < 0;
    const int cn0 = an == 0 ? 0 : sn, cn1 = an == 0 ? g.NS - 1 : sn;
    int cm0[2], cm1[2];
    double2 *row[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int m = c ? mB : mA;
      const int am = m < 0 ? -m : m, smn = m < 0;
      cm0[c] = am == 0 ? 0 : smn;
      cm1[c] = am == 0 ? 1 : smn;
      row[c] = Gp + ((int64_t)an * L + am) * C2 * L;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int mp = 4 * (l + 32 * j) + 2 * h + e;
      const int ip = (4096 - mp) >> 1;
      const double2 zmv = T2[(((ip >> 6) - 24) << 6) + ((ip & 1) << 5) + ((ip >> 1) & 31)];
      if (mp >= L) continue;
      const double2 zp = make_double2(v[j].x * sc, v[j].y * sc);
      const double2 zm = make_double2(zmv.x * sc, zmv.y * sc);
      double2 GA, GB;
      if (mp > 0) {
        GA = make_double2(zp.x + sA * zm.x, zp.y + sA * zm.y);
        GB = make_double2(zp.y - sA * zm.y, -(zp.x - sA * zm.x));
      } else {
        const double pp = kPi * kPi;
        const double2 dz = make_double2(z1.x - zm1.x, z1.y - zm1.y);
        if (!negA) {
          GA = zp;
          GB = make_double2(pp * dz.x, pp * dz.y);
        } else {
          GA = make_double2(-pp * dz.y, pp * dz.x);
          GB = make_double2(zp.y, -zp.x);
        }
      }
      double2 corr = cmul(p0, pole[mp]);
      if (!ODD) corr = cadd(corr, cmul(p1, pole[L + mp]));
      if (negA) GA = cadd(GA, corr);
      else GB = cadd(GB, corr);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        if (!(c ? okB : okA)) continue;
        const double2 G0 = c ? GB : GA;
#pragma nounroll
        for (int cn = cn0; cn <= cn1; ++cn) {
#pragma nounroll
          for (int cm = cm0[c]; cm <= cm1[c]; ++cm) {
            const double sg = ((cm != cn) && (mp & 1)) ? -1.0 : 1.0;
            row[c][(int64_t)(cm + 2 * cn) * L + mp] = make_double2(sg * G0.x, sg * G0.y);
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    r32_fence();  // the buffer is read before the next group's transforms
  }
}

// ---------------------------------------------------------------- paired beta, LDS engine
// The paired beta stages (see above) for the other DFT lengths: a
// workgroup transforms K pairs, 2K consecutive m from m_lo, in the LDS
// Stockham engine.  Reference counterparts as for k_inv_beta / k_fwd_beta.
__global__ __launch_bounds__(kThreads) void k_inv_beta_gp(Geo g, FftDesc fd, int K, int ld,
                                                          const double2 *__restrict__ X1,
                                                          double2 *__restrict__ X2,
                                                          const int64_t *__restrict__ nbase,
                                                          const double2 *__restrict__ phase) {
  extern __shared__ double2 lds[];
  double2 *A = lds;
  const double2 *tw = fft_load_twiddles(lds + K * ld, fd);
  const int L = g.L, Pb = g.Pb, C = 2 * K;
  const int mlo = -(L - 1) - g.ss;
  const int chunks = (g.nA + C - 1) / C;
  const int nslot = blockIdx.x / chunks;
  const int m0 = mlo + (blockIdx.x % chunks) * C;
  const int n = n_of_slot(g, nslot);
  const bool active = nbase[n + g.N - 1] != kNoBlock;
  for (int idx = threadIdx.x; idx < K * Pb; idx += blockDim.x) {
    const int s = idx / Pb, k = idx - s * Pb;
    const int mA = m0 + 2 * s, mB = mA + 1;
    const bool hA = active && mA >= -(L - 1) && mA <= L - 1;
    const bool hB = active && mB >= -(L - 1) && mB <= L - 1;
    const bool negA = ((mA + n) & 1) != 0;
    const int mp = k <= L - 1 ? k : k - Pb;
    double2 v = make_double2(0.0, 0.0);
    if (mp >= -(L - 1)) {
      const int am = mp < 0 ? -mp : mp;
      double2 a = hA ? X1[((int64_t)nslot * g.nA + m_slot_of(g, mA)) * L + am] : make_double2(0.0, 0.0);
      double2 b = hB ? X1[((int64_t)nslot * g.nA + m_slot_of(g, mB)) * L + am] : make_double2(0.0, 0.0);
      if (mp < 0) {
        if (negA) a = make_double2(-a.x, -a.y);
        else b = make_double2(-b.x, -b.y);
      }
      if (mp == 0) {  // the antisymmetric member vanishes at m' = 0
        if (negA) a = make_double2(0.0, 0.0);
        else b = make_double2(0.0, 0.0);
      }
      v = cmul(make_double2(a.x - b.y, a.y + b.x), phase[k]);
    }
    A[s * ld + k] = v;
  }
  __syncthreads();
  fft_run<+1>(A, K, ld, fd, tw);
  for (int idx = threadIdx.x; idx < K * g.nB; idx += blockDim.x) {
    const int b = idx / K, s = idx - b * K;
    const int mA = m0 + 2 * s, mB = mA + 1;
    const double sA = ((mA + n) & 1) ? -1.0 : 1.0;
    const int bp = g.ss ? (b == 0 ? 0 : Pb - b) : Pb - 1 - b;
    const double2 zb = A[s * ld + b], zp = A[s * ld + bp];
    const double2 su = make_double2(zb.x + sA * zp.x, zb.y + sA * zp.y);
    const double2 df = make_double2(zb.x - sA * zp.x, zb.y - sA * zp.y);
    double2 *row = X2 + x2_row(g, nslot, b) * g.ldA;
    // (the MW_SS Nyquist column m = -L holds no data: exact zeros)
    if (mA <= L - 1)
      row[m_slot_of(g, mA)] = mA >= -(L - 1) ? make_double2(0.5 * su.x, 0.5 * su.y) : make_double2(0.0, 0.0);
    if (mB <= L - 1) row[m_slot_of(g, mB)] = make_double2(0.5 * df.y, -0.5 * df.x);
  }
}

__global__ __launch_bounds__(kThreads) void k_fwd_beta_gp(Geo g, FftDesc fb, FftDesc fw, int K, int ld,
                                                          const double2 *__restrict__ X2,
                                                          double2 *__restrict__ Gp,
                                                          const int64_t *__restrict__ nbase,
                                                          const double2 *__restrict__ phase,
                                                          const double2 *__restrict__ wre,
                                                          const double2 *__restrict__ pole, int Pw) {
  constexpr double kPi = 3.141592653589793238462643383279502884;
  extern __shared__ double2 lds[];
  // one twiddle table serves both transforms when their lengths agree;
  // F~(+-1) of every pair after the twiddles
  double2 *A = lds, *twb = lds + K * ld, *tww = fw.len == fb.len ? twb : twb + tw_count(fb);
  double2 *zpm = (fw.len == fb.len ? twb + tw_count(fb) : tww + tw_count(fw));
  const int L = g.L, Pb = g.Pb, C = 2 * K;
  const int mlo = -(L - 1) - g.ss;
  const int chunks = (g.nA + C - 1) / C;
  const int nslot = blockIdx.x / chunks;
  const int m0 = mlo + (blockIdx.x % chunks) * C;
  const int n = n_of_slot(g, nslot);
  if (nbase[n + g.N - 1] == kNoBlock) return;  // block not transformed
  const double2 *twbp = fft_load_twiddles(twb, fb);
  const double2 *twwp = fw.len == fb.len ? twbp : fft_load_twiddles(tww, fw);
  for (int idx = threadIdx.x; idx < K * Pb; idx += blockDim.x) {
    const int k = idx / K, s = idx - k * K;
    const int mA = m0 + 2 * s, mB = mA + 1;
    const bool hA = mA >= -(L - 1) && mA <= L - 1, hB = mB >= -(L - 1) && mB <= L - 1;
    const bool negA = ((mA + n) & 1) != 0;
    // MW: F(2L-2-b) for b >= L; MW_SS: F(2L-b) for b > L
    const bool direct = g.ss ? k <= L : k < L;
    const int kr = direct ? k : (g.ss ? 2 * L - k : 2 * L - 2 - k);
    const double2 *row = X2 + x2_row(g, nslot, kr) * g.ldA;
    double2 a = hA ? row[m_slot_of(g, mA)] : make_double2(0.0, 0.0);
    double2 b = hB ? row[m_slot_of(g, mB)] : make_double2(0.0, 0.0);
    if (!direct) {
      if (negA) a = make_double2(-a.x, -a.y);
      else b = make_double2(-b.x, -b.y);
    }
    const bool pk = g.ss ? (k == 0 || k == L) : (k == L - 1);
    if (pk) {  // pole samples of the antisymmetric member go through the pole table
      if (negA) a = make_double2(0.0, 0.0);
      else b = make_double2(0.0, 0.0);
    }
    A[s * ld + k] = make_double2(a.x - b.y, a.y + b.x);
  }
  __syncthreads();
  fft_run<-1>(A, K, ld, fb, twbp);
  for (int s = threadIdx.x; s < K; s += blockDim.x) {
    zpm[2 * s] = cmul(A[s * ld + 1], phase[L]);
    zpm[2 * s + 1] = cmul(A[s * ld + Pb - 1], phase[L - 2]);
  }
  // F~(m'') = DFT/Pb * phase, placed at m'' mod Pw (zero elsewhere); the
  // remap runs through registers because it widens the sequence in place
  {
    constexpr int kMaxPer = kFftElems * 2;
    double2 v[kMaxPer];
#pragma unroll
    for (int b = 0; b < kMaxPer; ++b) {
      const int idx = threadIdx.x + b * kThreads;
      v[b] = make_double2(0.0, 0.0);
      if (idx < K * Pw) {
        const int s = idx / Pw, r = idx - s * Pw;
        const int mpp = r <= L - 1 ? r : r - Pw;
        if (mpp >= -(L - 1)) v[b] = cmul(A[s * ld + (mpp < 0 ? mpp + Pb : mpp)], phase[mpp + L - 1]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kMaxPer; ++b) {
      const int idx = threadIdx.x + b * kThreads;
      if (idx < K * Pw) {
        const int s = idx / Pw, r = idx - s * Pw;
        A[s * ld + r] = v[b];
      }
    }
  }
  __syncthreads();
  // convolution with the even weight taps
  fft_run<+1>(A, K, ld, fw, twwp);
  for (int idx = threadIdx.x; idx < K * Pw; idx += blockDim.x) {
    const int s = idx / Pw, r = idx - s * Pw;
    A[s * ld + r] = cmul(A[s * ld + r], wre[r]);
  }
  __syncthreads();
  fft_run<-1>(A, K, ld, fw, twwp);
  const double sc = kTwoPi / (double)Pw;
  const int C2 = 2 * g.NS;
  const int an = n < 0 ? -n : n, sn = n < 0;
  for (int idx = threadIdx.x; idx < K * L; idx += blockDim.x) {
    const int s = idx / L, mp = idx - s * L;
    const int mA = m0 + 2 * s, mB = mA + 1;
    const bool negA = ((mA + n) & 1) != 0;
    const double sA = negA ? -1.0 : 1.0;
    const double2 zp = cscale(A[s * ld + mp], sc), zm = cscale(A[s * ld + (mp == 0 ? 0 : Pw - mp)], sc);
    double2 GA, GB;
    if (mp > 0) {
      GA = make_double2(zp.x + sA * zm.x, zp.y + sA * zm.y);
      GB = make_double2(zp.y - sA * zm.y, -(zp.x - sA * zm.x));
    } else {
      const double pp = kPi * kPi;
      const double2 z1 = zpm[2 * s], zm1 = zpm[2 * s + 1];
      const double2 dz = make_double2(z1.x - zm1.x, z1.y - zm1.y);
      if (!negA) {
        GA = zp;
        GB = make_double2(pp * dz.x, pp * dz.y);
      } else {
        GA = make_double2(-pp * dz.y, pp * dz.x);
        GB = make_double2(zp.y, -zp.x);
      }
    }
    // pole samples of the antisymmetric member, through their responses
    const int mX = negA ? mA : mB;
    if (mX >= -(L - 1) && mX <= L - 1) {
      const double2 *colX = X2 + m_slot_of(g, mX);
      double2 corr = cmul(colX[x2_row(g, nslot, g.ss ? 0 : L - 1) * g.ldA], pole[mp]);
      if (g.ss) corr = cadd(corr, cmul(colX[x2_row(g, nslot, L) * g.ldA], pole[L + mp]));
      if (negA) GA = cadd(GA, corr);
      else GB = cadd(GB, corr);
    }
    for (int e = 0; e < 2; ++e) {
      const int m = e ? mB : mA;
      if (m < -(L - 1) || m > L - 1) continue;  // MW_SS Nyquist column / past the last m
      const double2 G0 = e ? GB : GA;
      const int am = m < 0 ? -m : m, sm = m < 0;
      double2 *dst = Gp + ((int64_t)an * L + am) * C2 * L + mp;
      for (int cm = 0; cm < 2; ++cm) {
        if (cm != sm && am != 0) continue;
        for (int cn = 0; cn < g.NS; ++cn) {
          if (cn != sn && an != 0) continue;
          dst[(int64_t)(cm + 2 * cn) * L] = ((cm != cn) && (mp & 1)) ? make_double2(-G0.x, -G0.y) : G0;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- adjoint of the forward: beta
__global__ __launch_bounds__(kThreads) void k_adjf_beta(Geo g, FftDesc fb, FftDesc fw, int K, int ld,
                                                        const double2 *__restrict__ X1,
                                                        double2 *__restrict__ X2,
                                                        const int64_t *__restrict__ nbase,
                                                        const double2 *__restrict__ phase,
                                                        const double2 *__restrict__ wr, int Pw) {
  extern __shared__ double2 lds[];
  // one twiddle table serves both transforms when their lengths agree
  double2 *A = lds, *twb = lds + K * ld, *tww = fw.len == fb.len ? twb : twb + tw_count(fb);
  const int chunks = (g.nA + K - 1) / K;
  const int nslot = blockIdx.x / chunks;
  const int ms0 = (blockIdx.x % chunks) * K;
  const int n = n_of_slot(g, nslot);
  const bool active = nbase[n + g.N - 1] != kNoBlock;
  const int Pb = g.Pb, L = g.L;
  const double2 *twbp = fft_load_twiddles(twb, fb);
  const double2 *twwp = fw.len == fb.len ? twbp : fft_load_twiddles(tww, fw);
  // G(m') for |m'| <= L-1 at m' mod Pw, F(-m') = (-1)^{m+n} F(m')
  for (int idx = threadIdx.x; idx < K * Pw; idx += blockDim.x) {
    const int s = idx / Pw, r = idx - s * Pw;
    const int mslot = ms0 + s;
    const int mp = r <= L - 1 ? r : r - Pw;
    double2 v = make_double2(0.0, 0.0);
    if (active && mslot < g.nA && mp >= -(L - 1)) {
      const int m = m_of_slot(g, mslot);
      v = X1[((int64_t)nslot * g.nA + mslot) * L + (mp < 0 ? -mp : mp)];
      if (mp < 0 && ((m + n) & 1)) v = make_double2(-v.x, -v.y);
    }
    A[s * ld + r] = v;
  }
  __syncthreads();
  fft_run<+1>(A, K, ld, fw, twwp);
  for (int idx = threadIdx.x; idx < K * Pw; idx += blockDim.x) {
    const int s = idx / Pw, r = idx - s * Pw;
    A[s * ld + r] = cmul(A[s * ld + r], wr[r]);
  }
  __syncthreads();
  fft_run<-1>(A, K, ld, fw, twwp);
  // x 4 pi^2 / Pw, phase exp(+i m' pi/(2L-1)) / Pb, to m' mod Pb (in place
  // through registers: the sequence shrinks from Pw to Pb)
  {
    constexpr int kMaxPer = kFftElems * 2;
    const double sc = 2.0 * kTwoPi * 3.14159265358979323846264338327950288 / (double)Pw;
    double2 v[kMaxPer];
#pragma unroll
    for (int b = 0; b < kMaxPer; ++b) {
      const int idx = threadIdx.x + b * kThreads;
      v[b] = make_double2(0.0, 0.0);
      if (idx < K * Pb) {
        const int s = idx / Pb, k = idx - s * Pb;
        const int mp = k <= L - 1 ? k : k - Pb;
        if (mp >= -(L - 1))
          v[b] = cscale(cmul(A[s * ld + (mp < 0 ? mp + Pw : mp)], phase[mp + L - 1]), sc);
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kMaxPer; ++b) {
      const int idx = threadIdx.x + b * kThreads;
      if (idx < K * Pb) {
        const int s = idx / Pb, k = idx - s * Pb;
        A[s * ld + k] = v[b];
      }
    }
  }
  __syncthreads();
  fft_run<+1>(A, K, ld, fb, twbp);
  // adjoint of the extension: out(b) = E(b) + (-1)^{m+n} E(2L-2-b), b < L-1
  for (int idx = threadIdx.x; idx < K * g.nB; idx += blockDim.x) {
    const int b = idx / K, s = idx - b * K;
    const int mslot = ms0 + s;
    if (mslot >= g.nA) continue;
    const int m = m_of_slot(g, mslot);
    double2 v = A[s * ld + b];
    if (b < L - 1) {
      const double2 e = A[s * ld + (2 * L - 2 - b)];
      v = ((m + n) & 1) ? csub(v, e) : cadd(v, e);
    }
    X2[x2_row(g, nslot, b) * g.ldA + mslot] = v;
  }
}

// ---------------------------------------------------------------- forward zero fill
// via SSHT (direct == 0): active blocks get zeros for l < max(L0,|n|)
//   (padded: from l = 0; compact: from l = |n|), as the per-n SSHT output.
// direct (direct == 1): every stored (l >= |n|) entry the contraction does not
//   write is zero (so3_core.c:1548-1553).
__global__ void k_fwd_zero(Geo g, double2 *__restrict__ flmn, const int64_t *__restrict__ allbase,
                           const int64_t *__restrict__ nbase, int direct, int compact) {
  const int nidx = blockIdx.y;  // n + N - 1 (complex) or n (real)
  const int n = g.real ? nidx : nidx - (g.N - 1);
  const int an = n < 0 ? -n : n;
  const int l = blockIdx.x;
  if (l >= g.L) return;
  const bool active = nbase[n + g.N - 1] != kNoBlock;
  const int64_t base = allbase[n + g.N - 1];
  bool zero;
  if (!direct) {
    if (!active) return;
    const int l0e = max(g.L0, an);
    if (compact && l < an) return;
    zero = l < l0e;
  } else {
    if (l < an) return;
    zero = !active || l < g.L0 || (g.nmodeL && l != an);
  }
  if (!zero) return;
  for (int m = -l + (int)threadIdx.x; m <= l; m += blockDim.x)
    flmn[base + (int64_t)l * l + l + m] = make_double2(0.0, 0.0);
}

// sequences per workgroup: every pass must fit kFftElems values per thread,
// and the LDS (data + twiddle tables) should leave room for two workgroups
// per CU when it can
// sequences per workgroup for a kernel that runs two descriptors (each must
// fit its own register capacity)
int choose_k2(int ld, const FftDesc &a, const FftDesc &b, int tables, int want);
int pow2_floor(int k);

int choose_k(int ld, int cap, int len, int tables, int want) {
  int k = std::min(want, cap / len);
  const int budget = std::max(lds_limit() / 2, (int)(ld * sizeof(double2)) + tables);
  while (k > 1 && (size_t)k * ld * sizeof(double2) + tables > (size_t)budget) --k;
  if (k < 1 || (size_t)k * ld * sizeof(double2) + tables > (size_t)lds_limit() || len > cap)
    throw HipError("DFT length " + std::to_string(len) + " exceeds the LDS / register budget");
  return k;
}

int pow2_floor(int k) {
  int p = 1;
  while (2 * p <= k) p *= 2;
  return p;
}

int choose_k2(int ld, const FftDesc &a, const FftDesc &b, int tables, int want) {
  const int ka = fft_capacity(a) / a.len, kb = fft_capacity(b) / b.len;
  if (ka < 1 || kb < 1)
    throw HipError("DFT length " + std::to_string(std::max(a.len, b.len)) + " exceeds the register budget");
  const int k = std::min(ka, kb);
  return choose_k(ld, k * std::max(a.len, b.len), std::max(a.len, b.len), tables, want);
}

}  // namespace

int lds_limit() {
  static int lim = [] {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, dev) != hipSuccess || pr.sharedMemPerBlock == 0) return 65536;
    return (int)pr.sharedMemPerBlock;
  }();
  return lim;
}

template <typename Kern>
static void set_lds(Kern k, size_t bytes) {
  if (bytes > 65536) SO3_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// grid of a persistent register-path kernel: the blocks that fit on the
// device at once, a multiple of 8 (xcd_tile), at most `groups`
template <typename Kern>
static unsigned reg_grid(Kern k, int threads, size_t lds, int64_t groups) {
  set_lds(k, lds);
  int dev = 0, cus = 0, per = 0;
  SO3_HIP(hipGetDevice(&dev));
  SO3_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  SO3_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k, threads, lds));
  int64_t n = (int64_t)std::max(per, 1) * std::max(cus, 1);
  if (n > groups) n = groups;
  if (n >= 8) n -= n % 8;
  return (unsigned)std::max<int64_t>(n, 1);
}

template <int MODE>
static void launch_gamma_tile(const Geo &g, const FftDesc &fd, const void *src, void *dst, double scale,
                              hipStream_t st) {
  if (g.real) throw HipError("gamma tile: complex signals only");
  const int64_t S = (int64_t)g.nB * g.nA;
  const size_t sh = reg_shared<kRegWaves>(kRegBlueTab, (int64_t)fd.n * 4);
  const unsigned grid = reg_grid(k_gamma_tile<MODE>, kThreads, sh, (S + kRegWaves - 1) / kRegWaves);
  k_gamma_tile<MODE><<<grid, kThreads, sh, st>>>(g, fd, src, dst, scale);
  SO3_LAUNCH_CHECK();
}

void launch_inv_beta(const Geo &g, const double2 *X1, double2 *X2, NBase nb, const double2 *phase,
                     hipStream_t st) {
  const FftDesc fd = fft_plan(g.Pb);
  if (fd.len == 1024 && fd.blue && g.nB <= 257) {
    constexpr int C = 4 * kRows32Waves;
    const size_t sh = sizeof(double2) * kR32BlueTab + sizeof(double) * kRows32Waves * kR32T;
    const int chunks = (g.nA + C - 1) / C;
    auto go = [&](auto kern) {
      const unsigned grid = reg_grid(kern, 64 * kRows32Waves, sh, (int64_t)g.nG * chunks);
      kern<<<grid, 64 * kRows32Waves, sh, st>>>(g, fd, X1, X2, nb.base, phase);
    };
    (g.Pb & 1) ? go(k_inv_beta_r32<1>) : go(k_inv_beta_r32<0>);
    SO3_LAUNCH_CHECK();
    return;
  }
  if (g.Pb >= 1025 && g.Pb <= 2048 && g.nB <= 1025) {
    const FftDesc f4 = fft_plan_reg(g.Pb);
    constexpr int W = kInvBeta4kWaves, C = W;
    const size_t sh = sizeof(double2) * kR4kTab + sizeof(double) * W * kR32T;
    const int chunks = (g.nA + C - 1) / C;
    auto go = [&](auto kern) {
      const unsigned grid = reg_grid(kern, 64 * W, sh, (int64_t)g.nG * chunks);
      kern<<<grid, 64 * W, sh, st>>>(g, f4, X1, X2, nb.base, phase);
    };
    (g.Pb & 1) ? go(k_inv_beta_r4k<1>) : go(k_inv_beta_r4k<0>);
    SO3_LAUNCH_CHECK();
    return;
  }
  if (g.Pb >= 513 && g.Pb <= 1024 && g.nB <= 513) {
    const FftDesc f2 = fft_plan_reg(g.Pb);
    constexpr int C = 2 * kRows32Waves;
    const size_t sh = sizeof(double2) * kR2kTab + sizeof(double) * kRows32Waves * kR32T;
    const int chunks = (g.nA + C - 1) / C;
    auto go = [&](auto kern) {
      const unsigned grid = reg_grid(kern, 64 * kRows32Waves, sh, (int64_t)g.nG * chunks);
      kern<<<grid, 64 * kRows32Waves, sh, st>>>(g, f2, X1, X2, nb.base, phase);
    };
    (g.Pb & 1) ? go(k_inv_beta_r2k<1>) : go(k_inv_beta_r2k<0>);
    SO3_LAUNCH_CHECK();
    return;
  }
  const int ld = fft_ld(fd);
  const int tb = tw_count(fd) * (int)sizeof(double2);
  const int K = choose_k(ld, fft_capacity(fd), fd.len, tb, 8);
  const size_t sh = (size_t)K * ld * sizeof(double2) + tb;
  set_lds(k_inv_beta_gp, sh);
  const int chunks = (g.nA + 2 * K - 1) / (2 * K);  // K pairs per workgroup
  k_inv_beta_gp<<<g.nG * chunks, kThreads, sh, st>>>(g, fd, K, ld, X1, X2, nb.base, phase);
  SO3_LAUNCH_CHECK();
}

// rows of length n on the register path (fft_plan(n).len == 1024)
static void launch_rows_reg(const FftDesc &fd, const double2 *src, double2 *dst, int64_t nrows, int n,
                            int64_t rs, int dir, double scale, const Geo &g, int map, hipStream_t st) {
  if (fd.blue) {
    const size_t sh = sizeof(double2) * kR32BlueTab + sizeof(double) * kRows32Waves * kR32T;
    auto go = [&](auto kern) {
      const int64_t groups = (nrows + 2 * kRows32Waves - 1) / (2 * kRows32Waves);
      const unsigned grid = reg_grid(kern, 64 * kRows32Waves, sh, groups);
      kern<<<grid, 64 * kRows32Waves, sh, st>>>(fd, src, dst, nrows, n, rs, scale, g, map);
    };
    if (n & 1)
      dir > 0 ? go(k_rows_r32<+1, 1>) : go(k_rows_r32<-1, 1>);
    else
      dir > 0 ? go(k_rows_r32<+1, 0>) : go(k_rows_r32<-1, 0>);
  } else {
    const size_t sh = reg_shared<kRowsWaves>(kRegTw, 0);
    auto go = [&](auto kern) {
      const unsigned grid = reg_grid(kern, 64 * kRowsWaves, sh, (nrows + kRowsWaves - 1) / kRowsWaves);
      kern<<<grid, 64 * kRowsWaves, sh, st>>>(fd, src, dst, nrows, rs, scale, g, map);
    };
    dir > 0 ? go(k_rows_reg<+1>) : go(k_rows_reg<-1>);
  }
  SO3_LAUNCH_CHECK();
}

bool reg_order(const Geo &g) {
  if (g.real) return false;
  const FftDesc fa = fft_plan(g.nA), fg = fft_plan(2 * g.N - 1);
  return fa.len == 1024 && fg.len == 1024 && fg.blue;
}

void launch_gamma_x2(const Geo &g, double2 *X2, int dir, double scale, hipStream_t st) {
  const FftDesc fd = fft_plan(2 * g.N - 1);
  if (!reg_order(g)) throw HipError("gamma in place on X2 needs the register path");
  {
    const size_t sh = sizeof(double2) * kR32BlueTab + sizeof(double) * kRows32Waves * kR32T;
    const int64_t S = (int64_t)g.nB * g.nA;
    auto go = [&](auto kern) {
      const unsigned grid = reg_grid(kern, 64 * kRows32Waves, sh, (S + 2 * kRows32Waves - 1) / (2 * kRows32Waves));
      kern<<<grid, 64 * kRows32Waves, sh, st>>>(g, fd, X2, scale);
    };
    dir > 0 ? go(k_gamma_r32<+1>) : go(k_gamma_r32<-1>);
    SO3_LAUNCH_CHECK();
  }
}

void launch_alpha_x2f(const Geo &g, const double2 *X2, double2 *f, int dir, double scale, hipStream_t st) {
  if (!reg_order(g)) throw HipError("alpha X2 -> f needs the register path");
  launch_rows_reg(fft_plan(g.nA), X2, f, (int64_t)g.nG * g.nB, g.nA, 0, dir, scale, g, kRowsX2F, st);
}

void launch_alpha_fx2(const Geo &g, const double2 *f, double2 *X2, int dir, double scale, hipStream_t st) {
  if (!reg_order(g)) throw HipError("alpha f -> X2 needs the register path");
  launch_rows_reg(fft_plan(g.nA), f, X2, (int64_t)g.nG * g.nB, g.nA, 0, dir, scale, g, kRowsFX2, st);
}

// rows of length n in [513, 1024] on the 2048-point register FFT
static void launch_rows_r2k(const double2 *src, double2 *dst, int64_t nrows, int n, int64_t rs, int dir,
                            double scale, const Geo &g, int map, hipStream_t st) {
  const FftDesc fd = fft_plan_reg(n);
  const size_t sh = sizeof(double2) * kR2kTab + sizeof(double) * kRows32Waves * kR32T;
  auto go = [&](auto kern) {
    const unsigned grid = reg_grid(kern, 64 * kRows32Waves, sh, (nrows + kRows32Waves - 1) / kRows32Waves);
    kern<<<grid, 64 * kRows32Waves, sh, st>>>(fd, src, dst, nrows, n, rs, scale, g, map);
  };
  dir > 0 ? go(k_rows_r2k<+1>) : go(k_rows_r2k<-1>);
  SO3_LAUNCH_CHECK();
}

// rows of length n in [1025, 2048] on the 4096-point register FFT
static void launch_rows_r4k(const double2 *src, double2 *dst, int64_t nrows, int n, int64_t rs, int dir,
                            double scale, const Geo &g, int map, hipStream_t st) {
  const FftDesc fd = fft_plan_reg(n);
  constexpr int W = kRows4kWaves;
  const size_t sh = sizeof(double2) * kR4kTab + sizeof(double) * W * kR32T;
  const int64_t groups = (nrows + W / 2 - 1) / (W / 2);
  auto go = [&](auto kern) {
    const unsigned grid = reg_grid(kern, 64 * W, sh, groups);
    kern<<<grid, 64 * W, sh, st>>>(fd, src, dst, nrows, n, rs, scale, g, map);
  };
  dir > 0 ? go(k_rows_r4k<+1>) : go(k_rows_r4k<-1>);
  SO3_LAUNCH_CHECK();
}

void launch_rows(double2 *X, int64_t nrows, int n, int ldr, int dir, double scale, hipStream_t st) {
  if (nrows <= 0) return;
  if (n >= 513 && n <= 1024) {
    launch_rows_r2k(X, X, nrows, n, ldr, dir, scale, Geo{}, kRowsFlat, st);
    return;
  }
  if (n >= 1025 && n <= 2048) {
    launch_rows_r4k(X, X, nrows, n, ldr, dir, scale, Geo{}, kRowsFlat, st);
    return;
  }
  const FftDesc fd = fft_plan(n);
  if (fd.len == 1024) {
    launch_rows_reg(fd, X, X, nrows, n, ldr, dir, scale, Geo{}, kRowsFlat, st);
    return;
  }
  const int ld = fft_ld(fd);
  const int tb = tw_count(fd) * (int)sizeof(double2);
  const int K = choose_k(ld, fft_capacity(fd), fd.len, tb, 8);
  const size_t sh = (size_t)K * ld * sizeof(double2) + tb;
  const unsigned grid = (unsigned)((nrows + K - 1) / K);
  if (dir > 0) {
    set_lds(k_rows<+1>, sh);
    k_rows<+1><<<grid, kThreads, sh, st>>>(fd, K, ld, X, nrows, n, (int64_t)ldr, scale);
  } else {
    set_lds(k_rows<-1>, sh);
    k_rows<-1><<<grid, kThreads, sh, st>>>(fd, K, ld, X, nrows, n, (int64_t)ldr, scale);
  }
  SO3_LAUNCH_CHECK();
}

void launch_alpha(const Geo &g, double2 *X2, int dir, double scale, hipStream_t st) {
  launch_rows(X2, (int64_t)g.nG * g.nB, g.nA, g.ldA, dir, scale, st);
}

// real gamma on the two-sequence register FFT (four columns per wavefront)
static void launch_gamma_real(const Geo &g, const FftDesc &fd, double2 *X2, double *f, int dir, double scale,
                              hipStream_t st) {
  const size_t sh = sizeof(double2) * kR32BlueTab + sizeof(double) * kRows32Waves * kR32T;
  const int64_t S = (int64_t)g.nB * g.nA;
  auto go = [&](auto kern) {
    const unsigned grid = reg_grid(kern, 64 * kRows32Waves, sh, (S + 4 * kRows32Waves - 1) / (4 * kRows32Waves));
    kern<<<grid, 64 * kRows32Waves, sh, st>>>(g, fd, X2, f, scale);
  };
  dir > 0 ? go(k_gamma_real_r32<+1>) : go(k_gamma_real_r32<-1>);
  SO3_LAUNCH_CHECK();
}

// The default rule for nG = 2N - 1 in [33, 127]: the register FFT for the
// lengths the workgroup engine runs through Bluestein (2-3x faster there),
// the workgroup engine for the smooth ones it runs directly (up to 2x faster
// than Bluestein on the register FFT at L = 256).
int gamma_engine(const Geo &g) {
  const int nG = 2 * g.N - 1;
  if (g.steer || nG < 33 || nG > 127 || (int64_t)g.nB * g.nA >= ((int64_t)1 << 30)) return 0;
  if (gamma_short_applies(g)) return 0;  // the direct DFT
  if (g.geng == 1 || g.geng == 2) return g.geng;
  return fft_plan(nG).blue ? 2 : 1;
}

// gamma for nG = 2N - 1 in [33, 127] on the 256-point register FFT
// (dir = +1: X2 -> f, -1: f -> X2); false when gamma_engine(g) is not it
static bool launch_gamma_r16(const Geo &g, double2 *X2, void *f, int dir, double scale, hipStream_t st) {
  if (gamma_engine(g) != 2) return false;
  const int nG = 2 * g.N - 1;
  const FftDesc fd = fft_plan_r256(nG);
  const size_t sh = sizeof(double2) * kR16Tab + sizeof(double) * kR16Waves * kR16T;
  const int64_t S = (int64_t)g.nB * g.nA;
  if (g.real) {
    auto go = [&](auto kern) {
      const unsigned grid = reg_grid(kern, 64 * kR16Waves, sh, (S + 8 * kR16Waves - 1) / (8 * kR16Waves));
      kern<<<grid, 64 * kR16Waves, sh, st>>>(g, fd, X2, (double *)f, scale);
    };
    dir > 0 ? go(k_gamma_real_r16<+1>) : go(k_gamma_real_r16<-1>);
  } else {
    auto go = [&](auto kern, const double2 *src, double2 *dst) {
      const unsigned grid = reg_grid(kern, 64 * kR16Waves, sh, (S + 4 * kR16Waves - 1) / (4 * kR16Waves));
      kern<<<grid, 64 * kR16Waves, sh, st>>>(g, fd, src, dst, scale);
    };
    dir > 0 ? go(k_gamma_r16<+1>, X2, (double2 *)f) : go(k_gamma_r16<-1>, (const double2 *)f, X2);
  }
  SO3_LAUNCH_CHECK();
  return true;
}

void launch_inv_gamma(const Geo &g, const double2 *X2, void *f, double scale, hipStream_t st) {
  if (launch_gamma_short(g, X2, f, +1, scale, st)) return;
  if (launch_gamma_r16(g, (double2 *)X2, f, +1, scale, st)) return;
  const FftDesc fd = fft_plan(2 * g.N - 1);
  if (fd.len == 1024 && fd.blue && g.real) {
    launch_gamma_real(g, fd, (double2 *)X2, (double *)f, +1, scale, st);
    return;
  }
  if (fd.len == 1024 && fd.blue) {
    launch_gamma_tile<kGxF>(g, fd, X2, f, scale, st);
    return;
  }
  const int ld = fft_ld(fd);
  const int tb = tw_count(fd) * (int)sizeof(double2);
  // a power of two: the kernel's threads keep one column each
  const int K = pow2_floor(choose_k(ld, fft_capacity(fd), fd.len, tb, 64));
  const size_t sh = (size_t)K * ld * sizeof(double2) + tb;
  set_lds(k_inv_gamma, sh);
  const int64_t S = (int64_t)g.nB * g.nA;
  k_inv_gamma<<<(unsigned)((S + K - 1) / K), kThreads, sh, st>>>(
      g, fd, K, ld, X2, g.real ? nullptr : (double2 *)f, g.real ? (double *)f : nullptr, scale);
  SO3_LAUNCH_CHECK();
}

void launch_fwd_gamma(const Geo &g, const void *f, double2 *X2, double scale, hipStream_t st) {
  if (launch_gamma_short(g, f, X2, -1, scale, st)) return;
  if (launch_gamma_r16(g, X2, (void *)f, -1, scale, st)) return;
  const FftDesc fd = fft_plan(2 * g.N - 1);
  if (fd.len == 1024 && fd.blue && g.real) {
    launch_gamma_real(g, fd, X2, (double *)f, -1, scale, st);
    return;
  }
  if (fd.len == 1024 && fd.blue) {
    launch_gamma_tile<kFxG>(g, fd, f, X2, scale, st);
    return;
  }
  const int ld = fft_ld(fd);
  const int tb = tw_count(fd) * (int)sizeof(double2);
  // a power of two: the kernel's threads keep one column each
  const int K = pow2_floor(choose_k(ld, fft_capacity(fd), fd.len, tb, 64));
  const size_t sh = (size_t)K * ld * sizeof(double2) + tb;
  set_lds(k_fwd_gamma, sh);
  const int64_t S = (int64_t)g.nB * g.nA;
  k_fwd_gamma<<<(unsigned)((S + K - 1) / K), kThreads, sh, st>>>(
      g, fd, K, ld, g.real ? nullptr : (const double2 *)f, g.real ? (const double *)f : nullptr, X2, scale);
  SO3_LAUNCH_CHECK();
}

// steerable gamma on the LDS engine (length 2N), complex only
void launch_gamma_steer(const Geo &g, const void *src, void *dst, int dir, double scale, hipStream_t st) {
  if (!g.steer || g.real) throw HipError("steerable gamma needs a complex steerable transform");
  const FftDesc fd = fft_plan_lds(2 * g.N);
  const int ld = fft_ld(fd);
  const int tb = tw_count(fd) * (int)sizeof(double2);
  const int K = choose_k(ld, fft_capacity(fd), fd.len, tb, 64);
  const size_t sh = (size_t)K * ld * sizeof(double2) + tb;
  const int64_t S = (int64_t)g.nB * g.nA;
  const unsigned grid = (unsigned)((S + K - 1) / K);
  if (dir > 0) {
    set_lds(k_inv_gamma_steer, sh);
    k_inv_gamma_steer<<<grid, kThreads, sh, st>>>(g, fd, K, ld, (const double2 *)src, (double2 *)dst, scale);
  } else {
    set_lds(k_fwd_gamma_steer, sh);
    k_fwd_gamma_steer<<<grid, kThreads, sh, st>>>(g, fd, K, ld, (const double2 *)src, (double2 *)dst, scale);
  }
  SO3_LAUNCH_CHECK();
}

void launch_fwd_beta(const Geo &g, const double2 *X2, double2 *Gp, NBase nb, const double2 *phase,
                     const double2 *wr, const double2 *wr_even,
                     const double2 *pole, int Pw, int adjoint, hipStream_t st) {
  if (!adjoint && wr_even && pole && Pw == 1024 && g.L <= 256 && fft_plan(g.Pb).len == 1024 &&
      fft_plan(g.Pb).blue) {
    const FftDesc fb = fft_plan(g.Pb);
    constexpr int C = 4 * kRows32Waves;
    const size_t sh = sizeof(double2) * kR32BlueTab + sizeof(double) * kRows32Waves * kR32T;
    const int chunks = (g.nA + C - 1) / C;
    auto go = [&](auto kern) {
      const unsigned grid = reg_grid(kern, 64 * kRows32Waves, sh, (int64_t)g.nG * chunks);
      kern<<<grid, 64 * kRows32Waves, sh, st>>>(g, fb, X2, Gp, nb.base, phase, wr_even, pole);
    };
    (g.Pb & 1) ? go(k_fwd_beta_r32<1>) : go(k_fwd_beta_r32<0>);
    SO3_LAUNCH_CHECK();
    return;
  }
  if (!adjoint && wr_even && pole && Pw == 4096 && g.L >= 513 && g.L <= 1024 && g.Pb >= 1025 && g.Pb <= 2048) {
    const FftDesc fb = fft_plan_reg(g.Pb);
    constexpr int C = kR4kWaves;
    const size_t sh = sizeof(double2) * kR4kTab + sizeof(double) * kR4kWaves * kR32T;
    const int chunks = (g.nA + C - 1) / C;
    auto go = [&](auto kern) {
      const unsigned grid = reg_grid(kern, 64 * kR4kWaves, sh, (int64_t)g.nG * chunks);
      kern<<<grid, 64 * kR4kWaves, sh, st>>>(g, fb, X2, Gp, nb.base, phase, wr_even, pole);
    };
    (g.Pb & 1) ? go(k_fwd_beta_r4k<1>) : go(k_fwd_beta_r4k<0>);
    SO3_LAUNCH_CHECK();
    return;
  }
  if (!adjoint && wr_even && pole && Pw == 2048 && g.L >= 257 && g.L <= 512 && g.Pb >= 513 && g.Pb <= 1024) {
    const FftDesc fb = fft_plan_reg(g.Pb);
    constexpr int C = 2 * kRows32Waves;
    const size_t sh = sizeof(double2) * kR2kTab + sizeof(double) * kRows32Waves * kR32T;
    const int chunks = (g.nA + C - 1) / C;
    auto go = [&](auto kern) {
      const unsigned grid = reg_grid(kern, 64 * kRows32Waves, sh, (int64_t)g.nG * chunks);
      kern<<<grid, 64 * kRows32Waves, sh, st>>>(g, fb, X2, Gp, nb.base, phase, wr_even, pole);
    };
    (g.Pb & 1) ? go(k_fwd_beta_r2k<1>) : go(k_fwd_beta_r2k<0>);
    SO3_LAUNCH_CHECK();
    return;
  }
  if (Pw == 1024 && g.L <= 256 && fft_plan(g.Pb).len == 1024 && fft_plan(g.Pb).blue) {
    const FftDesc fb = fft_plan(g.Pb), fw = fft_plan(Pw);
    constexpr int W = kBetaW;
    const size_t sh = reg_shared<W>(kRegBlueTab, (int64_t)g.nB * tile_ld<W>());
    const int chunks = (g.nA + W - 1) / W;
    const unsigned grid = reg_grid(k_fwd_beta_reg<W>, 64 * W, sh, (int64_t)g.nG * chunks);
    k_fwd_beta_reg<W><<<grid, 64 * W, sh, st>>>(g, fb, fw, X2, Gp, nb.base, phase, wr, adjoint);
    SO3_LAUNCH_CHECK();
    return;
  }
  // workgroup engine: smooth lengths run directly
  FftDesc fb = fft_plan_lds(g.Pb);
  FftDesc fw = fft_plan_lds(Pw);
  const int ld = std::max(fft_ld(fb), fft_ld(fw));
  int tb = (tw_count(fb) + (fw.len == fb.len ? 0 : tw_count(fw))) * (int)sizeof(double2);
  if (!adjoint && wr_even && pole) {
    // K pairs per workgroup, F~(+-1) of each after the tables
    int K = choose_k2(ld, fb, fw, tb + 16 * (int)sizeof(double2), 8);
    if (K == 1) {
      // one pair per workgroup (L > 256): tables of 1024 points and more
      // are read from global memory rather than staged for a single pair
      // (20% faster at L = 288 MW_SS, 368 MW and 384 MW_SS; 6% slower at
      // L = 320 MW)
      fb.twg = fb.len >= 1024;
      fw.twg = fw.len >= 1024;
      tb = (tw_count(fb) + (fw.len == fb.len ? 0 : tw_count(fw))) * (int)sizeof(double2);
      K = choose_k2(ld, fb, fw, tb + 16 * (int)sizeof(double2), 8);
    }
    const size_t sh = (size_t)K * ld * sizeof(double2) + tb + 2 * K * sizeof(double2);
    set_lds(k_fwd_beta_gp, sh);
    const int chunks = (g.nA + 2 * K - 1) / (2 * K);
    k_fwd_beta_gp<<<g.nG * chunks, kThreads, sh, st>>>(g, fb, fw, K, ld, X2, Gp, nb.base, phase, wr_even, pole,
                                                       Pw);
    SO3_LAUNCH_CHECK();
    return;
  }
  const int K = choose_k2(ld, fb, fw, tb, 8);
  const size_t sh = (size_t)K * ld * sizeof(double2) + tb;
  set_lds(k_fwd_beta, sh);
  const int chunks = (g.nA + K - 1) / K;
  k_fwd_beta<<<g.nG * chunks, kThreads, sh, st>>>(g, fb, fw, K, ld, X2, Gp, nb.base, phase, wr, Pw, adjoint);
  SO3_LAUNCH_CHECK();
}

void launch_adjf_beta(const Geo &g, const double2 *X1, double2 *X2, NBase nb, const double2 *phase,
                      const double2 *wr, int Pw, hipStream_t st) {
  const FftDesc fb = fft_plan(g.Pb);
  const FftDesc fw = fft_plan(Pw);
  const int ld = std::max(fft_ld(fb), fft_ld(fw));
  const int tb = (tw_count(fb) + (fw.len == fb.len ? 0 : tw_count(fw))) * (int)sizeof(double2);
  const int K = choose_k2(ld, fb, fw, tb, 8);
  const size_t sh = (size_t)K * ld * sizeof(double2) + tb;
  set_lds(k_adjf_beta, sh);
  const int chunks = (g.nA + K - 1) / K;
  k_adjf_beta<<<g.nG * chunks, kThreads, sh, st>>>(g, fb, fw, K, ld, X1, X2, nb.base, phase, wr, Pw);
  SO3_LAUNCH_CHECK();
}

void launch_fwd_zero(const Geo &g, double2 *flmn, NBase nb, int direct, int64_t /*flmn_size*/,
                     int storage, int /*order*/, hipStream_t st) {
  (void)storage;
  // nb.base is the activity-masked table; the unmasked one follows it
  const int nn = g.real ? g.N : 2 * g.N - 1;
  dim3 grid(g.L, nn);
  k_fwd_zero<<<grid, 64, 0, st>>>(g, flmn, nb.base + (2 * g.N - 1), nb.base, direct,
                                  storage == 1);
  SO3_LAUNCH_CHECK();
}

}  // namespace so3g
