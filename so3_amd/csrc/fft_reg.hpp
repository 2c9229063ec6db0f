// fft_reg.hpp -- register-resident 1024-point FFT, one sequence per wavefront.
//
// A wavefront holds one sequence of length 1024 in one of two layouts:
//   N (natural):    lane l, register q  <->  index l + 64 q
//   O (scrambled):  lane 4 k1 + a, register c  <->  index k1 + 16 c + 256 r(a),
//                   r(a) = 2-bit reversal of a
// with l < 64, q, k1, c < 16, a < 4.  Two transforms of 1024 = 16 x 16 x 4:
//   F_s : N -> O   DFT-16 over q, twiddle w1024^{s l k1}, LDS transpose,
//                  DFT-16, twiddle w64^{s a c}, radix-4 across each lane quad
//                  (two exchange steps through DPP)
//   G_s : O -> N   the same steps in reverse order (the conjugate-transpose
//                  flow of F_{-s})
// both unnormalised with kernel exp(s 2 pi i j k / 1024).  Bluestein's
// convolution F_{-1} -> kernel product -> G_{+1} and a chain such as
// F_{+1} -> weights -> G_{-1} therefore run on registers with one LDS
// transpose per transform.  A transpose moves the real parts and then the
// imaginary parts through a buffer of 16 rows x 68 doubles (half the LDS of
// a complex one); the padding keeps both the writes and the reads free of
// bank conflicts.  The twiddle tables are read from an LDS copy staged once
// per workgroup (reg_stage_tw): as global loads they were most of the
// vector-memory traffic of these kernels.
#pragma once
#include "fft_device.hpp"

namespace so3g {

constexpr int kRegLd = 68;            // row stride (doubles) of the transpose buffer
constexpr int kRegT = 16 * kRegLd;    // doubles per wavefront (8704 bytes)
constexpr int kRegTw = 1024 + 64;     // double2 of the twiddle tables tw1 | tw2
// double2 of the Bluestein tables: tw1 | tw2 | kernel spectrum / 1024 (1024,
// in register order: entry c * 64 + lane is the value at o_index(lane, c)) |
// chirp (n <= 512 entries: a length-1024 Bluestein plan has 257 <= n <= 512)
constexpr int kRegBlueTab = kRegTw + 1024 + 512;

__device__ __forceinline__ int bitrev2(int a) { return ((a & 1) << 1) | (a >> 1); }

// index of register c of lane `lane` in layout O
__device__ __forceinline__ int o_index(int lane, int c) {
  return (lane >> 2) + 16 * c + 256 * bitrev2(lane & 3);
}

// DFT-16 with natural output order
template <int DIR>
__device__ __forceinline__ void dft16n(double2 *v) {
  dft_small<16, DIR>(v);
  double2 t[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) t[k] = v[fft_out_slot<16>(k)];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = t[k];
}

// value of lane ^ 1 (CTRL 0xB1) or lane ^ 2 (CTRL 0x4E) within the quad
template <int CTRL>
__device__ __forceinline__ double2 quad_swap(double2 z) {
  const long long x = __double_as_longlong(z.x), y = __double_as_longlong(z.y);
  const int x0 = __builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, false);
  const int x1 = __builtin_amdgcn_mov_dpp((int)(x >> 32), CTRL, 0xF, 0xF, false);
  const int y0 = __builtin_amdgcn_mov_dpp((int)y, CTRL, 0xF, 0xF, false);
  const int y1 = __builtin_amdgcn_mov_dpp((int)(y >> 32), CTRL, 0xF, 0xF, false);
  return make_double2(__longlong_as_double(((long long)x1 << 32) | (unsigned)x0),
                      __longlong_as_double(((long long)y1 << 32) | (unsigned)y0));
}

// one exchange step of the quad radix-4: lanes with the bit clear take
// mine + partner, the others partner - mine
template <int CTRL>
__device__ __forceinline__ double2 quad_step(double2 z, double sg) {
  const double2 p = quad_swap<CTRL>(z);
  return make_double2(fma(sg, z.x, p.x), fma(sg, z.y, p.y));
}

// keeps the compiler from hoisting the table loads of later phases (which
// would hold five batches of 16 complex values live at once)
__device__ __forceinline__ void phase_fence() { asm volatile("" ::: "memory"); }

template <int DIR>
__device__ __forceinline__ double2 tw_load(const double2 *t, int i) {
  double2 w = t[i];
  if (DIR > 0) w.y = -w.y;
  return w;
}

// copies the twiddle tables of a length-1024 descriptor (tw1 | tw2) to LDS;
// every thread of the block takes part, the caller synchronises
__device__ __forceinline__ void reg_stage_tw(double2 *dst, const FftDesc &d) {
  for (int i = threadIdx.x; i < kRegTw; i += blockDim.x) dst[i] = d.tws[i];
}

// copies the tables of a Bluestein descriptor with inner length 1024 and
// sign DIR to LDS (layout of kRegBlueTab); the caller synchronises
template <int DIR>
__device__ __forceinline__ void reg_stage_blue(double2 *dst, const FftDesc &d) {
  const double2 *bh = DIR < 0 ? d.bhat_f : d.bhat_b;
  for (int i = threadIdx.x; i < kRegTw; i += blockDim.x) dst[i] = d.tws[i];
  // 1/1024 of the convolution is exact in the spectrum (a power of two)
  for (int i = threadIdx.x; i < 1024; i += blockDim.x)
    dst[kRegTw + i] = cscale(bh[o_index(i & 63, i >> 6)], 1.0 / 1024.0);
  for (int i = threadIdx.x; i < d.n; i += blockDim.x) dst[kRegTw + 1024 + i] = d.chirp[i];
}

// v[k] of lane l goes to row k, column l; lane 4 k1 + a then takes row k1,
// columns 4 j + a as v[j]
__device__ __forceinline__ void xpose_to_quads(double2 *v, double *T, int lane) {
  const int rd = (lane >> 2) * kRegLd + (lane & 3);
  double t[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k * kRegLd + lane] = v[k].x;
  __builtin_amdgcn_wave_barrier();
  phase_fence();
#pragma unroll
  for (int j = 0; j < 16; ++j) t[j] = T[rd + 4 * j];
  __builtin_amdgcn_wave_barrier();
  phase_fence();
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k * kRegLd + lane] = v[k].y;
  __builtin_amdgcn_wave_barrier();
  phase_fence();
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = make_double2(t[j], T[rd + 4 * j]);
  __builtin_amdgcn_wave_barrier();
  phase_fence();
}

// the inverse movement of xpose_to_quads
__device__ __forceinline__ void xpose_from_quads(double2 *v, double *T, int lane) {
  const int wr = (lane >> 2) * kRegLd + (lane & 3);
  double t[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) T[wr + 4 * j] = v[j].x;
  __builtin_amdgcn_wave_barrier();
  phase_fence();
#pragma unroll
  for (int k = 0; k < 16; ++k) t[k] = T[k * kRegLd + lane];
  __builtin_amdgcn_wave_barrier();
  phase_fence();
#pragma unroll
  for (int j = 0; j < 16; ++j) T[wr + 4 * j] = v[j].y;
  __builtin_amdgcn_wave_barrier();
  phase_fence();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = make_double2(t[k], T[k * kRegLd + lane]);
  __builtin_amdgcn_wave_barrier();
  phase_fence();
}

// F_DIR: layout N -> layout O.  T: this wavefront's transpose buffer;
// tw1[k1 * 64 + l] = exp(-2 pi i l k1 / 1024), tw2[c * 4 + a] = exp(-2 pi i a c / 64).
template <int DIR>
__device__ __forceinline__ void fft_F(double2 *v, double *T, const double2 *tw1, const double2 *tw2,
                                      int lane) {
  dft16n<DIR>(v);
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) v[k1] = cmul(v[k1], tw_load<DIR>(tw1, k1 * 64 + lane));
  xpose_to_quads(v, T, lane);
  dft16n<DIR>(v);
  const int a = lane & 3;
  const double sA = (a & 2) ? -1.0 : 1.0, sB = (a & 1) ? -1.0 : 1.0;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    double2 z = cmul(v[c], tw_load<DIR>(tw2, c * 4 + a));
    z = quad_step<0x4E>(z, sA);
    if (a == 3) z = jrot<DIR>(z);
    v[c] = quad_step<0xB1>(z, sB);
  }
  phase_fence();
}

// G_DIR: layout O -> layout N (tables as for fft_F)
template <int DIR>
__device__ __forceinline__ void fft_G(double2 *v, double *T, const double2 *tw1, const double2 *tw2,
                                      int lane) {
  const int a = lane & 3;
  const double sA = (a & 2) ? -1.0 : 1.0, sB = (a & 1) ? -1.0 : 1.0;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    double2 z = quad_step<0xB1>(v[c], sB);
    if (a == 3) z = jrot<DIR>(z);
    z = quad_step<0x4E>(z, sA);
    v[c] = cmul(z, tw_load<DIR>(tw2, c * 4 + a));
  }
  dft16n<DIR>(v);
  xpose_from_quads(v, T, lane);
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = cmul(v[r], tw_load<DIR>(tw1, r * 64 + lane));
  phase_fence();
  dft16n<DIR>(v);
}

// DFT of length n (Bluestein with inner length 1024, so 257 <= n <= 512)
// with sign DIR, unnormalised, layout N in and out; entries at or beyond n
// must be zero on entry and are undefined on exit.  tw: the twiddles (LDS),
// chirp: n entries; bh: the kernel spectrum, in register order and divided
// by 1024 (the LDS copy of reg_stage_blue, BHREG = 1) or in natural order
// (FftDesc's, BHREG = 0).  Slots q < 4 (k < 256) are always inside the
// sequence and slots q >= 8 (k >= 512) never, so only q = 4..7 test k < n.
// Slots q >= 8 are set to constant zero before the first transform, so its
// first radix-4 step has two zero inputs; callers that read only slots
// q < 8 afterwards let the compiler drop half of the last radix-4 step;
// LOWREG tests every slot, which serialises the chirp loads and keeps fewer
// registers live (for kernels near the register limit).
template <int DIR, int BHREG, int LOWREG = 0>
__device__ __forceinline__ void blue_reg(double2 *v, int n, const double2 *tw, const double2 *bh,
                                         const double2 *chirp, double *T, int lane) {
#pragma unroll
  for (int q = 0; q < (LOWREG ? 16 : 8); ++q) {
    const int k = lane + 64 * q;
    if ((!LOWREG && q < 4) || k < n) v[q] = cmul(v[q], tw_load<DIR>(chirp, k));
  }
#pragma unroll
  for (int q = 8; q < 16; ++q) v[q] = make_double2(0.0, 0.0);
  phase_fence();
  fft_F<-1>(v, T, tw, tw + 1024, lane);
#pragma unroll
  for (int c = 0; c < 16; ++c)
    v[c] = BHREG ? cmul(v[c], bh[c * 64 + lane]) : cmul(v[c], cscale(bh[o_index(lane, c)], 1.0 / 1024.0));
  phase_fence();
  fft_G<+1>(v, T, tw, tw + 1024, lane);
  phase_fence();
#pragma unroll
  for (int q = 0; q < (LOWREG ? 16 : 8); ++q) {
    const int k = lane + 64 * q;
    if ((!LOWREG && q < 4) || k < n) v[q] = cmul(v[q], tw_load<DIR>(chirp, k));
  }
}

// the same with every table in LDS (reg_stage_blue<DIR> layout)
template <int DIR, int LOWREG = 0>
__device__ __forceinline__ void blue_reg(double2 *v, int n, const double2 *tab, double *T, int lane) {
  blue_reg<DIR, 1, LOWREG>(v, n, tab, tab + kRegTw, tab + kRegTw + 1024, T, lane);
}

}  // namespace so3g
