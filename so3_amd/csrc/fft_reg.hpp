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
// transpose per transform.  Transpose rows are padded to 68 entries so
// that neither the writes nor the reads conflict in LDS banks.
#pragma once
#include "fft_device.hpp"

namespace so3g {

constexpr int kRegLd = 68;            // LDS row stride (double2) of the transpose buffer
constexpr int kRegLds = 16 * kRegLd;  // double2 per wavefront

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

// F_DIR: layout N -> layout O.  T: this wavefront's transpose buffer;
// tw1[k1 * 64 + l] = exp(-2 pi i l k1 / 1024), tw2[c * 4 + a] = exp(-2 pi i a c / 64).
template <int DIR>
__device__ __forceinline__ void fft_F(double2 *v, double2 *T, const double2 *tw1, const double2 *tw2,
                                      int lane) {
  dft16n<DIR>(v);
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) T[k1 * kRegLd + lane] = cmul(v[k1], tw_load<DIR>(tw1, k1 * 64 + lane));
  __builtin_amdgcn_wave_barrier();
  phase_fence();
  const int k1 = lane >> 2, a = lane & 3;
#pragma unroll
  for (int j2 = 0; j2 < 16; ++j2) v[j2] = T[k1 * kRegLd + 4 * j2 + a];
  __builtin_amdgcn_wave_barrier();
  phase_fence();
  dft16n<DIR>(v);
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
__device__ __forceinline__ void fft_G(double2 *v, double2 *T, const double2 *tw1, const double2 *tw2,
                                      int lane) {
  const int k1 = lane >> 2, a = lane & 3;
  const double sA = (a & 2) ? -1.0 : 1.0, sB = (a & 1) ? -1.0 : 1.0;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    double2 z = quad_step<0xB1>(v[c], sB);
    if (a == 3) z = jrot<DIR>(z);
    z = quad_step<0x4E>(z, sA);
    v[c] = cmul(z, tw_load<DIR>(tw2, c * 4 + a));
  }
  dft16n<DIR>(v);
#pragma unroll
  for (int j2 = 0; j2 < 16; ++j2) T[k1 * kRegLd + 4 * j2 + a] = v[j2];
  __builtin_amdgcn_wave_barrier();
  phase_fence();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = cmul(T[r * kRegLd + lane], tw_load<DIR>(tw1, r * 64 + lane));
  __builtin_amdgcn_wave_barrier();
  phase_fence();
  dft16n<DIR>(v);
}

// DFT of length d.n <= 1024 (d.len == 1024, Bluestein) with sign DIR,
// unnormalised, layout N in and out; entries at or beyond d.n must be zero
// on entry and are undefined on exit.
template <int DIR>
__device__ __forceinline__ void blue_reg(double2 *v, const FftDesc &d, double2 *T, int lane) {
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int k = lane + 64 * q;
    if (k < d.n) v[q] = cmul(v[q], tw_load<DIR>(d.chirp, k));
  }
  phase_fence();
  fft_F<-1>(v, T, d.tws, d.tws + 1024, lane);
  const double2 *bh = DIR < 0 ? d.bhat_f : d.bhat_b;
#pragma unroll
  for (int c = 0; c < 16; ++c) v[c] = cmul(v[c], bh[o_index(lane, c)]);
  phase_fence();
  fft_G<+1>(v, T, d.tws, d.tws + 1024, lane);
  phase_fence();
  constexpr double inv = 1.0 / 1024.0;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int k = lane + 64 * q;
    if (k < d.n) v[q] = cmul(v[q], cscale(tw_load<DIR>(d.chirp, k), inv));
  }
}

}  // namespace so3g
