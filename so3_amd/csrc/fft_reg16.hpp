// fft_reg16.hpp -- register-resident 256-point FFT, four sequences per
// wavefront, 16 values per lane.
//
// Lane 16 h + l (h < 4, l < 16) holds register q of sequence h at index
// l + 16 q (layout P16, q < 16).  The transform of 256 = 16 x 16,
//   X[k1 + 16 k2] = sum_l w16^{l k2} [ w256^{l k1} sum_q x[l + 16 q] w16^{q k1} ],
// is a DFT-16 over the registers of each lane, a twiddle, one transpose of
// the 16 x 16 block of each sequence through LDS (lane l, register k1 ->
// lane k1, register l) and a second DFT-16 over the registers; the result is
// again in layout P16.  This is the layout and step order of the two-sequence
// 1024-point FFT (fft_reg32.hpp) at a quarter of the size: a lane holds 32
// VGPRs of data instead of 128, so the kernels on it run four wavefronts per
// SIMD.
//
// The transpose moves the real parts and then the imaginary parts through one
// buffer of 16 rows x 66 doubles per wavefront (8432 bytes): row k holds
// register k of the 64 lanes; a lane reads 16 consecutive doubles (16-byte
// reads) of one row.
//
// Bluestein's algorithm for a length n in [33, 128] runs on it with inner
// length M = 256 >= 2 n - 1.
#pragma once
#include "fft_reg32.hpp"

namespace so3g {

constexpr int kR16Ld = 66;                 // row stride (doubles) of the transpose buffer
constexpr int kR16T = 15 * kR16Ld + 64;    // doubles per wavefront
// LDS tables (double2): twiddles tw[(k1 - 1) * 16 + l] = exp(-2 pi i l k1 / 256)
// (k1 = 1..15, l < 16), then the chirp c_j / 16 (j < 128; zero for j >= n)
// with the DFT's sign, then the kernel spectrum (256 entries) of that sign.
// The two chirp products carry the 1/256 of the inverse FFT.
constexpr int kR16Tw = 240;
constexpr int kR16Ch = 128;
constexpr int kR16Bh = 256;
constexpr int kR16Tab = kR16Tw + kR16Ch + kR16Bh;

// register slot of natural output index k after dft_small<16>
__host__ __device__ constexpr int r16_slot(int k) { return 4 * (k & 3) + (k >> 2); }

// lane (h, l), register k  ->  lane (h, k), register l, for each sequence h;
// the input is in the slot order of dft_small<16> (natural k at r16_slot(k)),
// the output in natural order
__device__ __forceinline__ void xpose16(double2 *v, double *T, int lane) {
  const int rd = (lane & 15) * kR16Ld + (lane & 48);
  double t[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k * kR16Ld + lane] = v[r16_slot(k)].x;
  __builtin_amdgcn_wave_barrier();
  r32_fence();
#pragma unroll
  for (int j = 0; j < 16; ++j) t[j] = T[rd + j];
  __builtin_amdgcn_wave_barrier();
  r32_fence();
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k * kR16Ld + lane] = v[r16_slot(k)].y;
  __builtin_amdgcn_wave_barrier();
  r32_fence();
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = make_double2(t[j], T[rd + j]);
  __builtin_amdgcn_wave_barrier();
  r32_fence();
}

// unnormalised 256-point DFT with kernel exp(DIR 2 pi i j k / 256) of the
// four sequences of a wavefront, layout P16 in and out; tab: the twiddle
// table above; post(k, X) runs on output register k once it is final
template <int DIR, class Post>
__device__ __forceinline__ void fft16(double2 *v, double *T, const double2 *tab, int lane, Post post) {
  const int l = lane & 15;
  dft_small<16, DIR>(v);
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) {
    double2 w = tab[(k1 - 1) * 16 + l];
    if (DIR > 0) w.y = -w.y;
    v[r16_slot(k1)] = cmul(v[r16_slot(k1)], w);
    if ((k1 & 3) == 3) r32_fence();  // table reads issued 4 at a time
  }
  xpose16(v, T, lane);
  dft_small<16, DIR>(v);
  double2 t[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) t[k] = v[r16_slot(k)];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    v[k] = t[k];
    post(k, v[k]);
    if ((k & 3) == 3) r32_fence();
  }
}

// copies the LDS tables of a Bluestein descriptor with inner length 256 for
// the DFT sign DIR; the caller synchronises
template <int DIR>
__device__ __forceinline__ void r16_stage_blue(double2 *dst, const FftDesc &d) {
  const double2 *bh = DIR < 0 ? d.bhat_f : d.bhat_b;
  for (int i = threadIdx.x; i < kR16Tw; i += blockDim.x) dst[i] = d.tw[((i & 15) * ((i >> 4) + 1)) & 255];
  for (int i = threadIdx.x; i < kR16Ch; i += blockDim.x) {
    double2 c = i < d.n ? cscale(d.chirp[i], 1.0 / 16.0) : make_double2(0.0, 0.0);
    if (DIR > 0) c.y = -c.y;
    dst[kR16Tw + i] = c;
  }
  for (int i = threadIdx.x; i < kR16Bh; i += blockDim.x) dst[kR16Tw + kR16Ch + i] = bh[i];
}

// DFT of length n (Bluestein, inner length 256, 33 <= n <= 128) with the
// sign of the staged tables, unnormalised, of the four sequences of a
// wavefront in layout P16 (registers q < 8 hold entries j = l + 16 q < 128;
// registers q >= 8 are ignored); entries at or beyond n must be finite on
// entry.  out(q, X) receives output register q < 8 (entry l + 16 q, zero at
// or beyond n) as soon as it is final.
template <class Out>
__device__ __forceinline__ void blue16(double2 *v, const double2 *tab, double *T, int lane, Out out) {
  const int l = lane & 15;
  const double2 *ch = tab + kR16Tw, *bh = tab + kR16Tw + kR16Ch;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    v[q] = cmul(v[q], ch[l + 16 * q]);
    if ((q & 3) == 3) r32_fence();
  }
#pragma unroll
  for (int q = 8; q < 16; ++q) v[q] = make_double2(0.0, 0.0);
  fft16<-1>(v, T, tab, lane, [&](int k, double2 &z) { z = cmul(z, bh[l + 16 * k]); });
  fft16<+1>(v, T, tab, lane, [&](int k, double2 &z) {
    if (k < 8) out(k, cmul(z, ch[l + 16 * k]));
  });
}

}  // namespace so3g
