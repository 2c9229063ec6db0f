// fft_reg2k.hpp -- register-resident 2048-point FFT, one sequence per
// wavefront, built on the two-sequence 1024-point FFT of fft_reg32.hpp.
//
// Lane 32 h + l (h < 2, l < 32) holds register q of the half h:
//   * time domain (layout E):  x[2 (l + 32 q) + h], the even samples in the
//     lower half-wavefront and the odd ones in the upper, so that one access
//     instruction moves 64 consecutive entries;
//   * frequency domain (layout Q):  X[l + 32 q + 1024 h].
// Forward transform (kernel exp(-2 pi i jk / 2048)), decimation in time:
//   E = DFT1024(x_even), O = DFT1024(x_odd)   (fft32 on both halves at once)
//   X[k] = E[k] + w^k O[k],  X[k + 1024] = E[k] - w^k O[k],  w = exp(-2 pi i / 2048).
// Inverse (kernel exp(+2 pi i jk / 2048)), decimation in frequency:
//   A[k] = Z[k] + Z[k + 1024],  B[k] = (Z[k] - Z[k + 1024]) conj(w)^k
//   y[2 j] = DFT1024^+(A)[j],  y[2 j + 1] = DFT1024^+(B)[j].
// The butterflies pair register q of lane (0, l) with register q of lane
// (1, l).  They run in four rounds of kR2kRound = 8 registers through the
// wavefront's transpose buffer: each lane writes its 8 values and reads its
// partner's (lane ^ 32).  (Rounds of 16 kept 64 more VGPRs live and spilled.)  Each lane then computes
//   forward:  sigma own' + partner',  own' = own w^(h k)      (sigma = 1 - 2 h)
//   inverse:  (sigma own + partner) conj(w)^(h k)
// which gives X[k] / A[k] in the lower half and X[k + 1024] / B[k] in the
// upper, with uniform register indices (no per-lane selects); w^(h k) reads
// the factor tables at index 0 (= 1) in the lower half.
//
// Bluestein's algorithm for a length n in [513, 1024] runs on it with inner
// length M = 2048 (2 n - 1 <= 2047): chirp, forward FFT, product with the
// kernel spectrum (read from global memory: 2048 entries do not fit beside
// the transpose buffers), inverse FFT, chirp.
#pragma once
#include "fft_reg32.hpp"

namespace so3g {

// LDS tables (double2): the fft32 twiddles (kR32Tw entries, the 1024-point
// FFT), then w^l (l < 32) and w^(32 q) (q < 32), then the chirp half: 0 and
// then c_j / 32 for 1 <= j <= 512 (c_{n-j} = (-1)^n c_j; the table serves
// j < n <= 1024; j >= n falls on the leading zero).
constexpr int kR2kLo = kR32Tw;
constexpr int kR2kHi = kR32Tw + 32;
constexpr int kR2kCh = kR32Tw + 64;
constexpr int kR2kChN = 513;
constexpr int kR2kTab = kR2kCh + kR2kChN;
constexpr int kR2kRound = 8;  // registers per round of the radix-2 exchange

// copies the LDS tables of a Bluestein descriptor with inner length 2048 (the
// chirp unconjugated; the caller conjugates for DIR > 0); the caller
// synchronises
__device__ __forceinline__ void r2k_stage_blue(double2 *dst, const FftDesc &d) {
  for (int i = threadIdx.x; i < kR32Tw; i += blockDim.x) dst[i] = d.tw[(2 * (i & 31) * ((i >> 5) + 1)) & 2047];
  for (int i = threadIdx.x; i < 32; i += blockDim.x) {
    dst[kR2kLo + i] = d.tw[i];
    dst[kR2kHi + i] = d.tw[32 * i];
  }
  for (int i = threadIdx.x; i < kR2kChN; i += blockDim.x)
    dst[kR2kCh + i] = i >= 1 ? cscale(d.chirp[i], 1.0 / 32.0) : make_double2(0.0, 0.0);
}

// chirp factor of sample j = 2 (l + 32 q) + h (< 1024) of a sequence of
// length n with sign DIR: c_j / 32 for j < n, zero for j >= n.  Register
// q < 8 holds j < 512 (the table), q >= 8 holds j >= 512, which reads
// c_{n-j} (-1)^n (1 <= n - j for j < n; j >= n falls on the leading zero).
template <int DIR>
__device__ __forceinline__ double2 r2k_chirp(const double2 *ch, int n, int j, int q) {
  double2 c;
  if (q < 8) {
    c = ch[j];
    if (q == 0 && j == 0) c = make_double2(1.0 / 32.0, 0.0);
  } else {
    c = ch[max(n - j, 0)];
    if (n & 1) c = make_double2(-c.x, -c.y);
  }
  if (DIR > 0) c.y = -c.y;
  return c;
}

// w^(h k) for k = l + 32 q: the product of the two factor tables, index 0
// (= 1) in the lower half-wavefront
__device__ __forceinline__ double2 r2k_tw(const double2 *tab, int h, int l, int q) {
  const double2 lo = tab[kR2kLo + (h ? l : 0)], hi = tab[kR2kHi + (h ? q : 0)];
  return cmul(lo, hi);
}

// one round of the butterfly exchange over registers q0 .. q0 + kR2kRound - 1:
// the partner's values arrive in p[]
__device__ __forceinline__ void r2k_swap(const double2 *v, double2 *p, double *T, int lane) {
  double2 *T2 = (double2 *)T;
#pragma unroll
  for (int i = 0; i < kR2kRound; ++i) T2[i * 64 + lane] = v[i];
  __builtin_amdgcn_wave_barrier();
  r32_fence();
#pragma unroll
  for (int i = 0; i < kR2kRound; ++i) p[i] = T2[i * 64 + (lane ^ 32)];
  __builtin_amdgcn_wave_barrier();
  r32_fence();
}

// forward 2048-point transform, layout E in, layout Q out; tab: the
// r2k_stage_blue tables.  post(q, X) runs on each output register.
template <class Post>
__device__ __forceinline__ void fft2k_fwd(double2 *v, double *T, const double2 *tab, int lane, Post post) {
  const int h = lane >> 5, l = lane & 31;
  const double sg = h ? -1.0 : 1.0;
  fft32<-1>(v, T, tab, lane, NoPost());
  constexpr int R = kR2kRound;
#pragma unroll
  for (int r = 0; r < 32 / R; ++r) {
    double2 p[R];
#pragma unroll
    for (int i = 0; i < R; ++i) v[R * r + i] = cmul(v[R * r + i], r2k_tw(tab, h, l, R * r + i));
    r32_fence();  // table reads issued one round at a time
    r2k_swap(v + R * r, p, T, lane);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      double2 &z = v[R * r + i];
      z = make_double2(sg * z.x + p[i].x, sg * z.y + p[i].y);
      post(R * r + i, z);
    }
    r32_fence();
  }
}

// inverse (sign +1) 2048-point transform, layout Q in, layout E out,
// unnormalised; registers q >= 16 of the output (samples >= 1024) are not
// computed when HALF (the compiler drops them)
template <bool HALF, class Post>
__device__ __forceinline__ void fft2k_inv(double2 *v, double *T, const double2 *tab, int lane, Post post) {
  const int h = lane >> 5, l = lane & 31;
  const double sg = h ? -1.0 : 1.0;
  constexpr int R = kR2kRound;
#pragma unroll
  for (int r = 0; r < 32 / R; ++r) {
    double2 p[R];
    r2k_swap(v + R * r, p, T, lane);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      double2 &z = v[R * r + i];
      z = cmulc(make_double2(sg * z.x + p[i].x, sg * z.y + p[i].y), r2k_tw(tab, h, l, R * r + i));
    }
    r32_fence();
  }
  fft32<+1>(v, T, tab, lane, [&](int k, double2 &z) {
    if (!HALF || k < 16) post(k, z);
  });
}

// DFT of length n (Bluestein, inner length 2048, 513 <= n <= 1024) with sign
// DIR, unnormalised, of one sequence per wavefront in layout E (registers
// q < 16 hold samples j = 2 (l + 32 q) + h < 1024; registers q >= 16 are
// ignored); samples at or beyond n must be finite on entry and are zero on
// exit.  bh: the kernel spectrum of sign DIR (2048 entries, global memory).
// The result is twice the DFT (the chirp tables carry 1/1024 of the 1/2048
// of the inverse FFT): callers scale by 1/2.
// out(q, X) (blue2k_out) receives output register q < 16 as soon as it is
// final instead (the registers are not written back: for callers that store
// straight to memory, so that the outputs need not stay live).
template <int DIR, bool OUT, class Out>
__device__ __forceinline__ void blue2k_impl(double2 *v, int n, const double2 *tab, const double2 *__restrict__ bh,
                                            double *T, int lane, Out out) {
  const int h = lane >> 5, l = lane & 31;
  const double2 *ch = tab + kR2kCh;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    v[q] = cmul(v[q], r2k_chirp<DIR>(ch, n, 2 * (l + 32 * q) + h, q));
    if ((q & 7) == 7) r32_fence();
  }
#pragma unroll
  for (int q = 16; q < 32; ++q) v[q] = make_double2(0.0, 0.0);
  fft2k_fwd(v, T, tab, lane, [&](int q, double2 &z) { z = cmul(z, bh[l + 32 * q + 1024 * h]); });
  fft2k_inv<true>(v, T, tab, lane, [&](int q, double2 &z) {
    const double2 c = r2k_chirp<DIR>(ch, n, 2 * (l + 32 * q) + h, q);
    if (OUT) out(q, cmul(z, c));
    else z = cmul(z, c);
  });
}
template <int DIR>
__device__ __forceinline__ void blue2k(double2 *v, int n, const double2 *tab, const double2 *__restrict__ bh,
                                       double *T, int lane) {
  blue2k_impl<DIR, false>(v, n, tab, bh, T, lane, [](int, double2) {});
}
template <int DIR, class Out>
__device__ __forceinline__ void blue2k_out(double2 *v, int n, const double2 *tab, const double2 *__restrict__ bh,
                                           double *T, int lane, Out out) {
  blue2k_impl<DIR, true>(v, n, tab, bh, T, lane, out);
}

}  // namespace so3g
