// fft_reg4k.hpp -- register-resident 4096-point FFT, one sequence per pair
// of wavefronts, built on the 2048-point FFT of fft_reg2k.hpp.
//
// Wavefront e (= wave index & 1) of a pair holds the samples j = 2 i + e,
// i < 2048, in the layout E of fft_reg2k.hpp (lane (h, l), register q:
// i = 2 (l + 32 q) + h, so j = 4 (l + 32 q) + 2 h + e).
// Forward (kernel exp(-2 pi i jk / 4096)), decimation in time:
//   E = DFT2048(x_even) (wavefront 0), O = DFT2048(x_odd) (wavefront 1)
//   X[k] = E[k] + v^k O[k],  X[k + 2048] = E[k] - v^k O[k],  v = exp(-2 pi i / 4096)
// in layout QQ: wavefront e, lane (h, l), register q holds X[k + 2048 e],
// k = l + 32 q + 1024 h.  Inverse (kernel exp(+2 pi i jk / 4096)),
// decimation in frequency:
//   A[k] = Z[k] + Z[k + 2048] (wavefront 0),  B[k] = (Z[k] - Z[k + 2048]) conj(v)^k (1),
//   y[2 i] = DFT2048^+(A)[i],  y[2 i + 1] = DFT2048^+(B)[i].
// The butterflies pair a register of wavefront 0 with the same register and
// lane of wavefront 1; they run in four rounds of kR2kRound = 8 registers:
// each wavefront writes its values to its own transpose buffer, the workgroup
// synchronises, and each reads its partner's (so every wavefront of the
// workgroup must run them together).  (Rounds of 16 kept 64 more VGPRs live
// and spilled.)  The rounds alternate between the two 8 KB halves of the
// buffer, so one barrier per round suffices: a wavefront writes a half again
// only two rounds later, after a barrier that its partner passed once done
// reading it.  The forward transform's last round is followed by the inverse
// transform's first (the other half); the inverse transform ends its rounds
// with one more barrier, because the 2048-point transform after them
// rewrites the whole buffer.  As in fft_reg2k.hpp each lane then
// computes sigma own' + partner' (forward, own' = own v^(e k)) or
// (sigma own + partner) conj(v)^(e k) (inverse), sigma = 1 - 2 e.
//
// Bluestein's algorithm for a length n in [1025, 2048] runs on it with
// inner length M = 4096: chirp (read from global memory), forward FFT,
// product with the kernel spectrum (global memory), inverse FFT, chirp.
#pragma once
#include "fft_reg2k.hpp"

namespace so3g {

// LDS tables (double2): those of fft_reg2k.hpp up to its chirp (the fft32
// twiddles and the 2048-point factors), then v^(l + 1024 h) at 32 h + l and
// v^(32 q) (q < 32).
constexpr int kR4kLo = kR2kCh;
constexpr int kR4kHi = kR2kCh + 64;
constexpr int kR4kTab = kR2kCh + 96;
static_assert(2 * kR2kRound * 64 * sizeof(double2) <= kR32T * sizeof(double),
              "the two halves of the exchange must fit the transpose buffer");

// copies the LDS tables of a Bluestein descriptor with inner length 4096; the
// caller synchronises
__device__ __forceinline__ void r4k_stage_blue(double2 *dst, const FftDesc &d) {
  for (int i = threadIdx.x; i < kR32Tw; i += blockDim.x) dst[i] = d.tw[(4 * (i & 31) * ((i >> 5) + 1)) & 4095];
  for (int i = threadIdx.x; i < 32; i += blockDim.x) {
    dst[kR2kLo + i] = d.tw[2 * i];           // w2048^l
    dst[kR2kHi + i] = d.tw[64 * i];          // w2048^(32 q)
    dst[kR4kHi + i] = d.tw[32 * i];          // v^(32 q)
  }
  for (int i = threadIdx.x; i < 64; i += blockDim.x) dst[kR4kLo + i] = d.tw[(i & 31) + 1024 * (i >> 5)];
}

// v^(e k) for k = l + 32 q + 1024 h: index 0 (= 1) in wavefront 0
__device__ __forceinline__ double2 r4k_tw(const double2 *tab, int e, int h, int l, int q) {
  const double2 lo = tab[kR4kLo + (e ? 32 * h + l : 0)], hi = tab[kR4kHi + (e ? q : 0)];
  return cmul(lo, hi);
}

// one round of the cross-wavefront exchange over registers q0 .. q0 +
// kR2kRound - 1 through half `half` of the buffers: the partner wavefront's
// values arrive in p[].  Tp: the partner's buffer.  Every wavefront of the
// workgroup calls it together (one barrier).
__device__ __forceinline__ void r4k_swap(const double2 *v, double2 *p, double *T, const double *Tp, int lane,
                                         int half) {
  double2 *T2 = (double2 *)T + half * (kR2kRound * 64);
  const double2 *P2 = (const double2 *)Tp + half * (kR2kRound * 64);
#pragma unroll
  for (int i = 0; i < kR2kRound; ++i) T2[i * 64 + lane] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kR2kRound; ++i) p[i] = P2[i * 64 + lane];
}

// forward 4096-point transform of the pair, layout E in, layout QQ out.
// post(q, X) runs on each output register.
template <class Post>
__device__ __forceinline__ void fft4k_fwd(double2 *v, double *T, const double *Tp, const double2 *tab, int e,
                                          int lane, Post post) {
  const int h = lane >> 5, l = lane & 31;
  const double sg = e ? -1.0 : 1.0;
  fft2k_fwd(v, T, tab, lane, NoPost());
  constexpr int R = kR2kRound;
#pragma unroll
  for (int r = 0; r < 32 / R; ++r) {
    double2 p[R];
#pragma unroll
    for (int i = 0; i < R; ++i) v[R * r + i] = cmul(v[R * r + i], r4k_tw(tab, e, h, l, R * r + i));
    r32_fence();
    r4k_swap(v + R * r, p, T, Tp, lane, r & 1);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      double2 &z = v[R * r + i];
      z = make_double2(sg * z.x + p[i].x, sg * z.y + p[i].y);
      post(R * r + i, z);
    }
    r32_fence();
  }
}

// inverse (sign +1) 4096-point transform of the pair, layout QQ in, layout
// E out, unnormalised; with HALF only registers q < 16 (samples j < 2048)
template <bool HALF, class Post>
__device__ __forceinline__ void fft4k_inv(double2 *v, double *T, const double *Tp, const double2 *tab, int e,
                                          int lane, Post post) {
  const int h = lane >> 5, l = lane & 31;
  const double sg = e ? -1.0 : 1.0;
  constexpr int R = kR2kRound;
#pragma unroll
  for (int r = 0; r < 32 / R; ++r) {
    double2 p[R];
    r4k_swap(v + R * r, p, T, Tp, lane, r & 1);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      double2 &z = v[R * r + i];
      z = cmulc(make_double2(sg * z.x + p[i].x, sg * z.y + p[i].y), r4k_tw(tab, e, h, l, R * r + i));
    }
    r32_fence();
  }
  __syncthreads();  // the partner has read this buffer's last half before it is rewritten
  fft2k_inv<HALF>(v, T, tab, lane, post);
}

// chirp factor of sample j of a sequence of length n with sign DIR: c_j / 64
// for j < n (global table), zero for j >= n
template <int DIR>
__device__ __forceinline__ double2 r4k_chirp(const double2 *__restrict__ chirp, int n, int j) {
  double2 c = make_double2(0.0, 0.0);
  if (j < n) c = cscale(chirp[j], 1.0 / 64.0);
  if (DIR > 0) c.y = -c.y;
  return c;
}

// DFT of length n (Bluestein, inner length 4096, 1025 <= n <= 2048) with
// sign DIR, unnormalised, of one sequence per pair of wavefronts in layout E
// (registers q < 16 of wavefront e hold samples j = 4 (l + 32 q) + 2 h + e;
// registers q >= 16 are ignored); samples at or beyond n must be finite on
// entry and are zero on exit.  d: the descriptor (chirp and kernel spectrum
// of sign DIR in global memory).  The result is exactly the DFT (the two
// chirp products carry the 1/4096 of the inverse FFT).
// out(q, X) (blue4k_out) receives output register q < 16 as soon as it is
// final instead (the registers are not written back: for callers that store
// straight to memory).
template <int DIR, bool OUT, class Out>
__device__ __forceinline__ void blue4k_impl(double2 *v, int n, const double2 *tab, const FftDesc &d, double *T,
                                            const double *Tp, int e, int lane, Out out) {
  const int h = lane >> 5, l = lane & 31;
  const double2 *__restrict__ bh = DIR < 0 ? d.bhat_f : d.bhat_b;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    v[q] = cmul(v[q], r4k_chirp<DIR>(d.chirp, n, 4 * (l + 32 * q) + 2 * h + e));
    if ((q & 7) == 7) r32_fence();
  }
#pragma unroll
  for (int q = 16; q < 32; ++q) v[q] = make_double2(0.0, 0.0);
  fft4k_fwd(v, T, Tp, tab, e, lane, [&](int q, double2 &z) { z = cmul(z, bh[l + 32 * q + 1024 * h + 2048 * e]); });
  fft4k_inv<true>(v, T, Tp, tab, e, lane, [&](int q, double2 &z) {
    const double2 c = r4k_chirp<DIR>(d.chirp, n, 4 * (l + 32 * q) + 2 * h + e);
    if (OUT) out(q, cmul(z, c));
    else z = cmul(z, c);
  });
}
// blue4k_out with this lane's chirp factors (cpre[q], samples q < 16) and
// kernel spectrum entries (spre[q], outputs q < 32) already in registers
template <int DIR, class Out>
__device__ __forceinline__ void blue4k_pre_out(double2 *v, const double2 *tab, const double2 *cpre,
                                               const double2 *spre, double *T, const double *Tp, int e, int lane,
                                               Out out) {
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], cpre[q]);
#pragma unroll
  for (int q = 16; q < 32; ++q) v[q] = make_double2(0.0, 0.0);
  fft4k_fwd(v, T, Tp, tab, e, lane, [&](int q, double2 &z) { z = cmul(z, spre[q]); });
  fft4k_inv<true>(v, T, Tp, tab, e, lane, [&](int q, double2 &z) { out(q, cmul(z, cpre[q])); });
}
// blue4k with the tables in registers, as blue4k_pre_out
template <int DIR>
__device__ __forceinline__ void blue4k_pre(double2 *v, const double2 *tab, const double2 *cpre, const double2 *spre,
                                           double *T, const double *Tp, int e, int lane) {
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], cpre[q]);
#pragma unroll
  for (int q = 16; q < 32; ++q) v[q] = make_double2(0.0, 0.0);
  fft4k_fwd(v, T, Tp, tab, e, lane, [&](int q, double2 &z) { z = cmul(z, spre[q]); });
  fft4k_inv<true>(v, T, Tp, tab, e, lane, [&](int q, double2 &z) { z = cmul(z, cpre[q]); });
}
template <int DIR>
__device__ __forceinline__ void blue4k(double2 *v, int n, const double2 *tab, const FftDesc &d, double *T,
                                       const double *Tp, int e, int lane) {
  blue4k_impl<DIR, false>(v, n, tab, d, T, Tp, e, lane, [](int, double2) {});
}
template <int DIR, class Out>
__device__ __forceinline__ void blue4k_out(double2 *v, int n, const double2 *tab, const FftDesc &d, double *T,
                                           const double *Tp, int e, int lane, Out out) {
  blue4k_impl<DIR, true>(v, n, tab, d, T, Tp, e, lane, out);
}

}  // namespace so3g
