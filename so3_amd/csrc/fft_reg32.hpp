// fft_reg32.hpp -- register-resident 1024-point FFT, two sequences per
// wavefront, 32 values per lane.
//
// Lane 32 h + l (h < 2, l < 32) holds register q of sequence h at index
// l + 32 q (layout P, q < 32).  The transform of 1024 = 32 x 32,
//   X[k1 + 32 k2] = sum_l w32^{l k2} [ w1024^{l k1} sum_q x[l + 32 q] w32^{q k1} ],
// is a DFT-32 over the registers of each lane, a twiddle, one transpose of
// the 32 x 32 block of each sequence through LDS (lane l, register k1 ->
// lane k1, register l) and a second DFT-32 over the registers; the result is
// again in layout P.  No step crosses lanes inside a wavefront, so the whole
// transform is arithmetic plus one LDS round trip per value.  Both signs
// share the code (kernel exp(s 2 pi i j k / 1024), unnormalised).
//
// The transpose moves the real parts and then the imaginary parts through one
// buffer of 32 rows x 66 doubles per wavefront (16880 bytes): row k holds
// register k of the 64 lanes.  Writes are 64 consecutive doubles; a read
// takes 16-byte pairs of doubles from rows 66 doubles (132 dwords) apart,
// which lands 8 consecutive lanes on distinct banks.  (An odd stride of 65
// forces 8-byte reads, which measured 15% slower for the whole kernel.)  (A buffer of half
// the size, used by one half-wavefront after the other, needs divergent code,
// in which the compiler keeps the old and the new values of all 32 registers
// live at once.)
#pragma once
#include "fft_device.hpp"

namespace so3g {

constexpr int kR32Ld = 66;                 // row stride (doubles) of the transpose buffer
constexpr int kR32T = 31 * kR32Ld + 64;    // doubles per wavefront (the last row needs 64)
// double2: twiddles tw[(k1 - 1) * 32 + l] = exp(-2 pi i l k1 / 1024), k1 = 1..31
// (row k1 = 0 is all ones and never read)
constexpr int kR32Tw = 992;
// LDS tables of the Bluestein transform (double2): twiddles | chirp half |
// spectrum half.  Both halves rest on exact symmetries: the chirp
// c_j = exp(-pi i j^2 / n) has c_{n-j} = (-1)^n c_j, and the kernel spectrum
// (the DFT of an even sequence) has bhat_k = bhat_{1024-k}.  The chirp half
// holds 0 and then c_j / 32 for 1 <= j <= 256 (a length-1024 Bluestein plan
// has 257 <= n <= 512); the chirp products before and after the convolution
// together carry its 1/1024, exactly.  The spectrum half holds bhat_k for
// k <= 512.  (c_0 = 1 is not stored: entry 0 is the zero that the chirp
// products use past n.)  Halving them leaves room for eight transpose buffers.
constexpr int kR32Ch = 257;
constexpr int kR32Bh = 513;
constexpr int kR32BlueTab = kR32Tw + kR32Ch + kR32Bh;

// z * exp(DIR 2 pi i K / 32) for a constant K
template <int DIR, int K>
__device__ __forceinline__ double2 w32(double2 z) {
  constexpr double C[9] = {1.0,
                           0.980785280403230449126182236134239037,
                           0.923879532511286756128183189396788287,
                           0.831469612302545237078788377617905757,
                           0.707106781186547524400844362104849039,
                           0.555570233019602224742830813948532874,
                           0.382683432365089771728459984030398867,
                           0.195090322016128267848284868477022241,
                           0.0};
  constexpr int k = K & 31;
  if constexpr (k == 0) return z;
  else if constexpr (k == 8) return jrot<DIR>(z);
  else if constexpr (k == 16) return make_double2(-z.x, -z.y);
  else if constexpr (k == 24) return jrot<-DIR>(z);
  else {
    // cos and sin of 2 pi k / 32 from the first octant table
    constexpr int o = k & 7, qd = k >> 3;
    constexpr double c0 = C[o], s0 = C[8 - o];
    constexpr double c = qd == 0 ? c0 : qd == 1 ? -s0 : qd == 2 ? -c0 : s0;
    constexpr double s1 = qd == 0 ? s0 : qd == 1 ? c0 : qd == 2 ? -s0 : -c0;
    constexpr double s = DIR < 0 ? -s1 : s1;
    return make_double2(z.x * c - z.y * s, z.x * s + z.y * c);
  }
}

template <int DIR, int Q>
__device__ __forceinline__ void w32_row(double2 *a) {
  // a[Q + 8 k1] *= w32^{Q k1}, k1 = 1..3
  a[Q + 8] = w32<DIR, Q>(a[Q + 8]);
  a[Q + 16] = w32<DIR, 2 * Q>(a[Q + 16]);
  a[Q + 24] = w32<DIR, 3 * Q>(a[Q + 24]);
}

__device__ __forceinline__ void r32_fence() { asm volatile("" ::: "memory"); }

// DFT-32 over a[0..31], natural order in and out: 32 = 4 x 8,
// X[k1 + 4 k2] = sum_{q1} w8^{q1 k2} w32^{q1 k1} sum_{q2} a[q1 + 8 q2] w4^{q2 k1}.
// post(k, X[k]) runs on the outputs of each DFT-8 as soon as they exist,
// with a fence after each group, so that table loads of a post step (a
// twiddle or spectrum product) are issued 8 at a time rather than all at once.
template <int DIR, class Post>
__device__ __forceinline__ void dft32(double2 *a, Post post) {
#pragma unroll
  for (int q1 = 0; q1 < 8; ++q1) dft4<DIR>(a[q1], a[q1 + 8], a[q1 + 16], a[q1 + 24]);
  w32_row<DIR, 1>(a);
  w32_row<DIR, 2>(a);
  w32_row<DIR, 3>(a);
  w32_row<DIR, 4>(a);
  w32_row<DIR, 5>(a);
  w32_row<DIR, 6>(a);
  w32_row<DIR, 7>(a);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    dft_small<8, DIR>(a + 8 * k1);
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) post(k1 + 4 * k2, a[8 * k1 + k2]);
    r32_fence();
  }
  // a[8 k1 + k2] = X[k1 + 4 k2]
  double2 t[32];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) t[k1 + 4 * k2] = a[8 * k1 + k2];
#pragma unroll
  for (int k = 0; k < 32; ++k) a[k] = t[k];
}

struct NoPost {
  __device__ void operator()(int, double2 &) const {}
};

// lane (h, l), register k  ->  lane (h, k), register l, for each sequence h
__device__ __forceinline__ void xpose32(double2 *v, double *T, int lane) {
  const int rd = (lane & 31) * kR32Ld + (lane & 32);
  double t[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) T[k * kR32Ld + lane] = v[k].x;
  __builtin_amdgcn_wave_barrier();
  r32_fence();
#pragma unroll
  for (int j = 0; j < 32; ++j) t[j] = T[rd + j];
  __builtin_amdgcn_wave_barrier();
  r32_fence();
#pragma unroll
  for (int k = 0; k < 32; ++k) T[k * kR32Ld + lane] = v[k].y;
  __builtin_amdgcn_wave_barrier();
  r32_fence();
#pragma unroll
  for (int j = 0; j < 32; ++j) v[j] = make_double2(t[j], T[rd + j]);
  __builtin_amdgcn_wave_barrier();
  r32_fence();
}

// unnormalised 1024-point DFT with kernel exp(DIR 2 pi i j k / 1024) of the
// two sequences of a wavefront, layout P in and out; tw: the kR32Tw table;
// post(k, X) runs on output register k once it is final (see dft32)
template <int DIR, class Post>
__device__ __forceinline__ void fft32(double2 *v, double *T, const double2 *tw, int lane, Post post) {
  const int l = lane & 31;
  dft32<DIR>(v, [&](int k1, double2 &z) {
    if (k1 == 0) return;
    double2 w = tw[(k1 - 1) * 32 + l];
    if (DIR > 0) w.y = -w.y;
    z = cmul(z, w);
  });
  xpose32(v, T, lane);
  dft32<DIR>(v, post);
}

// copies the LDS tables of a Bluestein descriptor with inner length 1024 and
// sign DIR (layout of kR32BlueTab); the caller synchronises
template <int DIR>
__device__ __forceinline__ void r32_stage_blue(double2 *dst, const FftDesc &d) {
  const double2 *bh = DIR < 0 ? d.bhat_f : d.bhat_b;
  for (int i = threadIdx.x; i < kR32Tw; i += blockDim.x) dst[i] = d.tw[((i & 31) * ((i >> 5) + 1)) & 1023];
  for (int i = threadIdx.x; i < kR32Ch; i += blockDim.x)
    dst[kR32Tw + i] = i >= 1 && i <= 256 ? cscale(d.chirp[i], 1.0 / 32.0) : make_double2(0.0, 0.0);
  for (int i = threadIdx.x; i <= 512; i += blockDim.x) dst[kR32Tw + kR32Ch + i] = bh[i];
}

// chirp factor of entry j = l + 32 q of a sequence of length n with parity
// ODD: c_j / 32 for j < n, zero for j >= n.  The table holds 0 and then
// c_i / 32 at i = 1..256; entries j >= 256 use c_{n-j} (1 <= n - j <= 256),
// and j >= n falls on the leading zero.  c_0 = 1.
template <int DIR, int ODD>
__device__ __forceinline__ double2 r32_chirp(const double2 *ch, int n, int l, int q) {
  double2 c;
  if (q == 0) {
    c = ch[l];
    if (l == 0) c = make_double2(1.0 / 32.0, 0.0);
  } else if (q < 8) {
    c = ch[l + 32 * q];
  } else {
    c = ch[max(n - l - 32 * q, 0)];
    if (ODD) c = make_double2(-c.x, -c.y);
  }
  if (DIR > 0) c.y = -c.y;
  return c;
}

// DFT of length n (Bluestein with inner length 1024, so 257 <= n <= 512,
// n odd if ODD) with sign DIR, unnormalised, for the two sequences of a
// wavefront in layout P (registers q < 16: n <= 512); entries at or beyond n
// must be finite on entry (their zero chirp factors clear them) and are zero
// on exit.  tab: the r32_stage_blue<DIR> tables.
template <int DIR, int ODD>
__device__ __forceinline__ void blue32(double2 *v, int n, const double2 *tab, double *T, int lane) {
  const int l = lane & 31;
  const double2 *ch = tab + kR32Tw, *bh = tab + kR32Tw + kR32Ch;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    v[q] = cmul(v[q], r32_chirp<DIR, ODD>(ch, n, l, q));
    if ((q & 7) == 7) r32_fence();
  }
#pragma unroll
  for (int q = 16; q < 32; ++q) v[q] = make_double2(0.0, 0.0);
  fft32<-1>(v, T, tab, lane, [&](int k, double2 &z) {
    z = cmul(z, k < 16 ? bh[l + 32 * k] : bh[1024 - l - 32 * k]);
  });
  fft32<+1>(v, T, tab, lane, [&](int k, double2 &z) {
    if (k < 16) z = cmul(z, r32_chirp<DIR, ODD>(ch, n, l, k));
  });
}

}  // namespace so3g
