// fft_device.hpp -- batched complex FP64 DFTs staged in LDS (gfx950).
//
// A workgroup of kFftThreads threads transforms S sequences that live in one
// LDS buffer at X + s*ld.  Each mixed-radix Stockham pass reads every
// butterfly's inputs into registers, synchronises, and writes the outputs
// back in place, so one buffer suffices.  Twiddles come from an LDS copy of
// exp(-2 pi i k/len).  Radices 2, 3, 4, 5, 7, 8, 16 have unrolled kernels;
// lengths with a larger prime factor go through Bluestein's chirp-z
// algorithm on a power-of-two inner length M >= 2n-1.
// Convention (FFTW's): DIR = -1 forward  X_k = sum_j x_j exp(-2 pi i jk/n),
//                      DIR = +1 backward X_k = sum_j x_j exp(+2 pi i jk/n),
// both unnormalised.
#pragma once
#include "gpu_common.hpp"

namespace so3g {

constexpr int kMaxPass = 16;
constexpr int kFftThreads = 256;
constexpr int kFftElems = 8;  // complex values a thread holds per pass

// radix codes (4 bits per pass in FftDesc::radcode)
constexpr int kNumRadix = 7;
constexpr int kRadixOf[kNumRadix] = {2, 3, 4, 5, 7, 8, 16};
// butterflies a thread keeps in registers per pass of radix R
__host__ __device__ constexpr int fft_nb(int R) { return kFftElems / R > 1 ? kFftElems / R : 1; }

struct FftDesc {
  int n;                 // logical length
  int blue;              // 1: Bluestein
  int len;               // inner transform length (n, or M for Bluestein)
  int npass;             // radix passes of the inner transform
  unsigned long long radcode;  // pass p uses kRadixOf[(radcode >> 4p) & 15]
  const double2 *tw;     // tw[k] = exp(-2 pi i k / len), k < len (global)
  const double2 *chirp;  // exp(-pi i (j^2 mod 2n) / n), j < n
  const double2 *bhat_f; // FFT_len of the forward chirp kernel
  const double2 *bhat_b; // FFT_len of the backward chirp kernel
  const double2 *tws;    // len == 1024: twiddles of the register FFT
                         // (fft_reg.hpp), else null
  int twg;               // 1: the workgroup engine reads tw from global
                         // memory (set per launch; see tw_count)
};

// floor(x / d) for 0 <= x < 2^16, 1 <= d <= 2^16 with m = floor(2^32/d) + 1;
// d = 1 (whose m does not fit 32 bits) is encoded as m = 0
__device__ __forceinline__ int fdiv(int x, unsigned m) { return m ? (int)__umulhi((unsigned)x, m) : x; }
__host__ __device__ inline unsigned fdiv_magic(int d) {
  return d == 1 ? 0u : (unsigned)((0x100000000ULL / (unsigned long long)d) + 1ULL);
}

// i * z * DIR
template <int DIR>
__device__ __forceinline__ double2 jrot(double2 z) {
  return DIR < 0 ? make_double2(z.y, -z.x) : make_double2(-z.y, z.x);
}

template <int R>
struct OddConst;
template <>
struct OddConst<3> {
  static constexpr double c[1] = {-0.5};
  static constexpr double s[1] = {0.866025403784438646763723170752936183};
};
template <>
struct OddConst<5> {
  static constexpr double c[2] = {0.309016994374947424102293417182819059,
                                  -0.809016994374947424102293417182819059};
  static constexpr double s[2] = {0.951056516295153572116439333379382143,
                                  0.587785252292473129168705954639072769};
};
template <>
struct OddConst<7> {
  static constexpr double c[3] = {0.623489801858733530525004884004239811,
                                  -0.222520933956314404288902564496794759,
                                  -0.900968867902419126236102319507445051};
  static constexpr double s[3] = {0.78183148246802980870844452667405775,
                                  0.974927912181823607018131682993931217,
                                  0.433883739117558120475768332848358755};
};
// odd prime radix by symmetric pairs:
//   X_k = a_0 + sum_{q=1}^{(R-1)/2} [ (a_q + a_{R-q}) cos(2 pi qk/R)
//                                    + DIR i (a_q - a_{R-q}) sin(2 pi qk/R) ]
template <int R, int DIR>
__device__ __forceinline__ void dft_odd(double2 *a) {
  constexpr int H = (R - 1) / 2;
  double2 sum[H], dif[H];
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    sum[q - 1] = cadd(a[q], a[R - q]);
    dif[q - 1] = csub(a[q], a[R - q]);
  }
  double2 x0 = a[0];
#pragma unroll
  for (int q = 0; q < H; ++q) x0 = cadd(x0, sum[q]);
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    double2 re = a[0], im = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      const int r = (q * k) % R;  // cos(2 pi r/R), sin(2 pi r/R)
      const int rr = r <= H ? r : R - r;
      const double c = OddConst<R>::c[rr - 1];
      const double s = r <= H ? OddConst<R>::s[rr - 1] : -OddConst<R>::s[rr - 1];
      re.x += c * sum[q - 1].x;
      re.y += c * sum[q - 1].y;
      im.x += s * dif[q - 1].x;
      im.y += s * dif[q - 1].y;
    }
    const double2 j = jrot<DIR>(im);
    a[k] = cadd(re, j);
    a[R - k] = csub(re, j);
  }
  a[0] = x0;
}

template <int DIR>
__device__ __forceinline__ void dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3) {
  const double2 s0 = cadd(a0, a2), d0 = csub(a0, a2);
  const double2 s1 = cadd(a1, a3), d1 = jrot<DIR>(csub(a1, a3));
  a0 = cadd(s0, s1);
  a2 = csub(s0, s1);
  a1 = cadd(d0, d1);
  a3 = csub(d0, d1);
}

// z * exp(DIR 2 pi i k / 16) for k = 1..9 (constant k)
template <int DIR, int K>
__device__ __forceinline__ double2 w16(double2 z) {
  constexpr double C1 = 0.923879532511286756128183189396788287;  // cos(pi/8)
  constexpr double S1 = 0.382683432365089771728459984030398867;  // sin(pi/8)
  constexpr double H = 0.707106781186547524400844362104849039;
  double c, s;
  switch (K & 15) {
    case 0: return z;
    case 1: c = C1; s = S1; break;
    case 2: c = H; s = H; break;
    case 3: c = S1; s = C1; break;
    case 4: return jrot<DIR>(z);
    case 6: c = -H; s = H; break;
    case 9: c = -C1; s = -S1; break;
    default: c = 0; s = 0; break;
  }
  s = DIR < 0 ? -s : s;
  return make_double2(z.x * c - z.y * s, z.x * s + z.y * c);
}

template <int R, int DIR>
__device__ __forceinline__ void dft_small(double2 *a) {
  if constexpr (R == 2) {
    const double2 t = a[1];
    a[1] = csub(a[0], t);
    a[0] = cadd(a[0], t);
  } else if constexpr (R == 4) {
    dft4<DIR>(a[0], a[1], a[2], a[3]);
  } else if constexpr (R == 8) {
    constexpr double H = 0.707106781186547524400844362104849039;
    double2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    double2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    // o_k *= exp(DIR 2 pi i k / 8)
    const double sg = DIR < 0 ? -1.0 : 1.0;
    o1 = make_double2(H * (o1.x - sg * o1.y), H * (o1.y + sg * o1.x));
    o2 = jrot<DIR>(o2);
    o3 = make_double2(H * (-o3.x - sg * o3.y), H * (-o3.y + sg * o3.x));
    a[0] = cadd(e0, o0);
    a[4] = csub(e0, o0);
    a[1] = cadd(e1, o1);
    a[5] = csub(e1, o1);
    a[2] = cadd(e2, o2);
    a[6] = csub(e2, o2);
    a[3] = cadd(e3, o3);
    a[7] = csub(e3, o3);
  } else if constexpr (R == 16) {
    // 4 x 4 Cooley-Tukey in registers: X_{k1+4k2} = sum_q w16^{q k1} w4^{q k2} Y_q[k1];
    // the result stays in slot 4*k1 + k2 (see fft_out_slot)
#pragma unroll
    for (int q = 0; q < 4; ++q) dft4<DIR>(a[q], a[q + 4], a[q + 8], a[q + 12]);
    a[5] = w16<DIR, 1>(a[5]);
    a[9] = w16<DIR, 2>(a[9]);
    a[13] = w16<DIR, 3>(a[13]);
    a[6] = w16<DIR, 2>(a[6]);
    a[10] = w16<DIR, 4>(a[10]);
    a[14] = w16<DIR, 6>(a[14]);
    a[7] = w16<DIR, 3>(a[7]);
    a[11] = w16<DIR, 6>(a[11]);
    a[15] = w16<DIR, 9>(a[15]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4<DIR>(a[4 * k1 + 0], a[4 * k1 + 1], a[4 * k1 + 2], a[4 * k1 + 3]);
  } else {
    dft_odd<R, DIR>(a);
  }
}

// Register slot that holds natural output index q after dft_small<R>:
// radix 16 leaves X_{k1 + 4 k2} in slot 4 k1 + k2, the others are in order.
// (The first dft4 loop puts Y_q[k1] in slot q + 4 k1; the twiddled groups
// a[4 k1 .. 4 k1 + 3] then go through the second dft4 in place.)
template <int R>
__host__ __device__ constexpr int fft_out_slot(int q) {
  return R == 16 ? 4 * (q & 3) + (q >> 2) : q;
}

template <int DIR>
__device__ __forceinline__ double2 twl(const double2 *tw, int k) {
  double2 w = tw[k];
  if (DIR > 0) w.y = -w.y;
  return w;
}

// One in-place Stockham pass of radix R over S sequences of length len.
// Kept out of line: a pass is long, and separate functions keep the
// register allocation of one radix from inflating the others.
template <int R, int DIR>
__device__ __attribute__((noinline)) void fft_pass(double2 *X, int S, int ld, int len, int Ns, const double2 *tw) {
  constexpr int NB = fft_nb(R);
  const int nb = len / R;
  const int tot = S * nb;
  const int twstep = len / (Ns * R);
  const unsigned mnb = fdiv_magic(nb), mns = fdiv_magic(Ns);
  double2 v[NB][R];
  int ob[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int idx = threadIdx.x + b * kFftThreads;
    if (idx < tot) {
      const int s = fdiv(idx, mnb), t = idx - s * nb;
      const int tq = fdiv(t, mns), j = t - tq * Ns;
      const double2 *xs = X + s * ld;
#pragma unroll
      for (int q = 0; q < R; ++q) v[b][q] = xs[t + q * nb];
      if (Ns > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) v[b][q] = cmul(v[b][q], twl<DIR>(tw, q * j * twstep));
      }
      dft_small<R, DIR>(v[b]);
      ob[b] = s * ld + tq * Ns * R + j;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int idx = threadIdx.x + b * kFftThreads;
    if (idx < tot) {
#pragma unroll
      for (int q = 0; q < R; ++q) X[ob[b] + q * Ns] = v[b][fft_out_slot<R>(q)];
    }
  }
  __syncthreads();
}

// Mixed-radix transform of length d.len in place; tw = LDS twiddle table.
template <int DIR>
__device__ void fft_inner(double2 *X, int S, int ld, const FftDesc &d, const double2 *tw) {
  int Ns = 1;
  unsigned long long code = d.radcode;
  for (int p = 0; p < d.npass; ++p, code >>= 4) {
    switch ((int)(code & 15)) {
      case 0: fft_pass<2, DIR>(X, S, ld, d.len, Ns, tw); Ns *= 2; break;
      case 1: fft_pass<3, DIR>(X, S, ld, d.len, Ns, tw); Ns *= 3; break;
      case 2: fft_pass<4, DIR>(X, S, ld, d.len, Ns, tw); Ns *= 4; break;
      case 3: fft_pass<5, DIR>(X, S, ld, d.len, Ns, tw); Ns *= 5; break;
      case 4: fft_pass<7, DIR>(X, S, ld, d.len, Ns, tw); Ns *= 7; break;
      case 5: fft_pass<8, DIR>(X, S, ld, d.len, Ns, tw); Ns *= 8; break;
      default: fft_pass<16, DIR>(X, S, ld, d.len, Ns, tw); Ns *= 16; break;
    }
  }
}

// Twiddle tables of lengths from kTwGlobal up are read from global memory,
// where they stay in L1/L2: in LDS (32 KB and more) they left room for one
// workgroup per CU.  A launcher sets twg for shorter tables when a
// workgroup transforms too few sequences to repay staging the table.
// tw_count: the table's LDS entries.
constexpr int kTwGlobal = 2048;
__host__ __device__ inline bool tw_global(const FftDesc &d) { return d.len >= kTwGlobal || d.twg; }
__host__ __device__ inline int tw_count(const FftDesc &d) { return tw_global(d) ? 0 : d.len; }

// Copy the twiddle table of d into LDS, or not (tw_count); returns the table
// fft_run reads.  Call before the first fft_run; the caller synchronises.
__device__ inline const double2 *fft_load_twiddles(double2 *tw_lds, const FftDesc &d) {
  if (tw_global(d)) return d.tw;
  for (int k = threadIdx.x; k < d.len; k += kFftThreads) tw_lds[k] = d.tw[k];
  return tw_lds;
}

// Full DFT of length d.n in place: input at X[s*ld + j], j < d.n, output
// at the same place.  ld >= d.len.  The caller has synchronised after
// writing X and after loading the twiddle table.
template <int DIR>
__device__ void fft_run(double2 *X, int S, int ld, const FftDesc &d, const double2 *tw) {
  if (!d.blue) {
    fft_inner<DIR>(X, S, ld, d, tw);
    return;
  }
  const int n = d.n, M = d.len;
  for (int j = threadIdx.x; j < M; j += kFftThreads) {
    double2 c = make_double2(0.0, 0.0);
    if (j < n) {
      c = d.chirp[j];
      if (DIR > 0) c.y = -c.y;
    }
    for (int s = 0; s < S; ++s) X[s * ld + j] = j < n ? cmul(X[s * ld + j], c) : c;
  }
  __syncthreads();
  fft_inner<-1>(X, S, ld, d, tw);
  const double2 *bh = DIR < 0 ? d.bhat_f : d.bhat_b;
  for (int j = threadIdx.x; j < M; j += kFftThreads) {
    const double2 b = bh[j];
    for (int s = 0; s < S; ++s) X[s * ld + j] = cmul(X[s * ld + j], b);
  }
  __syncthreads();
  fft_inner<+1>(X, S, ld, d, tw);
  const double inv = 1.0 / (double)M;
  for (int j = threadIdx.x; j < n; j += kFftThreads) {
    double2 c = d.chirp[j];
    if (DIR > 0) c.y = -c.y;
    c = cscale(c, inv);
    for (int s = 0; s < S; ++s) X[s * ld + j] = cmul(X[s * ld + j], c);
  }
  __syncthreads();
}

}  // namespace so3g
