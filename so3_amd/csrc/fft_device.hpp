// fft_device.hpp -- batched complex FP64 DFTs staged in LDS (gfx950).
//
// A workgroup of kFftThreads threads transforms S sequences that live in one
// LDS buffer at X + s*ld.  Each mixed-radix Stockham pass reads every
// butterfly's inputs into registers, synchronises, and writes the outputs
// back in place, so one buffer suffices.  Twiddles come from an LDS copy of
// exp(-2 pi i k/len).  Radices 2, 3, 4, 5, 7, 8, 11, 13, 16 have unrolled
// kernels; lengths with a larger prime factor go through Bluestein's
// chirp-z algorithm on a power-of-two inner length M >= 2n-1.
// Convention (FFTW's): DIR = -1 forward  X_k = sum_j x_j exp(-2 pi i jk/n),
//                      DIR = +1 backward X_k = sum_j x_j exp(+2 pi i jk/n),
// both unnormalised.
#pragma once
#include "gpu_common.hpp"

namespace so3g {

constexpr int kMaxPass = 24;
constexpr int kFftThreads = 256;
constexpr int kFftElems = 16;  // complex values a thread holds per pass

struct FftDesc {
  int n;                 // logical length
  int blue;              // 1: Bluestein
  int len;               // inner transform length (n, or M for Bluestein)
  int npass;             // radix passes of the inner transform
  int rad[kMaxPass];
  const double2 *tw;     // tw[k] = exp(-2 pi i k / len), k < len (global)
  const double2 *chirp;  // exp(-pi i (j^2 mod 2n) / n), j < n
  const double2 *bhat_f; // FFT_len of the forward chirp kernel
  const double2 *bhat_b; // FFT_len of the backward chirp kernel
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
template <>
struct OddConst<11> {
  static constexpr double c[5] = {0.841253532831181168861811648919367718,
                                  0.415415013001886425529274149229623204,
                                  -0.142314838273285140443792668616369669,
                                  -0.654860733945285064056925072466293553,
                                  -0.959492973614497389890368057066327699};
  static constexpr double s[5] = {0.540640817455597582107635954318691695,
                                  0.90963199535451837141171538307902846,
                                  0.989821441880932732376092037776718787,
                                  0.75574957435425828377403584397234442,
                                  0.281732556841429697711417915346616899};
};
template <>
struct OddConst<13> {
  static constexpr double c[6] = {0.885456025653209895900375522015098879,
                                  0.568064746731155802511807559127516625,
                                  0.120536680255323053349067687452543582,
                                  -0.354604887042535625969637892600018474,
                                  -0.748510748171101098634630599701351384,
                                  -0.970941817426052027156982276293789227};
  static constexpr double s[6] = {0.464723172043768545656015335133104778,
                                  0.822983865893656394579617423439381991,
                                  0.992708874098053992800751649492520179,
                                  0.935016242685414823439784599837830729,
                                  0.66312265824079520237678549266676628,
                                  0.239315664287557767148753726260211895};
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
    // 4 x 4 Cooley-Tukey in registers: X_{k1+4k2} = sum_q w16^{q k1} w4^{q k2} Y_q[k1]
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
    double2 t[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      double2 y0 = a[4 * k1 + 0], y1 = a[4 * k1 + 1], y2 = a[4 * k1 + 2], y3 = a[4 * k1 + 3];
      dft4<DIR>(y0, y1, y2, y3);
      t[k1] = y0;
      t[k1 + 4] = y1;
      t[k1 + 8] = y2;
      t[k1 + 12] = y3;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = t[k];
  } else {
    dft_odd<R, DIR>(a);
  }
}

// Note on dft16 layout: after the first loop a[q + 4*k1] holds Y_q[k1]; the
// twiddle w16^{q k1} is applied to a[q + 4 k1] for q, k1 >= 1 -- the second
// loop reads a[4 k1 + q] as the k1-th group.  For that the first loop must
// have stored Y_q[k1] at index 4*k1 + q, which dft4 on (q, q+4, q+8, q+12)
// does (output k1 lands at q + 4*k1).

template <int DIR>
__device__ __forceinline__ double2 twl(const double2 *tw, int k) {
  double2 w = tw[k];
  if (DIR > 0) w.y = -w.y;
  return w;
}

// One in-place Stockham pass of radix R over S sequences of length len.
template <int R, int DIR>
__device__ void fft_pass(double2 *X, int S, int ld, int len, int Ns, const double2 *tw) {
  constexpr int NB = (kFftElems + R - 1) / R;
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
      for (int q = 0; q < R; ++q) X[ob[b] + q * Ns] = v[b][q];
    }
  }
  __syncthreads();
}

// Mixed-radix transform of length d.len in place; tw = LDS twiddle table.
template <int DIR>
__device__ void fft_inner(double2 *X, int S, int ld, const FftDesc &d, const double2 *tw) {
  int Ns = 1;
  for (int p = 0; p < d.npass; ++p) {
    const int R = d.rad[p];
    switch (R) {
      case 2: fft_pass<2, DIR>(X, S, ld, d.len, Ns, tw); break;
      case 3: fft_pass<3, DIR>(X, S, ld, d.len, Ns, tw); break;
      case 4: fft_pass<4, DIR>(X, S, ld, d.len, Ns, tw); break;
      case 5: fft_pass<5, DIR>(X, S, ld, d.len, Ns, tw); break;
      case 7: fft_pass<7, DIR>(X, S, ld, d.len, Ns, tw); break;
      case 8: fft_pass<8, DIR>(X, S, ld, d.len, Ns, tw); break;
      case 11: fft_pass<11, DIR>(X, S, ld, d.len, Ns, tw); break;
      case 13: fft_pass<13, DIR>(X, S, ld, d.len, Ns, tw); break;
      default: fft_pass<16, DIR>(X, S, ld, d.len, Ns, tw); break;
    }
    Ns *= R;
  }
}

// Copy the twiddle table of d into LDS (call before the first fft_run;
// the caller synchronises).
__device__ inline void fft_load_twiddles(double2 *tw_lds, const FftDesc &d) {
  for (int k = threadIdx.x; k < d.len; k += kFftThreads) tw_lds[k] = d.tw[k];
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
  for (int idx = threadIdx.x; idx < S * M; idx += kFftThreads) {
    const int s = idx / M, j = idx - s * M;
    if (j < n) {
      double2 c = d.chirp[j];
      if (DIR > 0) c.y = -c.y;
      X[s * ld + j] = cmul(X[s * ld + j], c);
    } else {
      X[s * ld + j] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  fft_inner<-1>(X, S, ld, d, tw);
  const double2 *bh = DIR < 0 ? d.bhat_f : d.bhat_b;
  for (int idx = threadIdx.x; idx < S * M; idx += kFftThreads) {
    const int s = idx / M, j = idx - s * M;
    X[s * ld + j] = cmul(X[s * ld + j], bh[j]);
  }
  __syncthreads();
  fft_inner<+1>(X, S, ld, d, tw);
  const double inv = 1.0 / (double)M;
  for (int idx = threadIdx.x; idx < S * n; idx += kFftThreads) {
    const int s = idx / n, j = idx - s * n;
    double2 c = d.chirp[j];
    if (DIR > 0) c.y = -c.y;
    X[s * ld + j] = cscale(cmul(X[s * ld + j], c), inv);
  }
  __syncthreads();
}

}  // namespace so3g
