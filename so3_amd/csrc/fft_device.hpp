// fft_device.hpp -- batched complex FP64 DFTs staged in LDS (gfx950).
//
// A workgroup transforms `nseq` sequences that live in LDS at A + s*ld.
// Mixed-radix Stockham passes (radices 2, 3, 4, 5, 7, 8, 11, 13) ping-pong
// between two LDS buffers; lengths with a prime factor above 13 go through
// Bluestein's chirp-z algorithm on a power-of-two inner length M >= 2n-1.
// Convention (FFTW's): DIR = -1 forward  X_k = sum_j x_j exp(-2 pi i jk/n),
//                      DIR = +1 backward X_k = sum_j x_j exp(+2 pi i jk/n),
// both unnormalised.
#pragma once
#include "gpu_common.hpp"

namespace so3g {

constexpr int kMaxPass = 24;

struct FftDesc {
  int n;                 // logical length
  int blue;              // 1: Bluestein
  int len;               // inner transform length (n, or M for Bluestein)
  int npass;             // radix passes of the inner transform
  int rad[kMaxPass];
  const double2 *tw;     // tw[k] = exp(-2 pi i k / len), k < len
  const double2 *chirp;  // exp(-pi i (j^2 mod 2n) / n), j < n
  const double2 *bhat_f; // FFT_len of the forward chirp kernel
  const double2 *bhat_b; // FFT_len of the backward chirp kernel
};

template <int DIR>
__device__ inline double2 twid(const double2 *tw, int k) {
  double2 w = tw[k];
  if (DIR > 0) w.y = -w.y;
  return w;
}

template <int R, int DIR>
__device__ inline void dft_small(double2 *a, const double2 *tw, int len) {
  if constexpr (R == 2) {
    double2 t = a[1];
    a[1] = csub(a[0], t);
    a[0] = cadd(a[0], t);
  } else if constexpr (R == 4) {
    double2 s0 = cadd(a[0], a[2]), d0 = csub(a[0], a[2]);
    double2 s1 = cadd(a[1], a[3]), d1 = csub(a[1], a[3]);
    // d1 * (DIR * i)
    double2 r = DIR < 0 ? make_double2(d1.y, -d1.x) : make_double2(-d1.y, d1.x);
    a[0] = cadd(s0, s1);
    a[2] = csub(s0, s1);
    a[1] = cadd(d0, r);
    a[3] = csub(d0, r);
  } else if constexpr (R == 8) {
    const double h = 0.70710678118654752440084436210484903928;
    double2 e[4] = {a[0], a[2], a[4], a[6]};
    double2 o[4] = {a[1], a[3], a[5], a[7]};
    dft_small<4, DIR>(e, tw, len);
    dft_small<4, DIR>(o, tw, len);
    // o[k] *= exp(DIR * 2 pi i k / 8)
    double2 o1 = o[1], o2 = o[2], o3 = o[3];
    if (DIR < 0) {
      o[1] = make_double2(h * (o1.x + o1.y), h * (o1.y - o1.x));
      o[2] = make_double2(o2.y, -o2.x);
      o[3] = make_double2(h * (-o3.x + o3.y), h * (-o3.y - o3.x));
    } else {
      o[1] = make_double2(h * (o1.x - o1.y), h * (o1.y + o1.x));
      o[2] = make_double2(-o2.y, o2.x);
      o[3] = make_double2(h * (-o3.x - o3.y), h * (-o3.y + o3.x));
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = cadd(e[k], o[k]);
      a[k + 4] = csub(e[k], o[k]);
    }
  } else {
    // generic odd radix: O(R^2) with roots exp(DIR 2 pi i k / R) = tw[k*len/R]
    double2 w[R];
    const int st = len / R;
#pragma unroll
    for (int k = 0; k < R; ++k) w[k] = twid<DIR>(tw, k * st);
    double2 out[R];
#pragma unroll
    for (int p = 0; p < R; ++p) {
      double2 acc = a[0];
#pragma unroll
      for (int q = 1; q < R; ++q) {
        const double2 ww = w[(p * q) % R];
        acc.x += a[q].x * ww.x - a[q].y * ww.y;
        acc.y += a[q].x * ww.y + a[q].y * ww.x;
      }
      out[p] = acc;
    }
#pragma unroll
    for (int p = 0; p < R; ++p) a[p] = out[p];
  }
}

// One Stockham pass of radix R over nseq sequences of length len.
template <int R, int DIR>
__device__ void stockham_pass(const double2 *__restrict__ x, double2 *__restrict__ y, int nseq,
                              int ld, int len, int Ns, const double2 *__restrict__ tw) {
  const int nb = len / R;
  const int tot = nseq * nb;
  const int twstep = len / (Ns * R);
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int s = idx / nb, t = idx - s * nb;
    const int j = t % Ns;
    const double2 *xs = x + s * ld;
    double2 *ys = y + s * ld;
    double2 a[R];
#pragma unroll
    for (int q = 0; q < R; ++q) a[q] = xs[t + q * nb];
    if (Ns > 1) {
#pragma unroll
      for (int q = 1; q < R; ++q) a[q] = cmul(a[q], twid<DIR>(tw, q * j * twstep));
    }
    dft_small<R, DIR>(a, tw, len);
    const int out = (t / Ns) * Ns * R + j;
#pragma unroll
    for (int q = 0; q < R; ++q) ys[out + q * Ns] = a[q];
  }
}

// Inner mixed-radix transform; returns the buffer holding the result.
template <int DIR>
__device__ double2 *fft_inner(double2 *A, double2 *B, int nseq, int ld, const FftDesc &d) {
  int Ns = 1;
  for (int p = 0; p < d.npass; ++p) {
    const int R = d.rad[p];
    switch (R) {
      case 2: stockham_pass<2, DIR>(A, B, nseq, ld, d.len, Ns, d.tw); break;
      case 3: stockham_pass<3, DIR>(A, B, nseq, ld, d.len, Ns, d.tw); break;
      case 4: stockham_pass<4, DIR>(A, B, nseq, ld, d.len, Ns, d.tw); break;
      case 5: stockham_pass<5, DIR>(A, B, nseq, ld, d.len, Ns, d.tw); break;
      case 7: stockham_pass<7, DIR>(A, B, nseq, ld, d.len, Ns, d.tw); break;
      case 8: stockham_pass<8, DIR>(A, B, nseq, ld, d.len, Ns, d.tw); break;
      case 11: stockham_pass<11, DIR>(A, B, nseq, ld, d.len, Ns, d.tw); break;
      default: stockham_pass<13, DIR>(A, B, nseq, ld, d.len, Ns, d.tw); break;
    }
    __syncthreads();
    double2 *t = A;
    A = B;
    B = t;
    Ns *= R;
  }
  return A;
}

// Full DFT of length d.n: input at A (n values per sequence, stride ld),
// B is scratch of the same size; ld >= d.len.  Caller has synchronised
// after writing A.  Returns the buffer (A or B) holding the n outputs.
template <int DIR>
__device__ double2 *fft_run(double2 *A, double2 *B, int nseq, int ld, const FftDesc &d) {
  if (!d.blue) return fft_inner<DIR>(A, B, nseq, ld, d);
  const int n = d.n, M = d.len;
  for (int idx = threadIdx.x; idx < nseq * M; idx += blockDim.x) {
    const int s = idx / M, j = idx - s * M;
    double2 v = make_double2(0.0, 0.0);
    if (j < n) {
      double2 c = d.chirp[j];
      if (DIR > 0) c.y = -c.y;
      v = cmul(A[s * ld + j], c);
    }
    B[s * ld + j] = v;
  }
  __syncthreads();
  double2 *r = fft_inner<-1>(B, A, nseq, ld, d);
  double2 *o = (r == A) ? B : A;
  const double2 *bh = DIR < 0 ? d.bhat_f : d.bhat_b;
  for (int idx = threadIdx.x; idx < nseq * M; idx += blockDim.x) {
    const int s = idx / M, j = idx - s * M;
    r[s * ld + j] = cmul(r[s * ld + j], bh[j]);
  }
  __syncthreads();
  double2 *q = fft_inner<+1>(r, o, nseq, ld, d);
  const double inv = 1.0 / (double)M;
  for (int idx = threadIdx.x; idx < nseq * n; idx += blockDim.x) {
    const int s = idx / n, j = idx - s * n;
    double2 c = d.chirp[j];
    if (DIR > 0) c.y = -c.y;
    q[s * ld + j] = cscale(cmul(q[s * ld + j], c), inv);
  }
  __syncthreads();
  return q;
}

}  // namespace so3g
