// conv.hip -- harmonic-space convolutions (reference src/c/so3_conv.c).
//
//   so3_conv_s2toso3_harmonic_convolution (so3_conv.c:186-215):
//     h_lmn = f_lm conj(g_ln) 8 pi^2 / (2l+1)  for |m|, |n| <= l, else 0,
//     over every storage slot of h (padded blocks included).
//   so3_conv_harmonic_convolution (so3_conv.c:44-90):
//     h_lmn = sum_k f_lmk conj(g_lnk) over the k of f's n-loop for which
//     (l, n, k) is non-zero in g; outputs outside h's loops are untouched.
// Both are HBM-bound gathers: one thread per output coefficient.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "conv.hpp"
#include "so3/so3_sampling.h"

namespace so3g {

namespace {

constexpr double kEightPi2 = 78.956835208714863153;  // 8 pi^2

__global__ void k_s2toso3(int L, int N, int real, int compact, const int64_t *__restrict__ nbase,
                          const double2 *__restrict__ flm, const double2 *__restrict__ glm,
                          double2 *__restrict__ h) {
  const int nidx = blockIdx.y;  // n + N - 1 (complex), n (real)
  const int n = real ? nidx : nidx - (N - 1);
  const int an = n < 0 ? -n : n;
  const int l = blockIdx.x;
  if (compact && l < an) return;
  const int64_t b = nbase[n + N - 1];
  for (int m = -l + (int)threadIdx.x; m <= l; m += blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (an <= l) {
      const double2 a = flm[(int64_t)l * l + l + m];
      const double2 c = glm[(int64_t)l * l + l + n];
      v = cscale(cmulc(a, c), kEightPi2 / (2 * l + 1));
    }
    h[b + (int64_t)l * l + l + m] = v;
  }
}

// n-loop of a parameter block (so3_sampling.c:677-758)
struct NLoop {
  int start, stop, inc;
};

__device__ inline bool in_loop(int i, NLoop r) {
  return i >= r.start && i <= r.stop && (i - r.start) % r.inc == 0;
}

struct Side {
  int L, L0, N, n_mode;
  NLoop nl;
  const int64_t *base;  // indexed n + N - 1
};

// so3_sampling_is_elmn_non_zero for (el, m, n) under side s
__device__ inline bool nonzero(int el, int m, int n, const Side &s) {
  if (!in_loop(n, s.nl)) return false;
  const int an = n < 0 ? -n : n;
  const int lo = max(s.L0, an);
  const int hi = s.n_mode == SO3_N_MODE_L ? an : s.L - 1;
  if (el < lo || el > hi) return false;
  return m >= -el && m <= el;
}

__global__ void k_harmonic(Side hs, Side fs, Side gs, const double2 *__restrict__ f,
                           const double2 *__restrict__ g, double2 *__restrict__ h) {
  const int n = hs.nl.start + blockIdx.y * hs.nl.inc;
  if (n > hs.nl.stop) return;
  const int an = n < 0 ? -n : n;
  const int el = max(hs.L0, an) + blockIdx.x;
  const int hi = hs.n_mode == SO3_N_MODE_L ? an : hs.L - 1;
  if (el > hi) return;
  for (int m = -el + (int)threadIdx.x; m <= el; m += blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int k = fs.nl.start; k <= fs.nl.stop; k += fs.nl.inc) {
      if (!nonzero(el, n, k, gs)) continue;
      const double2 a = f[fs.base[k + fs.N - 1] + (int64_t)el * el + el + m];
      const double2 c = g[gs.base[k + gs.N - 1] + (int64_t)el * el + el + n];
      const double2 t = cmulc(a, c);
      acc.x += t.x;
      acc.y += t.y;
    }
    h[hs.base[n + hs.N - 1] + (int64_t)el * el + el + m] = acc;
  }
}

// index(l, m, n) = base(n) + l*l + l + m for every n of the parameter block
std::vector<int64_t> bases(const so3_parameters_t *p) {
  const int N = p->N;
  std::vector<int64_t> b(std::max(1, 2 * N - 1), 0);
  for (int n = -(N - 1); n <= N - 1; ++n) {
    if (p->reality && n < 0) continue;
    const int an = std::abs(n);
    int ind = 0;
    if (an < p->L) {
      if (p->reality) so3_sampling_elmn2ind_real(&ind, an, -an, n, p);
      else so3_sampling_elmn2ind(&ind, an, -an, n, p);
      b[n + N - 1] = (int64_t)ind - (int64_t)an * an;
    } else {
      if (p->reality) so3_sampling_elmn2ind_real(&ind, 0, 0, n, p);
      else so3_sampling_elmn2ind(&ind, 0, 0, n, p);
      b[n + N - 1] = ind;
    }
  }
  return b;
}

NLoop nloop(const so3_parameters_t *p) {
  NLoop r{};
  so3_sampling_n_loop_values(&r.start, &r.stop, &r.inc, p);
  return r;
}

int64_t *to_device(const std::vector<int64_t> &v, hipStream_t st) {
  int64_t *d = nullptr;
  SO3_HIP(hipMallocAsync((void **)&d, sizeof(int64_t) * v.size(), st));
  SO3_HIP(hipMemcpyAsync(d, v.data(), sizeof(int64_t) * v.size(), hipMemcpyHostToDevice, st));
  return d;
}

// device copies of bases(p), kept per (device, L, N, reality, storage, order)
std::mutex g_bmu;
std::map<std::tuple<int, int, int, int, int, int>, int64_t *> g_bases;

}  // namespace

const int64_t *cached_bases(const so3_parameters_t *hp) {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, hp->L, hp->N, hp->reality, (int)hp->storage, (int)hp->n_order);
  std::lock_guard<std::mutex> lk(g_bmu);
  auto it = g_bases.find(key);
  if (it != g_bases.end()) return it->second;
  const std::vector<int64_t> b = bases(hp);
  int64_t *d = nullptr;
  SO3_HIP(hipMalloc((void **)&d, sizeof(int64_t) * b.size()));
  SO3_HIP(hipMemcpy(d, b.data(), sizeof(int64_t) * b.size(), hipMemcpyHostToDevice));
  g_bases[key] = d;
  return d;
}

void release_conv_caches() {
  std::lock_guard<std::mutex> lk(g_bmu);
  for (auto &kv : g_bases) (void)hipFree(kv.second);
  g_bases.clear();
}

void s2toso3(const so3_parameters_t *hp, const void *flm, const void *glm, void *hlmn, hipStream_t st) {
  if (hp->storage != SO3_STORAGE_PADDED && hp->storage != SO3_STORAGE_COMPACT)
    throw HipError("Invalid storage method.");
  const int64_t *db = cached_bases(hp);
  const int nn = hp->reality ? hp->N : 2 * hp->N - 1;
  if (nn > 0 && hp->L > 0) {
    k_s2toso3<<<dim3(hp->L, nn), 64, 0, st>>>(hp->L, hp->N, hp->reality ? 1 : 0,
                                                hp->storage == SO3_STORAGE_COMPACT, db,
                                                (const double2 *)flm, (const double2 *)glm,
                                                (double2 *)hlmn);
    SO3_LAUNCH_CHECK();
  }
}

void harmonic_convolution(const so3_parameters_t *hp, const so3_parameters_t *fp,
                          const so3_parameters_t *gp, const void *f, const void *g, void *h,
                          hipStream_t st) {
  const std::vector<int64_t> bh = bases(hp), bf = bases(fp), bg = bases(gp);
  int64_t *dh = to_device(bh, st), *df = to_device(bf, st), *dg = to_device(bg, st);
  Side hs{hp->L, hp->L0, hp->N, (int)hp->n_mode, nloop(hp), dh};
  Side fs{fp->L, fp->L0, fp->N, (int)fp->n_mode, nloop(fp), df};
  Side gs{gp->L, gp->L0, gp->N, (int)gp->n_mode, nloop(gp), dg};
  const int ncount = hs.nl.inc > 0 && hs.nl.stop >= hs.nl.start ? (hs.nl.stop - hs.nl.start) / hs.nl.inc + 1 : 0;
  if (ncount > 0 && hp->L > 0) {
    k_harmonic<<<dim3(hp->L, ncount), 64, 0, st>>>(hs, fs, gs, (const double2 *)f, (const double2 *)g,
                                                     (double2 *)h);
    SO3_LAUNCH_CHECK();
  }
  SO3_HIP(hipStreamSynchronize(st));
  SO3_HIP(hipFreeAsync(dh, st));
  SO3_HIP(hipFreeAsync(df, st));
  SO3_HIP(hipFreeAsync(dg, st));
}

}  // namespace so3g
