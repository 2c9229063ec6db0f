// dl_tables.hip -- GPU generation of the Delta^l(pi/2) tables.
//
// Risbo (1996): d^l is grown from d^{l-1} by two spinor half-steps on the
// full (2l+1)^2 plane, in the gather form
//   d'[I][K] = ( sqrt((n-I)(n-K)) q d[I][K] - sqrt(I(n-K)) p d[I-1][K]
//              + sqrt((n-I)K) p d[I][K-1] + sqrt(IK) q d[I-1][K-1] ) / n,
// rows I = l + m', columns K = l + m, p = sin(b/2), q = cos(b/2), b = pi/2.
// Trapani & Navaza (2006): top row Delta^l_{l,c} along diagonals (their
// eqs. 9-10), then every column c independently by the three-term
// recursion in m' from m' = l down to m' = c, the other eighth by
// Delta^l_{c,m'} = (-1)^{m'-c} Delta^l_{m',c}.
// The reference obtains the same tables from SSHT
// (so3_core.c:983-1035, ssht_dl_beta_risbo_eighth_table /
// ssht_dl_halfpi_trapani_eighth_table).
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "dl_tables.hpp"

namespace so3g {

namespace {

struct Sink {
  double *inv;
  const int64_t *inv_off;
  double *fwd;
  const int64_t *fwd_off;
  int L;
};

__device__ inline void put(const Sink &s, int l, int mp, int c, double v) {
  const int row = (l + 64) & ~63;  // P64(l+1)
  s.inv[s.inv_off[l] + (int64_t)c * row + mp] = v;
  const int mx = mp > c ? mp : c;
  s.fwd[s.fwd_off[(int64_t)mp * s.L + c] + (l - (mx & ~63))] = v;
}

__global__ void trapani_top(double *top, int L) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;  // diagonal l - c
  if (d >= L) return;
  double v = 1.0;
  for (int k = 1; k <= d; ++k) v *= -sqrt((2.0 * k - 1.0) / (2.0 * k));
  top[(int64_t)d * L + 0] = v;
  for (int c = 1; d + c < L; ++c) {
    const int l = d + c;
    v = sqrt(l * (2.0 * l - 1.0) / (2.0 * (l + c) * (l + c - 1.0))) * v;
    top[(int64_t)l * L + c] = v;
  }
}

__global__ void trapani_cols(const double *top, Sink s) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int L = s.L;
  if (t >= (int64_t)L * (L + 1) / 2) return;
  // t -> (l, c) with c <= l, l = floor((sqrt(8t+1)-1)/2)
  int l = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((int64_t)(l + 1) * (l + 2) / 2 <= t) ++l;
  while ((int64_t)l * (l + 1) / 2 > t) --l;
  const int c = (int)(t - (int64_t)l * (l + 1) / 2);
  double cur = top[(int64_t)l * L + c], nxt = 0.0;
  put(s, l, l, c, cur);
  if (c != l) put(s, l, c, l, ((l - c) & 1) ? -cur : cur);
  for (int mp = l; mp > c; --mp) {
    const double v = (2.0 * c * cur - sqrt((double)(l - mp) * (l + mp + 1)) * nxt) /
                     sqrt((double)(l + mp) * (l - mp + 1));
    const int r = mp - 1;
    put(s, l, r, c, v);
    if (r != c) put(s, l, c, r, ((r - c) & 1) ? -v : v);
    nxt = cur;
    cur = v;
  }
}

__global__ void risbo_half(const double *d, double *dd, int n) {
  const int n1 = n + 1;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n1 * n1) return;
  const int I = (int)(t / n1), K = (int)(t % n1);
  const double p = 0.70710678118654752440084436210484903928;  // sin(pi/4)
  const double q = p;                                        // cos(pi/4)
  double acc = 0.0;
  if (I < n && K < n) acc += sqrt((double)(n - I) * (n - K)) * q * d[(int64_t)I * n + K];
  if (I > 0 && K < n) acc -= sqrt((double)I * (n - K)) * p * d[(int64_t)(I - 1) * n + K];
  if (I < n && K > 0) acc += sqrt((double)(n - I) * K) * p * d[(int64_t)I * n + K - 1];
  if (I > 0 && K > 0) acc += sqrt((double)I * K) * q * d[(int64_t)(I - 1) * n + K - 1];
  dd[t] = acc / n;
}

__global__ void risbo_scatter(const double *d, int l, Sink s) {
  const int n = 2 * l + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (l + 1) * (l + 1)) return;
  const int mp = t / (l + 1), c = t % (l + 1);
  put(s, l, mp, c, d[(int64_t)(l + mp) * n + (l + c)]);
}

DlTables build(int L, int method) {
  DlTables T;
  T.L = L;
  T.method = method;
  std::vector<int64_t> inv_off(L + 1), fwd_off((size_t)L * L);
  int64_t acc = 0;
  for (int l = 0; l < L; ++l) {
    inv_off[l] = acc;
    acc += (int64_t)(l + 1) * ((l + 64) & ~63);
  }
  inv_off[L] = acc;
  const int64_t ninv = acc;
  acc = 0;
  for (int mp = 0; mp < L; ++mp)
    for (int c = 0; c < L; ++c) {
      fwd_off[(size_t)mp * L + c] = acc;
      const int mx = mp > c ? mp : c;
      acc += L - (mx & ~63);
    }
  const int64_t nfwd = acc;
  SO3_HIP(hipMalloc(&T.inv, sizeof(double) * ninv));
  SO3_HIP(hipMalloc(&T.fwd, sizeof(double) * nfwd));
  SO3_HIP(hipMalloc(&T.inv_off, sizeof(int64_t) * (L + 1)));
  SO3_HIP(hipMalloc(&T.fwd_off, sizeof(int64_t) * L * L));
  SO3_HIP(hipMemset(T.inv, 0, sizeof(double) * ninv));
  SO3_HIP(hipMemset(T.fwd, 0, sizeof(double) * nfwd));
  SO3_HIP(hipMemcpy(T.inv_off, inv_off.data(), sizeof(int64_t) * (L + 1), hipMemcpyHostToDevice));
  SO3_HIP(hipMemcpy(T.fwd_off, fwd_off.data(), sizeof(int64_t) * L * L, hipMemcpyHostToDevice));
  T.bytes = sizeof(double) * (ninv + nfwd) + sizeof(int64_t) * (L + 1 + (size_t)L * L);
  Sink s{T.inv, T.inv_off, T.fwd, T.fwd_off, L};
  hipEvent_t e0, e1;
  SO3_HIP(hipEventCreate(&e0));
  SO3_HIP(hipEventCreate(&e1));
  SO3_HIP(hipEventRecord(e0, 0));
  if (method == 1) {  // Trapani
    double *top = nullptr;
    SO3_HIP(hipMalloc(&top, sizeof(double) * L * L));
    trapani_top<<<(L + 63) / 64, 64>>>(top, L);
    SO3_LAUNCH_CHECK();
    const int64_t ncol = (int64_t)L * (L + 1) / 2;
    trapani_cols<<<(unsigned)((ncol + 255) / 256), 256>>>(top, s);
    SO3_LAUNCH_CHECK();
    SO3_HIP(hipDeviceSynchronize());
    SO3_HIP(hipFree(top));
  } else {  // Risbo
    double *da = nullptr, *db = nullptr;
    const size_t cap = (size_t)(2 * L + 1) * (2 * L + 1);
    SO3_HIP(hipMalloc(&da, sizeof(double) * cap));
    SO3_HIP(hipMalloc(&db, sizeof(double) * cap));
    const double one = 1.0;
    SO3_HIP(hipMemcpy(da, &one, sizeof(double), hipMemcpyHostToDevice));
    int n = 1;
    for (int l = 0; l < L; ++l) {
      if (l > 0) {
        for (int h = 0; h < 2; ++h) {
          const int64_t tot = (int64_t)(n + 1) * (n + 1);
          risbo_half<<<(unsigned)((tot + 255) / 256), 256>>>(da, db, n);
          SO3_LAUNCH_CHECK();
          std::swap(da, db);
          ++n;
        }
      }
      const int tot = (l + 1) * (l + 1);
      risbo_scatter<<<(tot + 255) / 256, 256>>>(da, l, s);
      SO3_LAUNCH_CHECK();
    }
    SO3_HIP(hipDeviceSynchronize());
    SO3_HIP(hipFree(da));
    SO3_HIP(hipFree(db));
  }
  SO3_HIP(hipEventRecord(e1, 0));
  SO3_HIP(hipEventSynchronize(e1));
  SO3_HIP(hipEventElapsedTime(&T.build_ms, e0, e1));
  SO3_HIP(hipEventDestroy(e0));
  SO3_HIP(hipEventDestroy(e1));
  return T;
}

std::mutex g_mu;
std::map<std::tuple<int, int, int>, DlTables> g_cache;

}  // namespace

const DlTables &dl_tables(int L, int method) {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_mu);
  auto key = std::make_tuple(dev, L, method);
  auto it = g_cache.find(key);
  if (it != g_cache.end()) return it->second;
  return g_cache.emplace(key, build(L, method)).first->second;
}

void dl_tables_clear() {
  std::lock_guard<std::mutex> g(g_mu);
  for (auto &kv : g_cache) {
    (void)hipFree(kv.second.inv);
    (void)hipFree(kv.second.fwd);
    (void)hipFree(kv.second.inv_off);
    (void)hipFree(kv.second.fwd_off);
  }
  g_cache.clear();
}

}  // namespace so3g
