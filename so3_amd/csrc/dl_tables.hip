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
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "dl_tables.hpp"

namespace so3g {

namespace {

struct Sink {
  double *tf;
  const int64_t *offf;
  double *ti;
  const int64_t *offi;
  int L, nmt;
};

__host__ __device__ inline int64_t tri(int64_t k) { return k * (k + 1) / 2; }

// Delta^l_{mp, c} into both operand-tile layouts (mp, c <= l)
__device__ inline void put(const Sink &s, int l, int mp, int c, double v) {
  s.tf[s.offf[c] + 256 * (tri(l >> 4) - tri(c >> 4)) + 64 * (mp >> 2) + (mp & 3) * 16 + (l & 15)] = v;
  const int lcmin = max(c, mp & ~15) >> 2;
  s.ti[s.offi[(int64_t)c * s.nmt + (mp >> 4)] + 64 * ((l >> 2) - lcmin) + (l & 3) * 16 + (mp & 15)] = v;
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
  const int nlt = (L + 15) / 16, nlc = (L + 3) / 4;
  T.nmt = nlt;
  std::vector<int64_t> offf(L + 1), offi((size_t)L * nlt);
  int64_t acc = 0;
  for (int c = 0; c < L; ++c) {
    offf[c] = acc;
    acc += 256 * (tri(nlt) - tri(c >> 4));
  }
  offf[L] = acc;
  const int64_t nf = acc;
  acc = 0;
  for (int c = 0; c < L; ++c)
    for (int mt = 0; mt < nlt; ++mt) {
      offi[(size_t)c * nlt + mt] = acc;
      acc += 64 * (nlc - (std::max(c, 16 * mt) >> 2));
    }
  const int64_t ni = acc;
  SO3_HIP(hipMalloc(&T.tf, sizeof(double) * nf));
  SO3_HIP(hipMalloc(&T.ti, sizeof(double) * ni));
  SO3_HIP(hipMalloc(&T.offf, sizeof(int64_t) * offf.size()));
  SO3_HIP(hipMalloc(&T.offi, sizeof(int64_t) * offi.size()));
  SO3_HIP(hipMemset(T.tf, 0, sizeof(double) * nf));
  SO3_HIP(hipMemset(T.ti, 0, sizeof(double) * ni));
  SO3_HIP(hipMemcpy(T.offf, offf.data(), sizeof(int64_t) * offf.size(), hipMemcpyHostToDevice));
  SO3_HIP(hipMemcpy(T.offi, offi.data(), sizeof(int64_t) * offi.size(), hipMemcpyHostToDevice));
  T.bytes = sizeof(double) * (nf + ni) + sizeof(int64_t) * (offf.size() + offi.size());
  Sink s{T.tf, T.offf, T.ti, T.offi, L, nlt};
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
std::map<std::tuple<int, int, int>, unsigned long long> g_used;  // last use, per table
std::map<std::tuple<int, int, int>, int> g_pins;                 // holders that forbid eviction
unsigned long long g_clock = 0;
size_t g_limit = (size_t)32 << 30;  // bytes of tables kept per process
std::map<int, double *> g_zero;

void free_tables(DlTables &t) {
  (void)hipFree(t.tf);
  (void)hipFree(t.ti);
  (void)hipFree(t.offf);
  (void)hipFree(t.offi);
}

// evict least-recently-used tables of other sizes until `incoming` more bytes
// fit the limit; waits for the device first, since queued kernels may still
// read the tables being dropped
void make_room(size_t incoming, const std::tuple<int, int, int> &keep) {
  size_t total = incoming;
  for (auto &kv : g_cache) total += kv.second.bytes;
  if (total <= g_limit) return;
  SO3_HIP(hipDeviceSynchronize());
  while (total > g_limit) {
    auto victim = g_cache.end();
    unsigned long long oldest = ~0ull;
    for (auto it = g_cache.begin(); it != g_cache.end(); ++it) {
      if (it->first == keep || g_pins.count(it->first)) continue;
      const unsigned long long u = g_used[it->first];
      if (u < oldest) oldest = u, victim = it;
    }
    if (victim == g_cache.end()) break;
    total -= victim->second.bytes;
    free_tables(victim->second);
    g_used.erase(victim->first);
    g_cache.erase(victim);
  }
}

size_t table_bytes(int L) {
  const int64_t nlt = (L + 15) / 16, nlc = (L + 3) / 4;
  int64_t nf = 0, ni = 0;
  for (int c = 0; c < L; ++c) {
    nf += 256 * (tri(nlt) - tri(c >> 4));
    for (int mt = 0; mt < nlt; ++mt) ni += 64 * (nlc - (std::max(c, 16 * mt) >> 2));
  }
  return sizeof(double) * (nf + ni);
}

}  // namespace

// The smallest cached table of (dev, method) that covers L (its own L' >= L),
// or end().  A table built for L' serves every L <= L': Delta^l does not
// depend on L, each column's tiles for l < L are a prefix of its tiles for
// l < L', and the tile entries with L <= l < L' meet only zero coefficients
// (inverse) or outputs l >= L that are never stored (forward), so every
// transform's result is bit-identical with either table.
using TKey = std::tuple<int, int, int>;
std::map<TKey, DlTables>::iterator covering(int dev, int L, int method) {
  auto best = g_cache.end();
  for (auto it = g_cache.begin(); it != g_cache.end(); ++it) {
    const int d = std::get<0>(it->first), l = std::get<1>(it->first), m = std::get<2>(it->first);
    if (d != dev || m != method || l < L) continue;
    if (best == g_cache.end() || l < std::get<1>(best->first)) best = it;
  }
  return best;
}

const DlTables &dl_tables(int L, int method) {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_mu);
  auto it = covering(dev, L, method);
  if (it != g_cache.end()) {
    g_used[it->first] = ++g_clock;
    return it->second;
  }
  auto key = std::make_tuple(dev, L, method);
  g_used[key] = ++g_clock;
  make_room(table_bytes(L), key);
  return g_cache.emplace(key, build(L, method)).first->second;
}

const DlTables *dl_tables_find(int L, int method) {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_mu);
  auto it = covering(dev, L, method);
  if (it == g_cache.end()) return nullptr;
  g_used[it->first] = ++g_clock;
  return &it->second;
}

int dl_tables_pin(int L, int method) {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_mu);
  auto it = covering(dev, L, method);
  if (it == g_cache.end()) throw HipError("no Delta table to pin");
  ++g_pins[it->first];
  return std::get<1>(it->first);
}

void dl_tables_unpin(int table_L, int method) {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_mu);
  auto key = std::make_tuple(dev, table_L, method);
  auto it = g_pins.find(key);
  if (it == g_pins.end()) return;
  if (--it->second <= 0) g_pins.erase(it);
}

void dl_tables_limit(size_t bytes) {
  std::lock_guard<std::mutex> g(g_mu);
  g_limit = bytes;
}

const double *zero_page() {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_mu);
  auto it = g_zero.find(dev);
  if (it != g_zero.end()) return it->second;
  double *z = nullptr;
  SO3_HIP(hipMalloc(&z, sizeof(double) * kZeroPage));
  SO3_HIP(hipMemset(z, 0, sizeof(double) * kZeroPage));
  g_zero[dev] = z;
  return z;
}

void dl_tables_clear() {
  std::lock_guard<std::mutex> g(g_mu);
  for (auto &kv : g_zero) (void)hipFree(kv.second);
  g_zero.clear();
  for (auto &kv : g_cache) free_tables(kv.second);
  g_cache.clear();
  g_used.clear();
  g_pins.clear();
}

}  // namespace so3g
