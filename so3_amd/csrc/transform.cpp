// transform.cpp -- plans, workspace and the stage sequence of the inverse
// and forward transforms.
//
// Stage order (reference so3_core.c, via-SSHT and direct forms):
//   inverse: contraction (l) -> beta DFT (m') -> alpha DFT (m) -> gamma DFT (n)
//   forward: gamma DFT (g) -> alpha DFT (a) -> beta extension, DFT and weights
//            -> contraction (m')
// The Delta tables and DFT plans are cached per device; one workspace arena
// per device serves every single transform.  Its use is stream-ordered: each
// call records an event on its stream after its last stage, and the next
// call, on whatever stream, waits for that event before its first stage, so
// calls on different streams take turns on the arena instead of overwriting
// each other's intermediates.
#include "transform.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft_plan.hpp"
#include "so3/so3_sampling.h"

namespace so3g {

namespace {

using cld = std::complex<long double>;
const long double kPiL = 3.141592653589793238462643383279502884L;

struct Plan {
  Geo g{};
  int Pw = 0;
  double2 *phase_inv = nullptr;  // Pb entries
  double2 *phase_fwd = nullptr;  // 2L-1 entries, includes 1/Pb; then 3 x 16 block factors
  double2 *phase_ai = nullptr;   // 2L-1: exp(-i m'' pi/(2L-1)), adjoint of the inverse
  double2 *phase_af = nullptr;   // 2L-1: exp(+i m' pi/(2L-1)) / Pb, adjoint of the forward
  double2 *wr = nullptr;         // Pw entries
  double2 *wr_even = nullptr;    // Pw entries: the even taps of the weights only
  double2 *pole = nullptr;       // 2L entries: folded responses of the pole samples
  int64_t *nbase = nullptr;      // 2*(2N-1): masked, then unmasked
  size_t x1 = 0, x2 = 0;         // workspace sizes (double2)
};

// Length of the weight-convolution DFTs: the smallest 2^a or 3 * 2^a that is
// at least n (any length >= 4L - 3 gives the exact linear convolution; these
// run in one workgroup up to 4096 and keep 1024 for L = 256)
int conv_length(int n) {
  int best = 1;
  while (best < n) best *= 2;
  for (int m = 3; m < best; m *= 2)
    if (m >= n && m <= 1536) return m;
  return best;
}

double2 *upload(const std::vector<double2> &h) {
  double2 *d = nullptr;
  SO3_HIP(hipMalloc(&d, sizeof(double2) * std::max<size_t>(1, h.size())));
  SO3_HIP(hipMemcpy(d, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice));
  return d;
}

bool n_active_ssht(int n, int N, int mode) {
  if (mode == SO3_N_MODE_EVEN && (n % 2)) return false;
  if (mode == SO3_N_MODE_ODD && !(n % 2)) return false;
  if (mode == SO3_N_MODE_MAXIMUM && std::abs(n) < N - 1) return false;
  return true;
}

Geo make_geo(const so3_parameters_t *p, Method method) {
  Geo g{};
  g.L = p->L;
  g.N = p->N;
  g.L0 = p->L0;
  g.ss = (method == kViaSsht && p->sampling_scheme == SO3_SAMPLING_MW_SS) ? 1 : 0;
  g.real = p->reality ? 1 : 0;
  g.nA = g.ss ? 2 * g.L : 2 * g.L - 1;
  // X2 rows unpadded: padding rows to 8 entries (a power-of-two stride at
  // L = 256) was measured slower overall, the gamma columns most, and so was
  // a stride of 520 at L = 256 (128-byte aligned rows, not a power of two:
  // gamma 23% slower, the inverse beta 1% faster)
  g.ldA = g.nA;
  g.nB = g.ss ? g.L + 1 : g.L;
  g.nG = g.real ? g.N : 2 * g.N - 1;
  g.Pb = g.ss ? 2 * g.L : 2 * g.L - 1;
  g.NS = g.real ? 1 : 2;
  g.nmodeL = (method == kDirect && p->n_mode == SO3_N_MODE_L) ? 1 : 0;
  // the direct transforms ignore steerable, as the reference's do
  g.steer = (method == kViaSsht && p->steerable) ? 1 : 0;
  g.clmode = 1;
  return g;
}

// ---- tuned choices per (L, N, reality, MW_SS): see tune() below
struct TuneVal {
  int corder_inv = 0, corder_fwd = 0;  // contraction workgroup orders
  int pw = 0;                          // weight-convolution length (0: the default rule)
  int geng = 0;                        // gamma engine (Geo::geng; 0: the default rule)
};
using TuneKey = std::tuple<int, int, int, int>;
std::map<TuneKey, TuneVal> g_tune;

TuneKey tune_key(const so3_parameters_t *p, Method method) {
  const int ss = (method == kViaSsht && p->sampling_scheme == SO3_SAMPLING_MW_SS);
  return TuneKey(p->L, p->N, p->reality ? 1 : 0, ss);
}

TuneVal tuned(const so3_parameters_t *p, Method method) {
  auto it = g_tune.find(tune_key(p, method));
  return it == g_tune.end() ? TuneVal{} : it->second;
}

using Key = std::tuple<int, int, int, int, int, int, int, int, int, int, int, int, int>;

Key key_of(const so3_parameters_t *p, Method method, int dev) {
  const int ss = (method == kViaSsht && p->sampling_scheme == SO3_SAMPLING_MW_SS);
  const int steer = (method == kViaSsht && p->steerable) ? 1 : 0;
  return Key(dev, p->L, p->N, p->L0, ss, p->reality ? 1 : 0, (int)p->storage, (int)p->n_order,
             (int)p->n_mode, (int)method, steer, tuned(p, method).pw, tuned(p, method).geng);
}

Plan build_plan(const so3_parameters_t *p, Method method) {
  Plan P;
  Geo &g = P.g;
  g = make_geo(p, method);
  const int L = g.L, N = g.N;
  // inverse phase exp(i m' pi/(2L-1)) at k (m' = k or k - Pb); MW_SS none
  std::vector<double2> ph(g.Pb);
  for (int k = 0; k < g.Pb; ++k) {
    const int mp = k <= L - 1 ? k : k - g.Pb;
    if (g.ss) {
      ph[k] = make_double2(1.0, 0.0);
    } else {
      const long double a = kPiL * mp / (2.0L * L - 1.0L);
      ph[k] = make_double2((double)cosl(a), (double)sinl(a));
    }
  }
  P.phase_inv = upload(ph);
  // forward: exp(-i m'' pi/(2L-1)) / Pb for m'' in [-(L-1), L-1], then
  // exp(-i S j pi/(2L-1)) for S = 32, 64, 128 and j = -8 .. 7 (no 1/Pb): the
  // 1024-, 2048- and 4096-point forward beta form the factor of m'' = d + S j
  // (d < S, one per lane) as pf(d) times that entry
  std::vector<double2> pf(2 * L - 1 + 48);
  for (int mpp = -(L - 1); mpp <= L - 1; ++mpp) {
    if (g.ss) {
      pf[mpp + L - 1] = make_double2(1.0 / g.Pb, 0.0);
    } else {
      const long double a = -kPiL * mpp / (2.0L * L - 1.0L);
      pf[mpp + L - 1] = make_double2((double)(cosl(a) / g.Pb), (double)(sinl(a) / g.Pb));
    }
  }
  for (int e = 0; e < 3; ++e)
    for (int j = -8; j < 8; ++j) {
      const long double a = g.ss ? 0.0L : -kPiL * (32 << e) * j / (2.0L * L - 1.0L);
      pf[2 * L - 1 + 16 * e + 8 + j] = make_double2((double)cosl(a), (double)sinl(a));
    }
  P.phase_fwd = upload(pf);
  {
    std::vector<double2> ai(2 * L - 1), af(2 * L - 1);
    for (int mp = -(L - 1); mp <= L - 1; ++mp) {
      const long double a = kPiL * mp / (2.0L * L - 1.0L);
      ai[mp + L - 1] = g.ss ? make_double2(1.0, 0.0) : make_double2((double)cosl(a), (double)-sinl(a));
      af[mp + L - 1] = g.ss ? make_double2(1.0 / g.Pb, 0.0)
                            : make_double2((double)(cosl(a) / g.Pb), (double)(sinl(a) / g.Pb));
    }
    P.phase_ai = upload(ai);
    P.phase_af = upload(af);
  }
  // weights w(p) = int_0^pi exp(-i p b) sin b db (so3_sampling_weight) at p mod Pw,
  // backward DFT of length Pw >= 4L-3 (any such length is exact)
  P.Pw = conv_length(4 * L - 3);
  // the register-resident forward beta (beta DFTs on Bluestein M = 1024, so
  // L = 129..256) convolves at 1024 points; any Pw >= 4L - 3 is exact
  if (L <= 256 && fft_plan(g.Pb).blue && fft_plan(g.Pb).len == 1024) P.Pw = 1024;
  // the 2048-point register forward beta (L = 257..512) convolves at 2048
  if (L >= 257 && L <= 512) P.Pw = 2048;
  // and the 4096-point one (L = 513..1024) at 4096
  if (L >= 513 && L <= 1024) P.Pw = 4096;
  if (const int pw = tuned(p, method).pw) P.Pw = pw;  // the tuner's choice (>= 4L - 3)
  g.geng = tuned(p, method).geng;
  {
    const int Pw = P.Pw;
    std::vector<cld> w(Pw, cld(0, 0));
    for (int q = -2 * (L - 1); q <= 2 * (L - 1); ++q) {
      cld v(0, 0);
      if (q == 1) v = cld(0, -kPiL / 2);
      else if (q == -1) v = cld(0, kPiL / 2);
      else if (q % 2 == 0) v = cld(2.0L / (1.0L - (long double)q * q), 0);
      w[((q % Pw) + Pw) % Pw] = v;
    }
    // wr[r] = sum_q w[q] exp(+2 pi i q r / Pw) (a backward DFT in long
    // double, Pw = 2^a or 3 * 2^a); w has 2L-1 non-zero even taps + 2 odd
    // ones (q = +-1); wr_even leaves the odd ones out
    std::vector<cld> acc(w), ev(w);
    if (L >= 2) {  // the odd taps q = +-1 exist
      ev[1] = cld(0, 0);
      ev[Pw - 1] = cld(0, 0);
    }
    host_dft_ld(acc, +1);
    host_dft_ld(ev, +1);
    std::vector<double2> wr(Pw), wre(Pw);
    for (int r = 0; r < Pw; ++r) {
      wr[r] = make_double2((double)acc[r].real(), (double)acc[r].imag());
      wre[r] = make_double2((double)ev[r].real(), (double)ev[r].imag());
    }
    P.wr = upload(wr);
    P.wr_even = upload(wre);
    // Pole samples (b = L-1 for MW; b = 0 and b = L for MW_SS) map to
    // themselves under the extension, so an antisymmetric sequence
    // ((-1)^{m+n} = -1) with a non-zero sample there is not antisymmetric.
    // pole[j L + m'] is the folded (sign -1) forward-beta response to a unit
    // sample at the j-th pole sample: G(m') - G(-m') for m' > 0, G(0) at 0,
    // G(m') = 2 pi sum_k F(k) w(m' - k), F(k) = exp(-2 pi i k b0 / Pb) * phase(k).
    {
      std::vector<double2> pole(2 * L, make_double2(0.0, 0.0));
      const int npole = g.ss ? 2 : 1;
      for (int j = 0; j < npole; ++j) {
        const int b0 = g.ss ? j * L : L - 1;
        std::vector<cld> F(2 * L - 1), G(2 * L - 1);
        for (int k = -(L - 1); k <= L - 1; ++k) {
          long long e = ((long long)k * b0) % g.Pb;
          if (e < 0) e += g.Pb;
          const long double a = -2.0L * kPiL * e / g.Pb;
          const long double ph = g.ss ? 0.0L : -kPiL * k / (2.0L * L - 1.0L);
          F[k + L - 1] = cld(cosl(a), sinl(a)) * cld(cosl(ph), sinl(ph)) / (long double)g.Pb;
        }
        for (int mp = -(L - 1); mp <= L - 1; ++mp) {
          cld acc(0, 0);
          for (int k = -(L - 1); k <= L - 1; ++k) {
            const int q = mp - k;
            const cld wq = w[((q % Pw) + Pw) % Pw];
            if (wq != cld(0, 0)) acc += F[k + L - 1] * wq;
          }
          G[mp + L - 1] = 2.0L * kPiL * acc;
        }
        for (int mp = 0; mp < L; ++mp) {
          const cld t = mp == 0 ? G[L - 1] : G[mp + L - 1] - G[-mp + L - 1];
          pole[j * L + mp] = make_double2((double)t.real(), (double)t.imag());
        }
      }
      P.pole = upload(pole);
    }
  }
  // per-n storage bases
  std::vector<int64_t> nb(2 * (2 * N - 1), kNoBlock);
  for (int n = -(N - 1); n <= N - 1; ++n) {
    if (g.real && n < 0) continue;
    const int an = std::abs(n);
    int ind = 0;
    if (an < L) {
      if (g.real) so3_sampling_elmn2ind_real(&ind, an, -an, n, p);
      else so3_sampling_elmn2ind(&ind, an, -an, n, p);
    } else {
      // |n| >= L: block holds nothing (only padded storage reaches here)
      if (g.real) so3_sampling_elmn2ind_real(&ind, 0, 0, n, p);
      else so3_sampling_elmn2ind(&ind, 0, 0, n, p);
      ind += an * an;
    }
    const int64_t base = (int64_t)ind - (int64_t)an * an;
    nb[(2 * N - 1) + n + N - 1] = base;
    bool act = n_active_ssht(n, N, p->n_mode);
    if (an >= L) act = false;
    nb[n + N - 1] = act ? base : kNoBlock;
  }
  {
    int64_t *d = nullptr;
    SO3_HIP(hipMalloc(&d, sizeof(int64_t) * nb.size()));
    SO3_HIP(hipMemcpy(d, nb.data(), sizeof(int64_t) * nb.size(), hipMemcpyHostToDevice));
    P.nbase = d;
  }
  // the forward contraction copies G' in 4-row chunks: up to 3 rows past the
  // last pair's stream
  P.x1 = std::max<size_t>((size_t)g.nG * g.nA * L, (size_t)N * L * L * 2 * g.NS + 8 * g.NS);
  P.x2 = (size_t)g.nG * g.nB * g.ldA;
  // DFT plans and the zero page are made now so that a timed (or captured)
  // call does not allocate
  zero_page();
  fft_plan(g.Pb);
  fft_plan(g.nA);
  fft_plan(2 * N - 1);
  gamma_short_prepare(g);
  if (2 * N - 1 >= 33 && 2 * N - 1 <= 127) fft_plan_r256(2 * N - 1);
  fft_plan(P.Pw);
  fft_plan_lds(g.Pb);
  fft_plan_lds(P.Pw);
  fft_plan_lds(2 * N);
  for (int n : {g.Pb, g.nA, 2 * N - 1})
    if (n >= 513 && n <= 2048) fft_plan_reg(n);
  return P;
}

std::mutex g_mu;
std::map<Key, Plan> g_plans;

// ---- stage profiler
bool g_timing = false;
struct Pending {
  int stage;
  hipEvent_t a, b;
};
std::vector<Pending> g_pending;
const char *kStageNames[] = {"contract_inv", "beta_inv", "alpha_inv", "gamma_inv",
                             "gamma_fwd",    "alpha_fwd", "beta_fwd", "zero_fwd",
                             "contract_fwd"};
constexpr int kNStages = 9;
double g_ms[kNStages];
int g_cnt[kNStages];

template <typename F>
void timed(int stage, hipStream_t st, F &&launch) {
  if (!g_timing) {
    launch();
    return;
  }
  Pending p{stage, nullptr, nullptr};
  SO3_HIP(hipEventCreate(&p.a));
  SO3_HIP(hipEventCreate(&p.b));
  SO3_HIP(hipEventRecord(p.a, st));
  launch();
  SO3_HIP(hipEventRecord(p.b, st));
  g_pending.push_back(p);
}

void resolve_pending() {
  for (auto &p : g_pending) {
    SO3_HIP(hipEventSynchronize(p.b));
    float ms = 0.f;
    SO3_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    g_ms[p.stage] += ms;
    g_cnt[p.stage] += 1;
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  g_pending.clear();
}

struct Arena {
  double2 *buf = nullptr;
  size_t cap = 0;              // double2 elements
  hipEvent_t done = nullptr;   // end of the last call that used buf
  bool recorded = false;
};
std::map<int, Arena> g_arena;

const Plan &get_plan(const so3_parameters_t *p, Method method) {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  const Key k = key_of(p, method, dev);
  auto it = g_plans.find(k);
  if (it != g_plans.end()) return it->second;
  return g_plans.emplace(k, build_plan(p, method)).first->second;
}

bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  SO3_HIP(hipStreamIsCapturing(st, &cs));
  return cs != hipStreamCaptureStatusNone;
}

// The device's arena, at least `elems` double2 long, ready for work
// enqueued next on `st`: ordered after every earlier user.  Pair with
// arena_release(st) after the call's last launch.
double2 *arena_acquire(size_t elems, hipStream_t st) {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  Arena &a = g_arena[dev];
  if (!a.done) SO3_HIP(hipEventCreateWithFlags(&a.done, hipEventDisableTiming));
  if (a.cap < elems) {
    if (capturing(st)) throw HipError("the workspace must be allocated before graph capture (call once eagerly)");
    if (a.buf) {
      SO3_HIP(hipDeviceSynchronize());
      SO3_HIP(hipFree(a.buf));
    }
    a.buf = nullptr;
    a.cap = 0;
    SO3_HIP(hipMalloc(&a.buf, sizeof(double2) * elems));
    a.cap = elems;
  }
  // a captured call is ordered by its graph; an event recorded outside the
  // capture cannot be waited on inside it
  if (a.recorded && !capturing(st)) SO3_HIP(hipStreamWaitEvent(st, a.done, 0));
  return a.buf;
}

void arena_release(hipStream_t st) {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  Arena &a = g_arena[dev];
  if (capturing(st)) return;
  SO3_HIP(hipEventRecord(a.done, st));
  a.recorded = true;
}

}  // namespace

void validate(const so3_parameters_t *p, Method method) {
  if (!p) throw HipError("parameters pointer is NULL");
  if (p->L < 1) throw HipError("L must be positive");
  if (p->N < 1) throw HipError("N must be positive");
  if (p->L0 < 0 || p->L0 >= p->L) throw HipError("L0 must be in [0, L)");
  if (p->sampling_scheme != SO3_SAMPLING_MW && p->sampling_scheme != SO3_SAMPLING_MW_SS)
    throw HipError("Invalid sampling scheme.");
  if (p->storage != SO3_STORAGE_PADDED && p->storage != SO3_STORAGE_COMPACT)
    throw HipError("Invalid storage method.");
  if (p->n_order != SO3_N_ORDER_ZERO_FIRST && p->n_order != SO3_N_ORDER_NEGATIVE_FIRST)
    throw HipError("Invalid n-order.");
  if ((int)p->n_mode < 0 || (int)p->n_mode >= (int)SO3_N_MODE_SIZE) throw HipError("Invalid n-mode.");
  if ((int)p->dl_method != 0 && (int)p->dl_method != 1) throw HipError("Invalid dl method");
  if (p->storage == SO3_STORAGE_COMPACT && p->N > p->L)
    throw HipError("compact storage needs N <= L");
  // the reference's real via-SSHT transforms have their steerable branches
  // commented out (so3_core.c:526-548, 743-813)
  if (method == kViaSsht && p->steerable && p->reality)
    throw HipError("steerable real signals are not supported");
  if ((size_t)(2 * p->N - 1) * (2 * p->L) * (p->L + 1) > (size_t)1 << 40)
    throw HipError("transform too large");
}

bool forward_writes_all(const so3_parameters_t *p, Method method) {
  if (method == kViaSsht) return p->n_mode == SO3_N_MODE_ALL || p->n_mode == SO3_N_MODE_L;
  return p->storage == SO3_STORAGE_COMPACT;  // direct leaves padded l < |n| alone
}

size_t workspace_bytes(const so3_parameters_t *p, Method method) {
  validate(p, method);
  std::lock_guard<std::mutex> lk(g_mu);
  const Plan &P = get_plan(p, method);
  return sizeof(double2) * (P.x1 + P.x2);
}

void prepare(const so3_parameters_t *p, Method method) {
  validate(p, method);
  std::lock_guard<std::mutex> lk(g_mu);
  get_plan(p, method);
  dl_tables(p->L, (int)p->dl_method);
}

namespace {

void inverse_post_locked(const Plan &P, void *f, double2 *ws, hipStream_t st);

void enqueue_inverse_locked(const Plan &P, const DlTables &dl, const void *flmn, void *f, double2 *ws,
                            hipStream_t st, int corder) {
  double2 *X1 = ws;
  Geo g = P.g;
  g.corder = corder;
  NBase nb{P.nbase};
  timed(0, st, [&] { launch_contract_inv(g, (const double2 *)flmn, nb, dl, X1, st); });
  inverse_post_locked(P, f, ws, st);
}

// the inverse's stages after the contraction (X1 at ws)
void inverse_post_locked(const Plan &P, void *f, double2 *ws, hipStream_t st) {
  double2 *X1 = ws, *X2 = ws + P.x1;
  const Geo &g = P.g;
  NBase nb{P.nbase};
  timed(1, st, [&] { launch_inv_beta(g, X1, X2, nb, P.phase_inv, st); });
  if (g.steer) {
    timed(2, st, [&] { launch_alpha(g, X2, +1, 1.0, st); });
    timed(3, st, [&] { launch_gamma_steer(g, X2, f, +1, 1.0, st); });
  } else if (reg_order(g)) {
    timed(3, st, [&] { launch_gamma_x2(g, X2, +1, 1.0, st); });
    timed(2, st, [&] { launch_alpha_x2f(g, X2, (double2 *)f, +1, 1.0, st); });
  } else {
    timed(2, st, [&] { launch_alpha(g, X2, +1, 1.0, st); });
    timed(3, st, [&] { launch_inv_gamma(g, X2, f, 1.0, st); });
  }
}

// the forward's stages before the contraction (G' left at ws)
void forward_pre_locked(const Plan &P, const so3_parameters_t *p, Method method, const void *f, void *flmn,
                        double2 *ws, hipStream_t st) {
  double2 *Gp = ws, *X2 = ws + P.x1;
  const Geo &g = P.g;
  NBase nb{P.nbase};
  if (g.steer) {
    timed(4, st, [&] { launch_gamma_steer(g, f, X2, -1, 2 * M_PI / g.N, st); });
    timed(5, st, [&] { launch_alpha(g, X2, -1, 1.0 / g.nA, st); });
  } else if (reg_order(g)) {
    timed(5, st, [&] { launch_alpha_fx2(g, (const double2 *)f, X2, -1, 1.0 / g.nA, st); });
    timed(4, st, [&] { launch_gamma_x2(g, X2, -1, 2 * M_PI / (2.0 * g.N - 1.0), st); });
  } else {
    timed(4, st, [&] { launch_fwd_gamma(g, f, X2, 2 * M_PI / (2.0 * g.N - 1.0), st); });
    timed(5, st, [&] { launch_alpha(g, X2, -1, 1.0 / g.nA, st); });
  }
  timed(6, st, [&] {
    launch_fwd_beta(g, X2, Gp, nb, P.phase_fwd, P.wr, P.wr_even, P.pole, P.Pw, 0, st);
  });
  timed(7, st, [&] {
    launch_fwd_zero(g, (double2 *)flmn, nb, method == kDirect, 0, (int)p->storage, (int)p->n_order, st);
  });
}

void enqueue_forward_locked(const Plan &P, const so3_parameters_t *p, Method method, const void *f, void *flmn,
                            double2 *ws, hipStream_t st) {
  const DlTables &dl = dl_tables(p->L, (int)p->dl_method);
  forward_pre_locked(P, p, method, f, flmn, ws, st);
  Geo g = P.g;
  g.corder = tuned(p, method).corder_fwd;
  NBase nb{P.nbase};
  timed(8, st, [&] { launch_contract_fwd(g, ws, nb, dl, (double2 *)flmn, st); });
}

}  // namespace

void run_inverse(const so3_parameters_t *p, Method method, const void *flmn, void *f, hipStream_t st) {
  validate(p, method);
  std::lock_guard<std::mutex> lk(g_mu);
  const Plan &P = get_plan(p, method);
  const DlTables &dl = dl_tables(p->L, (int)p->dl_method);  // built (synchronously) before the arena is taken
  enqueue_inverse_locked(P, dl, flmn, f, arena_acquire(P.x1 + P.x2, st), st, tuned(p, method).corder_inv);
  arena_release(st);
}

void run_forward(const so3_parameters_t *p, Method method, const void *f, void *flmn, hipStream_t st) {
  validate(p, method);
  std::lock_guard<std::mutex> lk(g_mu);
  const Plan &P = get_plan(p, method);
  dl_tables(p->L, (int)p->dl_method);
  enqueue_forward_locked(P, p, method, f, flmn, arena_acquire(P.x1 + P.x2, st), st);
  arena_release(st);
}

size_t workspace_elems(const so3_parameters_t *p, Method method) {
  validate(p, method);
  std::lock_guard<std::mutex> lk(g_mu);
  const Plan &P = get_plan(p, method);
  dl_tables(p->L, (int)p->dl_method);
  return P.x1 + P.x2;
}

void enqueue_inverse(const so3_parameters_t *p, Method method, const void *flmn, void *f, void *ws,
                     hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  enqueue_inverse_locked(get_plan(p, method), dl_tables(p->L, (int)p->dl_method), flmn, f, (double2 *)ws, st,
                         tuned(p, method).corder_inv);
}

void enqueue_forward(const so3_parameters_t *p, Method method, const void *f, void *flmn, void *ws,
                     hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  enqueue_forward_locked(get_plan(p, method), p, method, f, flmn, (double2 *)ws, st);
}

// ---- the allocation-free inverse (so3x_inverse_noalloc)
//
// Its workspace is this file-scope device array, one copy per GPU that the
// library runs on, held until the process ends.  In host code `scratch` is
// only a stand-in: the device address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device (with that device current) on
// the entry point's first call there, and only that address is passed to
// kernels.  The transform's kernels write every workspace entry they read,
// so nothing is cleared between calls.  All calls on one GPU share the
// array: calls that overlap on one GPU (different streams, no ordering)
// race.
constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB
constexpr size_t kScratchElems = kScratchBytes / sizeof(double2);
static __device__ __align__(256) double2 scratch[kScratchElems];
std::map<int, double2 *> g_scratch;  // device -> address of its copy of scratch

double2 *scratch_address() {
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) return it->second;
  void *ptr = nullptr;
  SO3_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) throw HipError("hipGetSymbolAddress returned no address for the static workspace");
  g_scratch[dev] = (double2 *)ptr;
  return (double2 *)ptr;
}

size_t scratch_limit_bytes() { return kScratchBytes; }

void run_inverse_noalloc(const so3_parameters_t *p, Method method, const void *flmn, void *f, hipStream_t st) {
  validate(p, method);
  std::lock_guard<std::mutex> lk(g_mu);
  int dev = 0;
  SO3_HIP(hipGetDevice(&dev));
  auto it = g_plans.find(key_of(p, method, dev));
  if (it == g_plans.end())
    throw HipError("parameters not prepared on this device: call so3x_prepare (or so3x_inverse) with them first");
  const Plan &P = it->second;
  const DlTables *dl = dl_tables_find(p->L, (int)p->dl_method);
  if (!dl) throw HipError("Delta table not resident: call so3x_prepare with these parameters first");
  const size_t need = sizeof(double2) * (P.x1 + P.x2);
  if (need > kScratchBytes)
    throw HipError("workspace of " + std::to_string(need) + " bytes exceeds the " +
                   std::to_string(kScratchBytes) + "-byte static workspace of the allocation-free entry point");
  double2 *ws = scratch_address();
  const bool timing = g_timing;  // the stage profiler creates events: off for this call
  g_timing = false;
  try {
    enqueue_inverse_locked(P, *dl, flmn, f, ws, st, tuned(p, method).corder_inv);
  } catch (...) {
    g_timing = timing;
    throw;
  }
  g_timing = timing;
}

// ---- autotuner
//
// For one (L, N, reality, sampling) the tuner times, on zero-filled buffers
// of the real size (kernel times do not depend on the values):
//   * the inverse and the forward contraction in both workgroup orders;
//   * the forward beta stage with each candidate weight-convolution length
//     Pw: the default rule and every other 2^a or 3 * 2^a in [4L - 3, 4096]
//     up to twice the default that the workgroup FFT engine runs directly
//     without Bluestein (any Pw >= 4L - 3 gives the exact linear
//     convolution; the length decides which FFT engine and plan run it);
//   * for gamma lengths 2N - 1 in [33, 127], the inverse and forward gamma
//     stages on the workgroup engine (its plan: direct for a smooth length,
//     else Bluestein) and on the 256-point register FFT (Bluestein).
// The fastest of each is kept and applies to every later transform with
// these parameters, alone or batched (so batched results stay bit-identical
// to single ones).  The workgroup order never changes a result; another Pw
// or gamma engine changes the rounding only.
namespace {

// the default, then the other lengths the workgroup engine runs directly
std::vector<int> pw_candidates(int L, int def) {
  std::vector<int> c{def};
  const int lo = 4 * L - 3, hi = std::min(4096, 2 * def);
  for (int b = 1; b <= hi; b *= 2)
    for (int m : {b, 3 * b})
      if (m >= lo && m <= hi && std::find(c.begin(), c.end(), m) == c.end() && !fft_plan_lds(m).blue)
        c.push_back(m);
  std::sort(c.begin() + 1, c.end());
  return c;
}

template <typename F>
float best_ms(hipStream_t st, hipEvent_t a, hipEvent_t b, F &&fn) {
  fn();  // warm-up (sets launch attributes)
  float best = 1e30f;
  for (int r = 0; r < 2; ++r) {
    SO3_HIP(hipEventRecord(a, st));
    fn();
    SO3_HIP(hipEventRecord(b, st));
    SO3_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    SO3_HIP(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, ms);
  }
  return best;
}

}  // namespace

void tune(const so3_parameters_t *p, Method method, hipStream_t st, int *out, double *ms_out) {
  validate(p, method);
  std::lock_guard<std::mutex> lk(g_mu);
  const TuneKey tk = tune_key(p, method);
  g_tune.erase(tk);  // the default rule first
  const Plan &P0 = get_plan(p, method);
  const DlTables &dl = dl_tables(p->L, (int)p->dl_method);
  const size_t nl = (size_t)so3_sampling_flmn_size(p) + 64;
  double2 *flmn = nullptr, *ws = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  TuneVal best;
  double ms[4] = {0, 0, 0, 0};
  double2 *fbuf = nullptr;
  try {
    SO3_HIP(hipMalloc(&flmn, sizeof(double2) * nl));
    SO3_HIP(hipMalloc(&ws, sizeof(double2) * (P0.x1 + P0.x2)));
    SO3_HIP(hipMemsetAsync(flmn, 0, sizeof(double2) * nl, st));
    SO3_HIP(hipMemsetAsync(ws, 0, sizeof(double2) * (P0.x1 + P0.x2), st));
    SO3_HIP(hipEventCreate(&a));
    SO3_HIP(hipEventCreate(&b));
    NBase nb{P0.nbase};
    double2 *X1 = ws, *X2 = ws + P0.x1;
    float bi = 1e30f, bf = 1e30f;
    for (int order = 0; order < 2; ++order) {
      Geo g = P0.g;
      g.corder = order;
      const float ti = best_ms(st, a, b, [&] { launch_contract_inv(g, flmn, nb, dl, X1, st); });
      const float tf = best_ms(st, a, b, [&] { launch_contract_fwd(g, X1, nb, dl, flmn, st); });
      if (ti < bi) bi = ti, best.corder_inv = order;
      if (tf < bf) bf = tf, best.corder_fwd = order;
    }
    ms[0] = bi;
    ms[1] = bf;
    float bw = 1e30f;
    for (int pw : pw_candidates(p->L, P0.Pw)) {
      TuneVal t;
      t.pw = pw == P0.Pw ? 0 : pw;
      g_tune[tk] = t;
      const Plan &Pc = get_plan(p, method);
      const float tw = best_ms(st, a, b, [&] {
        launch_fwd_beta(Pc.g, X2, X1, nb, Pc.phase_fwd, Pc.wr, Pc.wr_even, Pc.pole, Pc.Pw, 0, st);
      });
      if (tw < bw) bw = tw, best.pw = t.pw;
    }
    ms[2] = bw;
    if (gamma_engine(P0.g) != 0) {
      const size_t nf = (size_t)so3_sampling_f_size(p);
      SO3_HIP(hipMalloc(&fbuf, sizeof(double2) * nf));
      SO3_HIP(hipMemsetAsync(fbuf, 0, sizeof(double2) * nf, st));
      float bg = 1e30f;
      for (int e : {1, 2}) {
        TuneVal t = best;
        t.geng = e;
        g_tune[tk] = t;
        const Plan &Pc = get_plan(p, method);
        const float tg = best_ms(st, a, b, [&] {
          launch_inv_gamma(Pc.g, X2, fbuf, 1.0, st);
          launch_fwd_gamma(Pc.g, fbuf, X2, 1.0, st);
        });
        if (tg < bg) bg = tg, best.geng = e;
      }
      ms[3] = bg;
    }
  } catch (...) {
    g_tune.erase(tk);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipStreamSynchronize(st);
    (void)hipFree(flmn);
    (void)hipFree(ws);
    (void)hipFree(fbuf);
    throw;
  }
  g_tune[tk] = best;
  SO3_HIP(hipStreamSynchronize(st));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  SO3_HIP(hipFree(flmn));
  SO3_HIP(hipFree(ws));
  SO3_HIP(hipFree(fbuf));
  if (out) {
    out[0] = best.corder_inv;
    out[1] = best.corder_fwd;
    out[2] = best.pw ? best.pw : P0.Pw;
    out[3] = gamma_engine(get_plan(p, method).g);
  }
  if (ms_out)
    for (int i = 0; i < 4; ++i) ms_out[i] = ms[i];
}

void tune_clear() {
  std::lock_guard<std::mutex> lk(g_mu);
  g_tune.clear();
}

void tune_get(const so3_parameters_t *p, Method method, int *out) {
  validate(p, method);
  std::lock_guard<std::mutex> lk(g_mu);
  const TuneVal t = tuned(p, method);
  out[0] = t.corder_inv;
  out[1] = t.corder_fwd;
  out[2] = t.pw ? t.pw : get_plan(p, method).Pw;
  out[3] = gamma_engine(get_plan(p, method).g);
}

// one line per entry: "L N reality mw_ss corder_inv corder_fwd pw gamma" (pw,
// gamma 0: the default rule); tune_load also reads lines without the gamma
// column
void tune_save(const char *path) {
  std::lock_guard<std::mutex> lk(g_mu);
  FILE *fp = fopen(path, "w");
  if (!fp) throw HipError(std::string("cannot write ") + path);
  fprintf(fp, "# so3 MI355X tuning table: L N reality mw_ss corder_inv corder_fwd pw gamma\n");
  for (auto &kv : g_tune)
    fprintf(fp, "%d %d %d %d %d %d %d %d\n", std::get<0>(kv.first), std::get<1>(kv.first),
            std::get<2>(kv.first), std::get<3>(kv.first), kv.second.corder_inv, kv.second.corder_fwd,
            kv.second.pw, kv.second.geng);
  if (fclose(fp) != 0) throw HipError(std::string("cannot write ") + path);
}

int tune_load(const char *path) {
  std::lock_guard<std::mutex> lk(g_mu);
  FILE *fp = fopen(path, "r");
  if (!fp) throw HipError(std::string("cannot read ") + path);
  char line[256];
  int n = 0;
  while (fgets(line, sizeof line, fp)) {
    if (line[0] == '#') continue;
    int L, N, r, ss, ci, cf, pw, ge = 0;
    const int got = sscanf(line, "%d %d %d %d %d %d %d %d", &L, &N, &r, &ss, &ci, &cf, &pw, &ge);
    if (got != 7 && got != 8) continue;
    if (L < 1 || N < 1 || (ci != 0 && ci != 1) || (cf != 0 && cf != 1) || ge < 0 || ge > 2 ||
        (pw != 0 && (pw < 4 * L - 3 || pw > 4096 || (pw & (pw - 1) && (pw % 3 || (pw / 3) & (pw / 3 - 1)))))) {
      fclose(fp);
      throw HipError(std::string("invalid tuning entry in ") + path + ": " + line);
    }
    TuneVal t;
    t.corder_inv = ci;
    t.corder_fwd = cf;
    t.pw = pw;
    t.geng = ge;
    g_tune[TuneKey(L, N, r, ss)] = t;
    ++n;
  }
  fclose(fp);
  return n;
}

// ---- grouped batches: a transform split at its contraction
void contract_item(const so3_parameters_t *p, Method method, bool inverse, const void *in, void *out, void *ws,
                   CItem *it) {
  std::lock_guard<std::mutex> lk(g_mu);
  const Plan &P = get_plan(p, method);
  const DlTables &dl = dl_tables(p->L, (int)p->dl_method);
  const TuneVal tv = tuned(p, method);
  *it = CItem{};
  it->g = P.g;
  it->g.corder = inverse ? tv.corder_inv : tv.corder_fwd;
  it->in = inverse ? in : ws;
  it->out = inverse ? (double2 *)ws : (double2 *)out;
  it->nbase = P.nbase;
  it->tab = inverse ? dl.ti : dl.tf;
  it->off_tab = inverse ? dl.offi : dl.offf;
  it->nmt = dl.nmt;
  it->nblk = contract_blocks(it->g);
}

void enqueue_inverse_post(const so3_parameters_t *p, Method method, void *f, void *ws, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  const Plan &P = get_plan(p, method);
  launch_zero_rows(P.g, (double2 *)ws, st);
  inverse_post_locked(P, f, (double2 *)ws, st);
}

void enqueue_forward_pre(const so3_parameters_t *p, Method method, const void *f, void *flmn, void *ws,
                         hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  forward_pre_locked(get_plan(p, method), p, method, f, flmn, (double2 *)ws, st);
}

void set_stage_timing(bool on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_timing = on;
}

void reset_stage_times() {
  std::lock_guard<std::mutex> lk(g_mu);
  resolve_pending();
  for (int i = 0; i < kNStages; ++i) g_ms[i] = 0.0, g_cnt[i] = 0;
}

int stage_times(const char **names, double *ms, int *count, int cap) {
  std::lock_guard<std::mutex> lk(g_mu);
  resolve_pending();
  const int n = std::min(cap, kNStages);
  for (int i = 0; i < n; ++i) {
    if (names) names[i] = kStageNames[i];
    if (ms) ms[i] = g_ms[i];
    if (count) count[i] = g_cnt[i];
  }
  return kNStages;
}

// adjoint of so3_core_inverse_direct: f -> flmn (so3_adjoint.c:34-400)
void run_adjoint_inverse(const so3_parameters_t *p, const void *f, void *flmn, hipStream_t st) {
  validate(p, kDirect);
  std::lock_guard<std::mutex> lk(g_mu);
  const Plan &P = get_plan(p, kDirect);
  const DlTables &dl = dl_tables(p->L, (int)p->dl_method);
  double2 *ws = arena_acquire(P.x1 + P.x2, st);
  double2 *Gp = ws, *X2 = ws + P.x1;
  Geo g = P.g;
  NBase nb{P.nbase};
  if (reg_order(g)) {
    launch_alpha_fx2(g, (const double2 *)f, X2, -1, 1.0, st);
    launch_gamma_x2(g, X2, -1, 1.0, st);
  } else {
    launch_fwd_gamma(g, f, X2, 1.0, st);
    launch_alpha(g, X2, -1, 1.0, st);
  }
  launch_fwd_beta(g, X2, Gp, nb, P.phase_ai, P.wr, nullptr, nullptr, P.Pw, 1, st);
  launch_fwd_zero(g, (double2 *)flmn, nb, 1, 0, (int)p->storage, (int)p->n_order, st);
  g.clmode = 2;
  launch_contract_fwd(g, Gp, nb, dl, (double2 *)flmn, st);
  arena_release(st);
}

// adjoint of so3_core_forward_direct: flmn -> f (so3_adjoint.c:402-937)
void run_adjoint_forward(const so3_parameters_t *p, const void *flmn, void *f, hipStream_t st) {
  validate(p, kDirect);
  std::lock_guard<std::mutex> lk(g_mu);
  const Plan &P = get_plan(p, kDirect);
  const DlTables &dl = dl_tables(p->L, (int)p->dl_method);
  double2 *ws = arena_acquire(P.x1 + P.x2, st);
  double2 *X1 = ws, *X2 = ws + P.x1;
  Geo g = P.g;
  g.clmode = 0;
  NBase nb{P.nbase};
  launch_contract_inv(g, (const double2 *)flmn, nb, dl, X1, st);
  launch_adjf_beta(g, X1, X2, nb, P.phase_af, P.wr, P.Pw, st);
  if (reg_order(g)) {
    launch_gamma_x2(g, X2, +1, 1.0 / (2.0 * g.N - 1.0), st);
    launch_alpha_x2f(g, X2, (double2 *)f, +1, 1.0 / g.nA, st);
  } else {
    launch_alpha(g, X2, +1, 1.0 / g.nA, st);
    launch_inv_gamma(g, X2, f, 1.0 / (2.0 * g.N - 1.0), st);
  }
  arena_release(st);
}

void release_all() {
  std::lock_guard<std::mutex> lk(g_mu);
  (void)hipDeviceSynchronize();
  for (auto &kv : g_plans) {
    (void)hipFree(kv.second.phase_inv);
    (void)hipFree(kv.second.phase_fwd);
    (void)hipFree(kv.second.phase_ai);
    (void)hipFree(kv.second.phase_af);
    (void)hipFree(kv.second.wr);
    (void)hipFree(kv.second.wr_even);
    (void)hipFree(kv.second.pole);
    (void)hipFree(kv.second.nbase);
  }
  g_plans.clear();
  for (auto &kv : g_arena) {
    (void)hipFree(kv.second.buf);
    if (kv.second.done) (void)hipEventDestroy(kv.second.done);
  }
  g_arena.clear();
  dl_tables_clear();
}

}  // namespace so3g
