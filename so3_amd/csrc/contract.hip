// contract.hip -- the FP64 Delta-Delta contraction of the Wigner transform.
//
// Inverse (reference direct form so3_core.c:1081-1114, via-SSHT form inside
// ssht_core_mw_lb_inverse_sov_sym called at so3_core.c:204):
//   F(m,n,m') = i^{n-m} sum_{l >= max(L0,|m|,|n|,m')} c_l D^l_{m'm} D^l_{m'n} f_lmn,
//   c_l = (2l+1)/(8 pi^2), for 0 <= m' < L.
// Forward (so3_core.c:1589-1617, ssht_core_mw_lb_forward_sov_conv_sym at :410):
//   f_lmn = i^{m-n} sum_{0 <= m' <= l} D^l_{m'm} D^l_{m'n} G'(m,n,m').
//
// The four sign combinations (+-m, +-n) of a pair (x, y) = (|m|, |n|) are
// treated together: with D^l_{m',-m} = (-1)^{l+m'} D^l_{m',m}, one product
// Q(l, m') = D^l_{m'x} D^l_{m'y} feeds all four.  The factor (-1)^l of the
// mixed-sign combinations (and c_l of the synthesis) multiplies the
// coefficients as they enter the matrix instruction (inverse) or the result
// at the store (forward); (-1)^{m'} is applied at the store (inverse) or
// folded into G' by the forward beta stage.
//
// For one pair the contraction is a small matrix product over the reduction
// index r (m' forward, l inverse) with the lane index s (l forward, m'
// inverse):  out(s, col) = sum_r Q(s, r) coef(r, col), col = (combination,
// re/im), 8 columns (complex) or 4 (real).  The pairs (x, y) and (y, x) share
// Q, so the two together fill the 16 columns of one 16x16x4 FP64 matrix
// instruction; it runs as D'(col, s) = coef^T Q^T so that each lane holds the
// re/im parts of its combinations in adjacent accumulator registers.  Only
// the product Q = Dx * Dy is a vector instruction.
//
// Work decomposition: a wavefront owns one y, four consecutive x >= y (the
// pairs (x, y) and, when x < N, (y, x)) and 64 values of the lane index
// (four 16-wide tiles); the four wavefronts of a workgroup take four
// consecutive y with the same x group and lane block, so the Dx tiles they
// read meet in the cache.
#include "kernels.hpp"

namespace so3g {

namespace {

constexpr double kInv8Pi2 = 0.012665147955292221430953884060716700591;  // 1/(8 pi^2)
constexpr int kXG = 4;   // x values per wavefront
constexpr int kNT = 4;   // 16-wide lane-index tiles per wavefront

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__host__ __device__ inline int64_t tri(int64_t k) { return k * (k + 1) / 2; }

// Block order.  A workgroup reads the Delta columns of its 4 x and 4 y; the
// same columns serve every workgroup with the same x group (or y group) and
// lane block, so those should run on one XCD at about the same time and meet
// in its L2 (in plain grid order 2/3 of the reads missed L2).  Workgroups go
// round-robin over the 8 XCDs (block p on XCD p % 8), an XCD holding about
// 64 at once, so the work is cut into tiles of 8 x 8 (x group, y group) of
// one lane block and each XCD runs whole tiles back to back: tile T on XCD
// T % 8.  Only tiles with x >= y hold work (xg >= yg); they are numbered row
// by row (yt, then xt >= yt), lane block (z) outermost, so neighbouring tiles
// of one row, which carry similar work, land on different XCDs.
// Order 1 (chosen per (L, N) by the tuner) keeps the lane block innermost on
// each XCD instead: the XCD runs the nz lane blocks of one tile back to
// back, so the tile's coefficient rows (read once per lane block) are still
// in its L2 for the later ones; the slots of tile q are q % 8 + 8 (nz (q / 8)
// + z).  Both orders compute every output with the same arithmetic.
struct CBlock {
  int xg, yg, z;
  bool ok;
};
__host__ __device__ inline int cblock_tiles(int nxg, int nyg) {  // per lane block
  const int tx = (nxg + 7) / 8, ty = (nyg + 7) / 8;
  int n = 0;
  for (int yt = 0; yt < ty; ++yt) n += tx > yt ? tx - yt : 0;
  return n;
}
__host__ __device__ inline int64_t cblock_grid(int nxg, int nyg, int nz, int order) {
  const int64_t per = cblock_tiles(nxg, nyg);
  if (order == 1) return 64 * 8 * ((per + 7) / 8) * nz;
  const int64_t nt = (int64_t)nz * per;
  return 64 * ((nt + 7) / 8 * 8);
}
__device__ __forceinline__ CBlock cblock(unsigned p, int nxg, int nyg, int nz, int order) {
  const unsigned k = p >> 3;
  const int T = (int)((p & 7) + 8 * (k >> 6)), w = (int)(k & 63);
  const int tx = (nxg + 7) / 8, ty = (nyg + 7) / 8;
  const int per = cblock_tiles(nxg, nyg);
  CBlock b;
  int r;
  if (order == 1) {
    const int j = T >> 3;  // slot on this XCD
    b.z = j % nz;
    r = (T & 7) + 8 * (j / nz);
    if (r >= per) r = per, b.z = nz;  // past the last tile: no work
  } else {
    b.z = T / per;
    r = T - b.z * per;
  }
  int yt = 0;
  while (yt < ty && r >= tx - yt) {  // row yt holds tiles xt = yt .. tx - 1
    r -= tx - yt;
    ++yt;
  }
  const int xt = yt + r;
  b.xg = xt * 8 + (w & 7);
  b.yg = yt * 8 + (w >> 3);
  b.ok = b.z < nz && yt < ty && b.xg < nxg && b.yg < nyg;
  return b;
}

// Operand-row map of the coefficient matrix: lane i supplies row rho = i & 15
// of coef^T at reduction offset i >> 4.  Row rho = 4 j + q of the product
// lands in accumulator register j of the lanes with i >> 4 == q, so
//   q >> 1      : pair half h (0: (x, y), 1: (y, x)),
//   complex (C = 4): combination 2 (q & 1) + (j >> 1), part j & 1,
//   real    (C = 2): combination q & 1, part j & 1 (registers 2, 3 unused).
template <int C>
__device__ __forceinline__ int coef_col(int rho, bool &used) {
  const int j = rho >> 2, q = rho & 3;
  if (C == 4) {
    used = true;
    return (2 * (q & 1) + (j >> 1)) * 2 + (j & 1);
  }
  used = j < 2;
  return (q & 1) * 2 + (j & 1);
}

// ------------------------------------------------------------------ staging
// Operand staging shared by both kernels.  A workgroup of four wavefronts
// owns the pairs (x0 + a, y0 + w), a = 0..3, for wavefront w.  Per
// reduction chunk (4 values of m' forward, of l inverse) it needs
//   DX[a][t]  the Delta tile of column x0 + a, lane tile t   (16 x 64 doubles)
//   DY[w][t]  the Delta tile of column y0 + w, lane tile t   (16 x 64 doubles)
//   CV[w][a][h] four coefficient rows of pair half h           (8 x 4*2C doubles)
// Every one of them is a contiguous run in HBM, so staging is a plain copy:
// each wavefront copies its quarter (DX of a = w, its DY and its CV) with
// 16-byte direct-to-LDS loads into a ring of kRing buffers, kRing - 1 chunks
// ahead of the matrix instructions.  Regions outside the transform are
// copied from a page of zeros, so the inner loop has no per-pair branches.
constexpr int kRing = 3;
constexpr int kBufD = 3 * 1024;  // doubles per ring buffer: DX, DY, CV

typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ void dma16(const double *src, double *lds_base) {
  __builtin_amdgcn_global_load_lds((const void *)src, (lds_void *)lds_base, 16, 0, 0);
}

// wait until at most N vector-memory operations of this wavefront are
// outstanding (expcnt and lgkmcnt left unconstrained)
template <int N>
__device__ __forceinline__ void wait_vm() {
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

template <int C>
struct Stage {
  static constexpr int kRegD = 4 * 2 * C;      // doubles per coefficient region
  static constexpr int kCvI = kRegD / 16;       // copy instructions for CV per wave
  static constexpr int kInstr = 4 + kCvI;       // copy instructions per chunk per wave
};

__device__ __forceinline__ void wait_ring(int ahead, int n) {
  // outstanding copies allowed = chunks issued after the one needed now
  if (n == 6) {
    if (ahead >= 2) wait_vm<12>(); else if (ahead == 1) wait_vm<6>(); else wait_vm<0>();
  } else {
    if (ahead >= 2) wait_vm<10>(); else if (ahead == 1) wait_vm<5>(); else wait_vm<0>();
  }
}

// ------------------------------------------------------------------ inverse
// Workgroup: x group x0 .. x0+3, y0 + w for wavefront w, m' block mpb ..
// mpb+63 (four 16-wide tiles); reduction over l in chunks of 4.
// Block `blk` of the transform's grid; sm is the workgroup's LDS ring
// (kRing * kBufD doubles).  Returns early for blocks without work.
template <int NS>
__device__ __forceinline__ void contract_inv_block(unsigned blk, const Geo &g, const double2 *__restrict__ flmn,
                                                   const int64_t *__restrict__ nbase,
                                                   const double *__restrict__ ti,
                                                   const int64_t *__restrict__ offi, int nmt,
                                                   const double *__restrict__ zero,
                                                   double2 *__restrict__ X1, double *sm) {
  constexpr int C = 2 * NS;
  using St = Stage<C>;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nmb = (g.L + 63) >> 6;
  const int Np = min(g.N, g.L);
  const CBlock cb = cblock(blk, (g.L + kXG - 1) / kXG, (Np + 3) / 4, nmb, g.corder);
  const int mb = cb.z;
  const int x0 = cb.xg * kXG, y0 = cb.yg * 4;
  if (!cb.ok || y0 >= Np || x0 + kXG - 1 < y0 || x0 >= g.L) return;  // whole workgroup
  const int y = y0 + w;
  const bool yv = y < Np;
  const int mpb = mb * 64;

  bool pv[kXG], pp[kXG];
  int pmax[kXG];
#pragma unroll
  for (int a = 0; a < kXG; ++a) {
    const int x = x0 + a;
    pv[a] = yv && x < g.L && x >= y;
    pp[a] = pv[a] && x < Np && x != y;
    pmax[a] = max(x, y);
  }

  // ---- copy streams of this wavefront (per lane: source, lcmin, validity)
  const int e2 = 2 * (lane & 31);
  const int th = lane >> 5;  // tile within a DX / DY copy instruction
  const double *dsrc[4];     // DX (j = 0, 1), DY (j = 2, 3): base so that chunk lc is at + 64 lc
  int dlc[4];                // first valid chunk, INT_MAX if never valid
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int t = 2 * (j & 1) + th;
    const int mp0 = mpb + 16 * t;
    const int c = j < 2 ? x0 + w : y;
    const bool ok = (j < 2 ? c < g.L : yv) && mp0 < g.L;
    const int cc = min(c, g.L - 1), mpc = min(mp0, g.L - 1);
    const int lcmin = max(cc, mpc) >> 2;
    dsrc[j] = ti + offi[(int64_t)cc * nmt + (mpc >> 4)] - 64 * lcmin + e2;
    dlc[j] = ok ? lcmin : 0x7fffffff;
  }
  // coefficient regions: lane i of a region copies f_{l, +-m, +-n} for row
  // l = l0 + i / C and sign combination i % C straight from the caller's
  // flmn; entries outside the transform come from the zero page
  const double *cbase[St::kCvI];  // flmn + 2 (base(n) + m): row l at + 2 l (l + 1)
  int crow[St::kCvI], clmin[St::kCvI], cnj[St::kCvI];
#pragma unroll
  for (int j = 0; j < St::kCvI; ++j) {
    const int lpr = St::kRegD / 2;                // lanes per region
    const int rho = j * (64 / lpr) + lane / lpr;  // region = 2 a + h
    const int a = rho >> 1, h = rho & 1;
    const int x = x0 + a;
    const int li = lane % lpr, combo = li % C;
    const int mi = h == 0 ? x : y, nj = h == 0 ? y : x;
    const int m = (combo & 1) ? -mi : mi, n = (combo >> 1) ? -nj : nj;
    const bool ok = h == 0 ? pv[a] : pp[a];
    const int64_t b = ok ? nbase[n + g.N - 1] : kNoBlock;
    crow[j] = li / C;
    cnj[j] = nj;
    clmin[j] = b == kNoBlock ? 0x7fffffff : max(g.L0, max(mi, nj));
    cbase[j] = (const double *)flmn + 2 * (b == kNoBlock ? 0 : b + m);
  }
  auto issue = [&](int l0, double *buf) {
    const int lc = l0 >> 2;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double *src = lc >= dlc[j] ? dsrc[j] + 64 * lc : zero + e2;
      const int row = j < 2 ? 4 * w + 2 * j : 1024 / 64 + 4 * w + 2 * (j - 2);
      dma16(src, buf + 64 * row);
    }
#pragma unroll
    for (int j = 0; j < St::kCvI; ++j) {
      const int l = l0 + crow[j];
      const bool ok = l >= clmin[j] && l < g.L && (!g.nmodeL || l == cnj[j]);
      const double *src = ok ? cbase[j] + 2 * (int64_t)l * (l + 1) : zero + 2 * (lane % (St::kRegD / 2));
      dma16(src, buf + 2048 + w * 8 * St::kRegD + j * 128);
    }
  };

  // ---- operand reads of this lane
  bool used;
  const int col = coef_col<C>(lane & 15, used);
  const int h = ((lane & 15) & 3) >> 1;
  const int k = lane >> 4;
  const int cvo = 2048 + (w * 8 + h) * St::kRegD + k * (2 * C) + col;  // + 2 a kRegD
  const bool mixed = ((col >> 1) & 1) != ((col >> 1) >> 1);  // combination col / 2: sm != sn

  d4 acc[kXG][kNT];
#pragma unroll
  for (int a = 0; a < kXG; ++a)
#pragma unroll
    for (int t = 0; t < kNT; ++t) acc[a][t] = d4{0.0, 0.0, 0.0, 0.0};

  const int lend = (g.L + 3) & ~3;
  const int lbeg = max(max(y0, x0), mpb) & ~3;
  const int nch = (lend - lbeg) >> 2;
  for (int r = 0; r < kRing - 1 && r < nch; ++r) issue(lbeg + 4 * r, sm + r * kBufD);
  for (int q = 0; q < nch; ++q) {
    // this wavefront's copies of chunk q have landed (those of chunk q + 1
    // may still be in flight) and its reads of chunk q - 1 have returned; the
    // one barrier per chunk then makes chunk q visible to everyone and frees
    // the buffer of chunk q - 1, which the copies of chunk q + 2 refill (no
    // fence: it would wait for the chunk in flight)
    wait_ring(min(nch - 1 - q, kRing - 2), St::kInstr);
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    if (q + kRing - 1 < nch) issue(lbeg + 4 * (q + kRing - 1), sm + ((q + kRing - 1) % kRing) * kBufD);
    const double *buf = sm + (q % kRing) * kBufD;
    const int l0 = lbeg + 4 * q;
    // c_l and the (-1)^l of the mixed-sign combinations on this lane's row
    const int lk = l0 + k;
    double fac = (mixed && (lk & 1)) ? -1.0 : 1.0;
    if (g.clmode == 1) fac *= (2.0 * lk + 1.0) * kInv8Pi2;
    if (!used) fac = 0.0;
    double cv[kXG], dy[kNT];
    // operands read through a volatile LDS pointer: otherwise the compiler
    // pairs the reads of two tiles 512 bytes apart into ds_read2st64_b64,
    // whose two addresses fall in the same banks.  Separate reads of 64
    // consecutive doubles cut the bank-conflict cycles from 1.2 to 0.33 per
    // active LDS cycle and measured 0-4% faster, with 25 fewer VGPRs
    const volatile __attribute__((address_space(3))) double *vb =
        (const volatile __attribute__((address_space(3))) double *)buf;
#pragma unroll
    for (int a = 0; a < kXG; ++a) cv[a] = fac * vb[cvo + 2 * a * St::kRegD];
#pragma unroll
    for (int t = 0; t < kNT; ++t) dy[t] = vb[1024 + (4 * w + t) * 64 + lane];
    // tiles with m' > l or m' >= L were copied from the zero page: the
    // matrix instructions run unconditionally, which keeps the accumulators
    // in place (a branch per tile made the compiler move them every chunk)
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
#pragma unroll
      for (int a = 0; a < kXG; ++a) acc[a][t] = mma(cv[a], vb[(4 * a + t) * 64 + lane] * dy[t], acc[a][t]);
    }
  }
  if (!yv) return;

  // store: lane holds m' = mpb + 16 t + (lane & 15), pair half q >> 1,
  // registers 2 e, 2 e + 1 of its combination.  The X1 row, the power of i
  // and the sign pattern depend on (x, combination) only, so they are set
  // once per pair and the tiles follow (per-store slot arithmetic with
  // divisions was a large part of the kernel's instructions).
  const int q = lane >> 4, hh = q >> 1;
  const int mp0 = mpb + (lane & 15);
#pragma unroll
  for (int a = 0; a < kXG; ++a) {
    if (!(hh == 0 ? pv[a] : pp[a])) continue;
    const int x = x0 + a;
    const int mi = hh == 0 ? x : y, nj = hh == 0 ? y : x;
#pragma unroll
    for (int e = 0; e < (C == 4 ? 2 : 1); ++e) {
      const int c = C == 4 ? 2 * (q & 1) + e : (q & 1);
      const int sm = c & 1, sn = c >> 1;
      if (sm && mi == 0) continue;
      if (sn && nj == 0) continue;
      const int m = sm ? -mi : mi, n = sn ? -nj : nj;
      // |n| < nG and |m| < nA: a negative index wraps by one period
      const int nslot = (g.real || n >= 0) ? n : n + g.nG;
      const int mslot = m >= 0 ? m : m + g.nA;
      double2 *row = X1 + ((int64_t)nslot * g.nA + mslot) * g.L;
      const int ip = (n - m) & 3;
      const bool alt = sm != sn;  // (-1)^{m'} of the mixed-sign combinations
#pragma unroll
      for (int t = 0; t < kNT; ++t) {
        const int mp = mp0 + 16 * t;
        if (mp >= g.L) continue;
        double2 v = make_double2(acc[a][t][2 * e], acc[a][t][2 * e + 1]);
        if (alt && (mp & 1)) v = make_double2(-v.x, -v.y);
        row[mp] = cmul_ipow(v, ip);
      }
    }
  }
}

template <int NS>
__global__ __launch_bounds__(256) void k_contract_inv(Geo g, const double2 *__restrict__ flmn,
                                                      const int64_t *__restrict__ nbase,
                                                      const double *__restrict__ ti,
                                                      const int64_t *__restrict__ offi, int nmt,
                                                      const double *__restrict__ zero,
                                                      double2 *__restrict__ X1) {
  __shared__ __attribute__((aligned(16))) double sm[kRing * kBufD];
  contract_inv_block<NS>(blockIdx.x, g, flmn, nbase, ti, offi, nmt, zero, X1, sm);
}

// X1 rows of |n| in [L, N): no l reaches them
__global__ void k_zero_rows(Geo g, double2 *__restrict__ X1) {
  const int an = g.L + blockIdx.y;
  for (int s = 0; s < (g.real ? 1 : 2); ++s) {
    if (s == 1 && an == 0) continue;
    const int n = s ? -an : an;
    const int nslot = g.real ? n : (n + g.nG) % g.nG;
    double2 *row = X1 + (int64_t)nslot * g.nA * g.L;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)g.nA * g.L;
         i += (int64_t)gridDim.x * blockDim.x)
      row[i] = make_double2(0.0, 0.0);
  }
}

// ------------------------------------------------------------------ forward
// Workgroup: x group x0 .. x0+3, y0 + w for wavefront w, l block lo .. lo+63
// (four 16-wide tiles); reduction over m' in chunks of 4.
template <int NS>
__device__ __forceinline__ void contract_fwd_block(unsigned blk, const Geo &g, const double *__restrict__ Gp,
                                                   const int64_t *__restrict__ nbase,
                                                   const double *__restrict__ tf,
                                                   const int64_t *__restrict__ offf,
                                                   const double *__restrict__ zero,
                                                   double2 *__restrict__ flmn, double *sm) {
  constexpr int C = 2 * NS;
  using St = Stage<C>;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // l blocks are the outermost tile dimension, highest first: their work
  // grows with l, so the longest workgroups start first and the tail is short
  const int nlb = (g.L + 63) >> 6;
  const int Np = min(g.N, g.L);
  const CBlock cb = cblock(blk, (g.L + kXG - 1) / kXG, (Np + 3) / 4, nlb, g.corder);
  const int lb = nlb - 1 - cb.z;
  const int x0 = cb.xg * kXG, y0 = cb.yg * 4;
  const int lo = lb * 64;
  if (!cb.ok || y0 >= Np || x0 + kXG - 1 < y0 || x0 >= g.L || max(x0, y0) > lo + 63) return;
  const int y = y0 + w;
  const bool yv = y < Np;

  bool pv[kXG], pp[kXG];
  int pmax[kXG];
#pragma unroll
  for (int a = 0; a < kXG; ++a) {
    const int x = x0 + a;
    pv[a] = yv && x < g.L && x >= y;
    pp[a] = pv[a] && x < Np && x != y;
    pmax[a] = max(x, y);
  }

  // ---- copy streams of this wavefront
  const int e2 = 2 * (lane & 31);
  const int th = lane >> 5;
  const double *dsrc[4];  // DX (j = 0, 1), DY (j = 2, 3): chunk m0 / 4 at + 16 m0
  int dlast[4];           // last m0 of the tile (m' <= l0 + 15), -1 if none
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int t = 2 * (j & 1) + th;
    const int l0 = lo + 16 * t;
    const int c = j < 2 ? x0 + w : y;
    const bool ok = (j < 2 ? c < g.L : yv) && l0 < g.L && c <= l0 + 15;
    dsrc[j] = ok ? tf + offf[c] - 256 * tri(c >> 4) + 256 * tri(l0 >> 4) + e2 : tf;
    dlast[j] = ok ? l0 + 15 : -1;
  }
  const double *csrc[St::kCvI];
#pragma unroll
  for (int j = 0; j < St::kCvI; ++j) {
    const int lpr = St::kRegD / 2;
    const int rho = j * (64 / lpr) + lane / lpr;
    const int a = rho >> 1, h = rho & 1;
    const int x = x0 + a;
    const bool ok = h == 0 ? pv[a] : pp[a];
    const int mi = h == 0 ? x : y, nj = h == 0 ? y : x;
    // G' is stored [|n|][|m|][combination][m']: lane i of the region copies
    // row (i / C) of combination (i % C), one complex value
    const int li = lane % lpr;
    csrc[j] = ok ? Gp + 2 * ((((int64_t)nj * g.L + mi) * C + li % C) * g.L + li / C) : nullptr;
  }
  auto issue = [&](int m0, double *buf) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double *src = m0 <= dlast[j] ? dsrc[j] + 16 * m0 : zero + e2;
      const int row = j < 2 ? 4 * w + 2 * j : 1024 / 64 + 4 * w + 2 * (j - 2);
      dma16(src, buf + 64 * row);
    }
#pragma unroll
    for (int j = 0; j < St::kCvI; ++j) {
      const double *src = csrc[j] ? csrc[j] + 2 * m0 : zero + 2 * (lane % (St::kRegD / 2));
      dma16(src, buf + 2048 + w * 8 * St::kRegD + j * 128);
    }
  };

  bool used;
  const int col = coef_col<C>(lane & 15, used);
  const int h = ((lane & 15) & 3) >> 1;
  const int k = lane >> 4;
  const int cvo = 2048 + (w * 8 + h) * St::kRegD + k * (2 * C) + col;

  d4 acc[kXG][kNT];
#pragma unroll
  for (int a = 0; a < kXG; ++a)
#pragma unroll
    for (int t = 0; t < kNT; ++t) acc[a][t] = d4{0.0, 0.0, 0.0, 0.0};

  const int nch = (min(lo + 63, g.L - 1) >> 2) + 1;
  for (int r = 0; r < kRing - 1 && r < nch; ++r) issue(4 * r, sm + r * kBufD);
  for (int q = 0; q < nch; ++q) {
    // one barrier per chunk, as in the inverse
    wait_ring(min(nch - 1 - q, kRing - 2), St::kInstr);
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    if (q + kRing - 1 < nch) issue(4 * (q + kRing - 1), sm + ((q + kRing - 1) % kRing) * kBufD);
    const double *buf = sm + (q % kRing) * kBufD;
    const int m0 = 4 * q;
    double cv[kXG], dy[kNT];
    // separate LDS reads, as in the inverse
    const volatile __attribute__((address_space(3))) double *vb =
        (const volatile __attribute__((address_space(3))) double *)buf;
#pragma unroll
    for (int a = 0; a < kXG; ++a) cv[a] = used ? vb[cvo + 2 * a * St::kRegD] : 0.0;
#pragma unroll
    for (int t = 0; t < kNT; ++t) dy[t] = vb[1024 + (4 * w + t) * 64 + lane];
    // tiles with m' > l or l >= L were copied from the zero page: the matrix
    // instructions run unconditionally (see the inverse)
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
#pragma unroll
      for (int a = 0; a < kXG; ++a) acc[a][t] = mma(cv[a], vb[(4 * a + t) * 64 + lane] * dy[t], acc[a][t]);
    }
  }
  if (!yv) return;

  // store: lane holds l = lo + 16 t + (lane & 15), pair half q >> 1,
  // registers 2 e, 2 e + 1 of its combination.  The block base, the power of
  // i and the sign pattern depend on (x, combination) only, so they are set
  // (and the block base loaded) once per pair and the tiles follow.
  const int q = lane >> 4, hh = q >> 1;
  const int l0 = lo + (lane & 15);
#pragma unroll
  for (int a = 0; a < kXG; ++a) {
    if (!(hh == 0 ? pv[a] : pp[a])) continue;
    const int x = x0 + a;
    const int mi = hh == 0 ? x : y, nj = hh == 0 ? y : x;
#pragma unroll
    for (int e = 0; e < (C == 4 ? 2 : 1); ++e) {
      const int c = C == 4 ? 2 * (q & 1) + e : (q & 1);
      const int sm = c & 1, sn = c >> 1;
      if (sm && mi == 0) continue;
      if (sn && nj == 0) continue;
      const int m = sm ? -mi : mi, n = sn ? -nj : nj;
      const int64_t bb = nbase[n + g.N - 1];
      if (bb == kNoBlock) continue;
      double2 *blk = flmn + bb + m;
      const int ip = (m - n) & 3;
      const bool alt = sm != sn;  // (-1)^l of the mixed-sign combinations
      const int lmin = max(max(g.L0, max(mi, nj)), g.nmodeL ? nj : 0);
      const int lmax = g.nmodeL ? nj : g.L - 1;
#pragma unroll
      for (int t = 0; t < kNT; ++t) {
        const int l = l0 + 16 * t;
        if (l < lmin || l > lmax) continue;
        double2 v = make_double2(acc[a][t][2 * e], acc[a][t][2 * e + 1]);
        if (alt && (l & 1)) v = make_double2(-v.x, -v.y);
        if (g.clmode == 2) v = cscale(v, (2.0 * l + 1.0) * kInv8Pi2);
        blk[(int64_t)l * l + l] = cmul_ipow(v, ip);
      }
    }
  }
}

template <int NS>
__global__ __launch_bounds__(256) void k_contract_fwd(Geo g, const double *__restrict__ Gp,
                                                      const int64_t *__restrict__ nbase,
                                                      const double *__restrict__ tf,
                                                      const int64_t *__restrict__ offf,
                                                      const double *__restrict__ zero,
                                                      double2 *__restrict__ flmn) {
  __shared__ __attribute__((aligned(16))) double sm[kRing * kBufD];
  contract_fwd_block<NS>(blockIdx.x, g, Gp, nbase, tf, offf, zero, flmn, sm);
}

// ------------------------------------------------------------------ grouped
// One persistent launch runs the contraction blocks of many transforms (the
// items of a batch).  Item i owns virtual blocks off_i .. off_i + nblk_i - 1,
// in the order of its own grid (nblk_i and off_i are multiples of 8, so a
// virtual block v keeps the XCD v % 8 that the item's block order was laid
// out for).  The work is handed out by eight device queues, one per XCD:
// queue x holds the virtual blocks v = x + 8 k, k = 0, 1, ...; a workgroup
// takes the next k of its own XCD's queue with an atomic add, and once that
// queue is empty it takes from the other seven in turn, so every block runs
// exactly once whatever XCD a workgroup reports, and every workgroup exits
// after seeing all eight queues empty.  The queues are zeroed before each
// launch (k_queue_zero).  Each block runs the per-transform body above, so
// results are bit-identical to the transform's own launch.
__device__ __forceinline__ unsigned xcc_id() {
  unsigned x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  return x & 7u;
}

// thread 0 takes the next virtual block (or ~0u when all queues are empty)
// and finds its item; both are broadcast through LDS
__device__ __forceinline__ void grouped_next(const CItem *__restrict__ items, int nitems, unsigned per_queue,
                                             unsigned *__restrict__ queues, unsigned xcc, int &tried,
                                             unsigned *s_v, int *s_item) {
  if (threadIdx.x == 0) {
    unsigned v = ~0u;
    while (tried < 8) {
      const unsigned x = (xcc + (unsigned)tried) & 7u;
      const unsigned k = atomicAdd(&queues[x], 1u);
      if (k < per_queue) {
        v = x + 8u * k;
        break;
      }
      ++tried;
    }
    int it = 0;
    if (v != ~0u) {
      int lo = 0, hi = nitems - 1;  // last item with off <= v
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].off <= v) lo = mid;
        else hi = mid - 1;
      }
      it = lo;
    }
    *s_v = v;
    *s_item = it;
  }
  __syncthreads();
}

template <int NS, int DIR>
__global__ __launch_bounds__(256) void k_contract_grouped(const CItem *__restrict__ items, int nitems,
                                                          unsigned per_queue, unsigned *__restrict__ queues,
                                                          const double *__restrict__ zero) {
  __shared__ __attribute__((aligned(16))) double sm[kRing * kBufD];
  __shared__ unsigned s_v;
  __shared__ int s_item;
  const unsigned xcc = xcc_id();
  int tried = 0;
  for (;;) {
    grouped_next(items, nitems, per_queue, queues, xcc, tried, &s_v, &s_item);
    const unsigned v = __builtin_amdgcn_readfirstlane(s_v);
    const int i = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();  // every wave has read s_v / s_item before thread 0 writes them again
    if (v == ~0u) break;
    const CItem &it = items[i];
    const Geo g = it.g;
    if (DIR > 0)
      contract_inv_block<NS>(v - it.off, g, (const double2 *)it.in, it.nbase, it.tab, it.off_tab, it.nmt, zero,
                             it.out, sm);
    else
      contract_fwd_block<NS>(v - it.off, g, (const double *)it.in, it.nbase, it.tab, it.off_tab, zero, it.out, sm);
    // the block's last LDS reads are done before the next block's copies land
    __syncthreads();
  }
}

__global__ void k_queue_zero(unsigned *q, int n) {
  if ((int)threadIdx.x < n) q[threadIdx.x] = 0u;
}

}  // namespace

void launch_contract_inv(const Geo &g, const double2 *flmn, NBase nb, const DlTables &dl, double2 *X1,
                         hipStream_t st) {
  const int Np = std::min(g.N, g.L);
  const int nmb = (g.L + 63) / 64;
  const unsigned grid = (unsigned)cblock_grid((g.L + kXG - 1) / kXG, (Np + 3) / 4, nmb, g.corder);
  if (g.NS == 2)
    k_contract_inv<2><<<grid, 256, 0, st>>>(g, flmn, nb.base, dl.ti, dl.offi, dl.nmt, zero_page(), X1);
  else
    k_contract_inv<1><<<grid, 256, 0, st>>>(g, flmn, nb.base, dl.ti, dl.offi, dl.nmt, zero_page(), X1);
  SO3_LAUNCH_CHECK();
  if (g.N > g.L) {
    k_zero_rows<<<dim3(64, g.N - g.L), 256, 0, st>>>(g, X1);
    SO3_LAUNCH_CHECK();
  }
}

void launch_contract_fwd(const Geo &g, const double2 *Gp, NBase nb, const DlTables &dl,
                         double2 *flmn, hipStream_t st) {
  const int Np = std::min(g.N, g.L);
  const int nlb = (g.L + 63) / 64;
  const unsigned grid = (unsigned)cblock_grid((g.L + kXG - 1) / kXG, (Np + 3) / 4, nlb, g.corder);
  if (g.NS == 2)
    k_contract_fwd<2><<<grid, 256, 0, st>>>(g, (const double *)Gp, nb.base, dl.tf, dl.offf, zero_page(), flmn);
  else
    k_contract_fwd<1><<<grid, 256, 0, st>>>(g, (const double *)Gp, nb.base, dl.tf, dl.offf, zero_page(), flmn);
  SO3_LAUNCH_CHECK();
}

unsigned contract_blocks(const Geo &g) {
  const int Np = std::min(g.N, g.L);
  const int nz = (g.L + 63) / 64;
  return (unsigned)cblock_grid((g.L + kXG - 1) / kXG, (Np + 3) / 4, nz, g.corder);
}

void launch_contract_grouped(const CItem *items_dev, int nitems, unsigned total_blocks, int ns, int dir,
                             unsigned *queues, hipStream_t st) {
  if (nitems <= 0 || total_blocks == 0) return;
  if (total_blocks % 8) throw HipError("grouped contraction: block count not a multiple of 8");
  k_queue_zero<<<1, 64, 0, st>>>(queues, 8);
  SO3_LAUNCH_CHECK();
  // persistent grid: the workgroups that fit at once (two per CU), at most
  // one per block
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    SO3_HIP(hipGetDevice(&dev));
    SO3_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  }
  const unsigned grid = std::min<unsigned>(total_blocks, (unsigned)(2 * cus));
  const unsigned per_queue = total_blocks / 8;
  const double *z = zero_page();
  if (dir > 0) {
    if (ns == 2) k_contract_grouped<2, 1><<<grid, 256, 0, st>>>(items_dev, nitems, per_queue, queues, z);
    else k_contract_grouped<1, 1><<<grid, 256, 0, st>>>(items_dev, nitems, per_queue, queues, z);
  } else {
    if (ns == 2) k_contract_grouped<2, -1><<<grid, 256, 0, st>>>(items_dev, nitems, per_queue, queues, z);
    else k_contract_grouped<1, -1><<<grid, 256, 0, st>>>(items_dev, nitems, per_queue, queues, z);
  }
  SO3_LAUNCH_CHECK();
}

void launch_zero_rows(const Geo &g, double2 *X1, hipStream_t st) {
  if (g.N > g.L) {
    k_zero_rows<<<dim3(64, g.N - g.L), 256, 0, st>>>(g, X1);
    SO3_LAUNCH_CHECK();
  }
}

}  // namespace so3g
