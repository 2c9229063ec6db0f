// contract.hip -- the FP64 Delta-Delta contraction of the Wigner transform.
//
// Inverse (reference direct form so3_core.c:1081-1114, via-SSHT form inside
// ssht_core_mw_lb_inverse_sov_sym called at so3_core.c:204):
//   F(m,n,m') = i^{n-m} sum_{l >= max(L0,|m|,|n|,m')} c_l D^l_{m'm} D^l_{m'n} f_lmn,
//   c_l = (2l+1)/(8 pi^2), for 0 <= m' < L.
// Forward (so3_core.c:1589-1617, ssht_core_mw_lb_forward_sov_conv_sym at :410):
//   f_lmn = i^{m-n} sum_{0 <= m' <= l} D^l_{m'm} D^l_{m'n} G'(m,n,m').
//
// The four sign combinations (+-m, +-n) of a pair (|m|, |n|) are treated
// together: with D^l_{m',-m} = (-1)^{l+m'} D^l_{m',m}, one product
// D^l_{m'|m|} D^l_{m'|n|} feeds four complex FMAs.  The factor (-1)^l of the
// mixed-sign combinations sits in the packed coefficients (inverse) or is
// applied at the store (forward); (-1)^{m'} is applied at the store
// (inverse) or folded into G' by the forward beta stage.
//
// Work decomposition (both directions): a workgroup of four wavefronts owns
// a tile of 4 |m| x 4 |n| pairs and one 64-wide block of the lane index
// (m' for the inverse, l for the forward); each wavefront owns 2 x 2 pairs.
// Per step of the reduction loop (l for the inverse, m' for the forward)
// the workgroup stages the 8 Delta columns of the tile (4 |m| + 4 |n|, one
// 512-byte row each) in LDS, prefetched one step ahead, and every wavefront
// reads the coefficients of its pairs as one 64-byte scalar load per pair
// from a pair-contiguous stream, so f_lmn / G' never occupy vector registers.
#include "kernels.hpp"

namespace so3g {

namespace {

typedef const __attribute__((address_space(4))) double cdouble;

constexpr double kInv8Pi2 = 0.012665147955292221430953884060716700591;  // 1/(8 pi^2)
constexpr int kTile = 4;     // |m| and |n| per workgroup tile
constexpr int kWaves = 4;    // wavefronts per workgroup (2 x 2 pairs each)
constexpr int kCols = 2 * kTile;

__device__ __forceinline__ double2 ld2(cdouble *p, int64_t i) { return make_double2(p[2 * i], p[2 * i + 1]); }

// ------------------------------------------------------------------ packing
// Gi[(off(an,am) + (l - l0(an,am))) * C + c], l0 = max(an & ~3, am & ~3):
// c_l f_{l, +-am, +-an} with the (-1)^l of mixed signs folded in, zero where
// the coefficient is not part of the transform.
__global__ void k_pack_inv(Geo g, const double2 *__restrict__ flmn, const int64_t *__restrict__ nbase,
                           const int64_t *__restrict__ poff, double2 *__restrict__ Gi) {
  const int an = blockIdx.y;
  const int am = blockIdx.x;
  const int C = 2 * g.NS;
  const int l0 = max(an & ~3, am & ~3);
  const int64_t off = poff[(int64_t)an * g.L + am];
  for (int t = threadIdx.x; t < (g.L - l0) * C; t += blockDim.x) {
    const int l = l0 + t / C, c = t % C;
    const int sm = c & 1, sn = c >> 1;
    const int m = sm ? -am : am, n = sn ? -an : an;
    double2 v = make_double2(0.0, 0.0);
    const bool ok = an < g.N && l >= g.L0 && l >= am && l >= an && (!g.nmodeL || l == an);
    if (ok) {
      const int64_t b = nbase[n + g.N - 1];
      if (b != kNoBlock) {
        v = flmn[b + (int64_t)l * l + l + m];
        double s = g.clmode == 1 ? (2.0 * l + 1.0) * kInv8Pi2 : 1.0;
        if ((sm != sn) && (l & 1)) s = -s;
        v = cscale(v, s);
      }
    }
    Gi[(off + (l - l0)) * C + c] = v;
  }
}

// ------------------------------------------------------------------ inverse
template <int NS>
__global__ __launch_bounds__(256) void k_contract_inv(Geo g, const double2 *__restrict__ Gi,
                                                      const int64_t *__restrict__ poff,
                                                      const double *__restrict__ Dinv,
                                                      const int64_t *__restrict__ inv_off,
                                                      double2 *__restrict__ X1) {
  constexpr int C = 2 * NS;
  __shared__ double2 sD[2][kCols][32];  // [buffer][column][m' pairs] = 64 doubles per column
  const int tid = threadIdx.x;
  // the wavefront index is uniform: tell the compiler, so that everything
  // derived from it lives in scalar registers and feeds scalar loads
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int mpb = (g.L + 63) >> 6;
  const int mb = blockIdx.x % mpb;
  const int mt = blockIdx.x / mpb;
  const int m0 = mt * kTile, n0 = blockIdx.y * kTile;
  const int mp = mb * 64 + lane;
  // this wavefront's 2 x 2 pairs
  const int am0 = m0 + 2 * (w & 1), an0 = n0 + 2 * (w >> 1);
  cdouble *G = (cdouble *)Gi;
  const int lt = max(m0, n0);  // first l of the tile's coefficient streams
  int64_t gofs[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int am = am0 + a, an = an0 + b;
      gofs[a][b] = (am < g.L && an < g.N) ? poff[(int64_t)an * g.L + am] : -1;
    }
  double2 acc[2][2][C];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < C; ++c) acc[a][b][c] = make_double2(0.0, 0.0);

  // staging map: thread -> (column, two consecutive m')
  const int col = tid >> 5, pr = tid & 31;
  const int colv = col < kTile ? m0 + col : n0 + (col - kTile);
  auto fetch = [&](int l) -> double2 {
    const int rowlen = (l + 64) & ~63;
    const int cc = min(colv, l);
    const double *src = Dinv + inv_off[l] + (int64_t)cc * rowlen + mb * 64 + 2 * pr;
    return *(const double2 *)src;
  };
  const int lbeg = max(max(g.L0, mb * 64), lt);
  double2 nxt = make_double2(0.0, 0.0);
  if (lbeg < g.L) nxt = fetch(lbeg);
  int buf = 0;
  for (int l = lbeg; l < g.L; ++l) {
    sD[buf][col][pr] = nxt;
    __syncthreads();
    if (l + 1 < g.L) nxt = fetch(l + 1);
    const double *col_d = (const double *)&sD[buf][0][0];
    double dm[2], dn[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) dm[a] = col_d[(2 * (w & 1) + a) * 64 + lane];
#pragma unroll
    for (int b = 0; b < 2; ++b) dn[b] = col_d[(kTile + 2 * (w >> 1) + b) * 64 + lane];
    const int64_t li = l - lt;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (gofs[a][b] < 0) continue;
        const double wgt = dm[a] * dn[b];
        const int64_t q = (gofs[a][b] + li) * C;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const double2 v = ld2(G, q + c);
          acc[a][b][c].x += wgt * v.x;
          acc[a][b][c].y += wgt * v.y;
        }
      }
    buf ^= 1;
  }

  if (mp >= g.L) return;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int mi = am0 + a;
    if (mi >= g.L) continue;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int nj = an0 + b;
      if (nj >= g.N) continue;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int sm = c & 1, sn = c >> 1;
        if (sm && mi == 0) continue;
        if (sn && nj == 0) continue;
        const int m = sm ? -mi : mi, n = sn ? -nj : nj;
        double2 v = acc[a][b][c];
        if ((sm != sn) && (mp & 1)) v = make_double2(-v.x, -v.y);
        v = cmul_ipow(v, n - m);
        const int nslot = g.real ? n : (n + g.nG) % g.nG;
        const int mslot = (m + g.nA) % g.nA;
        X1[((int64_t)nslot * g.nA + mslot) * g.L + mp] = v;
      }
    }
  }
}

// ------------------------------------------------------------------ forward
template <int NS>
__global__ __launch_bounds__(256) void k_contract_fwd(Geo g, const double2 *__restrict__ Gp,
                                                      const int64_t *__restrict__ nbase,
                                                      const double *__restrict__ Dfwd,
                                                      const int64_t *__restrict__ fwd_off,
                                                      double2 *__restrict__ flmn) {
  constexpr int C = 2 * NS;
  __shared__ double2 sD[2][kCols][32];
  const int tid = threadIdx.x;
  // the wavefront index is uniform: tell the compiler, so that everything
  // derived from it lives in scalar registers and feeds scalar loads
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lbn = (g.L + 63) >> 6;
  const int lb = blockIdx.x % lbn;
  const int mt = blockIdx.x / lbn;
  const int m0 = mt * kTile, n0 = blockIdx.y * kTile;
  const int lo = lb * 64;
  if (max(m0, n0) > lo + 63) return;  // no l of this block reaches the tile
  const int l = lo + lane;
  const int am0 = m0 + 2 * (w & 1), an0 = n0 + 2 * (w >> 1);
  cdouble *G = (cdouble *)Gp;
  double2 acc[2][2][C];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < C; ++c) acc[a][b][c] = make_double2(0.0, 0.0);
  bool pv[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) pv[a][b] = (am0 + a < g.L) && (an0 + b < g.N) && (an0 + b < g.L);

  const int col = tid >> 5, pr = tid & 31;
  const int colv = col < kTile ? m0 + col : n0 + (col - kTile);
  const bool colok = colv < g.L && (col < kTile || colv < g.N);
  // Delta^l_{m', colv} for l = lo + 2 pr + {0, 1}
  auto fetch = [&](int mp) -> double2 {
    const int mx = max(mp, colv);
    if (!colok || mx > lo + 63) return make_double2(0.0, 0.0);
    const int64_t base = fwd_off[(int64_t)mp * g.L + colv] - (mx & ~63);
    const int l0 = lo + 2 * pr;
    const int la = min(l0, g.L - 1), lbb = min(l0 + 1, g.L - 1);
    return make_double2(Dfwd[base + la], Dfwd[base + lbb]);
  };
  const int mpmax = min(g.L - 1, lo + 63);
  double2 nxt = fetch(0);
  int buf = 0;
  for (int mp = 0; mp <= mpmax; ++mp) {
    sD[buf][col][pr] = nxt;
    __syncthreads();
    if (mp + 1 <= mpmax) nxt = fetch(mp + 1);
    const double *col_d = (const double *)&sD[buf][0][0];
    double dm[2], dn[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) dm[a] = col_d[(2 * (w & 1) + a) * 64 + lane];
#pragma unroll
    for (int b = 0; b < 2; ++b) dn[b] = col_d[(kTile + 2 * (w >> 1) + b) * 64 + lane];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (!pv[a][b]) continue;
        const double wgt = dm[a] * dn[b];
        const int64_t q = ((((int64_t)(an0 + b) * g.L + (am0 + a)) * g.L) + mp) * C;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const double2 v = ld2(G, q + c);
          acc[a][b][c].x += wgt * v.x;
          acc[a][b][c].y += wgt * v.y;
        }
      }
    buf ^= 1;
  }

  if (l >= g.L || l < g.L0) return;
  const int64_t ll = (int64_t)l * l + l;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int mi = am0 + a;
    if (!(mi < g.L) || mi > l) continue;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int nj = an0 + b;
      if (!pv[a][b] || nj > l) continue;
      if (g.nmodeL && nj != l) continue;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int sm = c & 1, sn = c >> 1;
        if (sm && mi == 0) continue;
        if (sn && nj == 0) continue;
        const int m = sm ? -mi : mi, n = sn ? -nj : nj;
        const int64_t bb = nbase[n + g.N - 1];
        if (bb == kNoBlock) continue;
        double2 v = acc[a][b][c];
        if ((sm != sn) && (l & 1)) v = make_double2(-v.x, -v.y);
        if (g.clmode == 2) v = cscale(v, (2.0 * l + 1.0) * kInv8Pi2);
        flmn[bb + ll + m] = cmul_ipow(v, m - n);
      }
    }
  }
}

}  // namespace

int64_t pack_inv_size(const Geo &g, int64_t *poff_host) {
  int64_t acc = 0;
  for (int an = 0; an < g.N; ++an)
    for (int am = 0; am < g.L; ++am) {
      if (poff_host) poff_host[(int64_t)an * g.L + am] = acc;
      const int l0 = std::max(an & ~3, am & ~3);
      acc += std::max(0, g.L - l0);
    }
  return acc * 2 * g.NS;  // double2 elements
}

void launch_contract_inv(const Geo &g, const double2 *flmn, NBase nb, const DlTables &dl,
                         const int64_t *poff, double2 *Gi, double2 *X1, hipStream_t st) {
  k_pack_inv<<<dim3(g.L, g.N), 256, 0, st>>>(g, flmn, nb.base, poff, Gi);
  SO3_LAUNCH_CHECK();
  const int mpb = (g.L + 63) / 64;
  dim3 grid(mpb * ((g.L + kTile - 1) / kTile), (g.N + kTile - 1) / kTile);
  if (g.NS == 2)
    k_contract_inv<2><<<grid, 256, 0, st>>>(g, Gi, poff, dl.inv, dl.inv_off, X1);
  else
    k_contract_inv<1><<<grid, 256, 0, st>>>(g, Gi, poff, dl.inv, dl.inv_off, X1);
  SO3_LAUNCH_CHECK();
}

void launch_contract_fwd(const Geo &g, const double2 *Gp, NBase nb, const DlTables &dl,
                         double2 *flmn, hipStream_t st) {
  const int lbn = (g.L + 63) / 64;
  dim3 grid(lbn * ((g.L + kTile - 1) / kTile), (g.N + kTile - 1) / kTile);
  if (g.NS == 2)
    k_contract_fwd<2><<<grid, 256, 0, st>>>(g, Gp, nb.base, dl.fwd, dl.fwd_off, flmn);
  else
    k_contract_fwd<1><<<grid, 256, 0, st>>>(g, Gp, nb.base, dl.fwd, dl.fwd_off, flmn);
  SO3_LAUNCH_CHECK();
}

}  // namespace so3g
