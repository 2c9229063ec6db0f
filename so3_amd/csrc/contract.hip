// contract.hip -- the FP64 Delta-Delta contraction of the Wigner transform.
//
// Inverse (reference direct form so3_core.c:1081-1114, via-SSHT form inside
// ssht_core_mw_lb_inverse_sov_sym called at so3_core.c:204):
//   F(m,n,m') = i^{n-m} sum_{l >= max(L0,|m|,|n|,m')} c_l D^l_{m'm} D^l_{m'n} f_lmn,
//   c_l = (2l+1)/(8 pi^2), for 0 <= m' < L.
// Forward (so3_core.c:1589-1617, ssht_core_mw_lb_forward_sov_conv_sym at :410):
//   f_lmn = i^{m-n} sum_{0 <= m' <= l} D^l_{m'm} D^l_{m'n} G'(m,n,m').
//
// Both kernels use one wavefront per workgroup and treat the four sign
// combinations (+-m, +-n) of a (|m|, |n|) pair together: with
//   D^l_{m',-m} = (-1)^{l+m'} D^l_{m',m},
// one product D^l_{m'|m|} D^l_{m'|n|} feeds four complex FMAs.  The factor
// (-1)^l of the mixed-sign pairs is applied per l (l unrolled by two), the
// factor (-1)^{m'} once at the end (inverse) or folded into G' by the
// forward beta stage.
//   inverse: lanes run over m' (64 per wave), the loop over l accumulates in
//            registers; f_lmn is wave-uniform and read with scalar loads.
//   forward: lanes run over l, the loop over m' accumulates in registers;
//            G' is wave-uniform and read with scalar loads.
// Delta columns come from the two layouts of dl_tables.hpp so that every
// vector load is a contiguous 512-byte row.
#include "kernels.hpp"

namespace so3g {

namespace {

typedef const __attribute__((address_space(4))) double cdouble;
typedef const __attribute__((address_space(4))) int64_t cint64;

constexpr double kInv8Pi2 = 0.012665147955292221430953884060716700591;  // 1/(8 pi^2)

template <int A, int B, int NS, bool ODD>
__device__ __forceinline__ void inv_step(int l, const Geo &g, const double *__restrict__ Dinv,
                                         int64_t offl, int mp, int m0, int n0, cdouble *fl,
                                         const int64_t (&bas)[B][NS], double2 (&acc)[A][B][2 * NS]) {
  constexpr int C = 2 * NS;
  const int rowlen = (l + 64) & ~63;
  const double *Dl = Dinv + offl + mp;
  double dm[A], dn[B];
#pragma unroll
  for (int i = 0; i < A; ++i) dm[i] = Dl[(int64_t)min(m0 + i, l) * rowlen];
  const double cl = (2.0 * l + 1.0) * kInv8Pi2;
#pragma unroll
  for (int j = 0; j < B; ++j) dn[j] = cl * Dl[(int64_t)min(n0 + j, l) * rowlen];
  const int64_t ll = (int64_t)l * l + l;
#pragma unroll
  for (int i = 0; i < A; ++i) {
    const int mi = m0 + i;
    if (mi > l) continue;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int nj = n0 + j;
      if (nj > l || nj >= g.N) continue;
      if (g.nmodeL && nj != l) continue;
      const double w = dm[i] * dn[j];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int sm = c & 1, sn = c >> 1;
        const int64_t b = bas[j][sn];
        if (b == kNoBlock) continue;
        const int64_t q = 2 * (b + ll + (sm ? -mi : mi));
        const double2 v = make_double2(fl[q], fl[q + 1]);
        if (ODD && (sm != sn)) {
          acc[i][j][c].x -= w * v.x;
          acc[i][j][c].y -= w * v.y;
        } else {
          acc[i][j][c].x += w * v.x;
          acc[i][j][c].y += w * v.y;
        }
      }
    }
  }
}

template <int A, int B, int NS>
__global__ __launch_bounds__(64) void k_contract_inv(Geo g, const double2 *__restrict__ flmn,
                                                     const int64_t *__restrict__ nbase,
                                                     const double *__restrict__ Dinv,
                                                     const int64_t *__restrict__ inv_off,
                                                     double2 *__restrict__ X1) {
  constexpr int C = 2 * NS;
  const int lane = threadIdx.x;
  const int mpb = (g.L + 63) >> 6;
  const int mb = blockIdx.x % mpb;
  const int mt = blockIdx.x / mpb;
  const int m0 = mt * A, n0 = blockIdx.y * B;
  const int mp = mb * 64 + lane;
  cdouble *fl = (cdouble *)flmn;
  cint64 *nbs = (cint64 *)nbase;
  cint64 *ioff = (cint64 *)inv_off;
  int64_t bas[B][NS];
#pragma unroll
  for (int j = 0; j < B; ++j) {
    const int nj = n0 + j;
    bas[j][0] = nj < g.N ? nbs[nj + g.N - 1] : kNoBlock;
    if (NS == 2) bas[j][NS - 1] = nj < g.N ? nbs[-nj + g.N - 1] : kNoBlock;
  }
  double2 acc[A][B][C];
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j)
#pragma unroll
      for (int c = 0; c < C; ++c) acc[i][j][c] = make_double2(0.0, 0.0);

  int l = max(max(g.L0, mb * 64), max(m0, n0));
  if (l < g.L && (l & 1)) {
    inv_step<A, B, NS, true>(l, g, Dinv, ioff[l], mp, m0, n0, fl, bas, acc);
    ++l;
  }
  for (; l + 1 < g.L; l += 2) {
    inv_step<A, B, NS, false>(l, g, Dinv, ioff[l], mp, m0, n0, fl, bas, acc);
    inv_step<A, B, NS, true>(l + 1, g, Dinv, ioff[l + 1], mp, m0, n0, fl, bas, acc);
  }
  if (l < g.L) inv_step<A, B, NS, false>(l, g, Dinv, ioff[l], mp, m0, n0, fl, bas, acc);

  if (mp >= g.L) return;
#pragma unroll
  for (int i = 0; i < A; ++i) {
    const int mi = m0 + i;
    if (mi >= g.L) continue;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int nj = n0 + j;
      if (nj >= g.N) continue;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int sm = c & 1, sn = c >> 1;
        if (sm && mi == 0) continue;
        if (sn && nj == 0) continue;
        const int m = sm ? -mi : mi, n = sn ? -nj : nj;
        double2 v = acc[i][j][c];
        if ((sm != sn) && (mp & 1)) v = make_double2(-v.x, -v.y);
        v = cmul_ipow(v, n - m);
        const int nslot = g.real ? n : (n + g.nG) % g.nG;
        const int mslot = (m + g.nA) % g.nA;
        X1[((int64_t)nslot * g.nA + mslot) * g.L + mp] = v;
      }
    }
  }
}

template <int A, int B, int NS>
__global__ __launch_bounds__(64) void k_contract_fwd(Geo g, const double2 *__restrict__ Gp,
                                                     const int64_t *__restrict__ nbase,
                                                     const double *__restrict__ Dfwd,
                                                     const int64_t *__restrict__ fwd_off,
                                                     double2 *__restrict__ flmn) {
  constexpr int C = 2 * NS;
  const int lane = threadIdx.x;
  const int lbn = (g.L + 63) >> 6;
  const int lb = blockIdx.x % lbn;
  const int mt = blockIdx.x / lbn;
  const int m0 = mt * A, n0 = blockIdx.y * B;
  const int lo = lb * 64;
  const int l = lo + lane;
  const int lc = min(l, g.L - 1);
  cdouble *G = (cdouble *)Gp;
  cint64 *foff = (cint64 *)fwd_off;
  double2 acc[A][B][C];
#pragma unroll
  for (int i = 0; i < A; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j)
#pragma unroll
      for (int c = 0; c < C; ++c) acc[i][j][c] = make_double2(0.0, 0.0);

  const int mpmax = min(g.L - 1, lo + 63);
  for (int mp = 0; mp <= mpmax; ++mp) {
    double dm[A], dn[B];
#pragma unroll
    for (int i = 0; i < A; ++i) {
      const int c = m0 + i;
      const int mx = max(mp, c);
      dm[i] = (c < g.L && mx <= lo + 63) ? Dfwd[foff[(int64_t)mp * g.L + c] + (lc - (mx & ~63))] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int c = n0 + j;
      const int mx = max(mp, c);
      dn[j] = (c < g.N && c < g.L && mx <= lo + 63) ? Dfwd[foff[(int64_t)mp * g.L + c] + (lc - (mx & ~63))] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < A; ++i) {
      const int mi = m0 + i;
      if (mi >= g.L) continue;
#pragma unroll
      for (int j = 0; j < B; ++j) {
        const int nj = n0 + j;
        if (nj >= g.N) continue;
        const double w = dm[i] * dn[j];
        const int64_t gb = (((int64_t)nj * g.L + mi) * g.L + mp) * C;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const double2 v = make_double2(G[2 * (gb + c)], G[2 * (gb + c) + 1]);
          acc[i][j][c].x += w * v.x;
          acc[i][j][c].y += w * v.y;
        }
      }
    }
  }
  if (l >= g.L) return;
  cint64 *nbs = (cint64 *)nbase;
  const int64_t ll = (int64_t)l * l + l;
#pragma unroll
  for (int i = 0; i < A; ++i) {
    const int mi = m0 + i;
    if (mi >= g.L || mi > l) continue;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int nj = n0 + j;
      if (nj >= g.N || nj > l || l < g.L0) continue;
      if (g.nmodeL && nj != l) continue;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int sm = c & 1, sn = c >> 1;
        if (sm && mi == 0) continue;
        if (sn && nj == 0) continue;
        const int m = sm ? -mi : mi, n = sn ? -nj : nj;
        const int64_t b = nbs[n + g.N - 1];
        if (b == kNoBlock) continue;
        double2 v = acc[i][j][c];
        if ((sm != sn) && (l & 1)) v = make_double2(-v.x, -v.y);
        flmn[b + ll + m] = cmul_ipow(v, m - n);
      }
    }
  }
}

constexpr int kA = 2, kB = 4;

}  // namespace

void launch_contract_inv(const Geo &g, const double2 *flmn, NBase nb, const DlTables &dl,
                         double2 *X1, hipStream_t st) {
  const int mpb = (g.L + 63) / 64;
  dim3 grid(mpb * ((g.L + kA - 1) / kA), (g.N + kB - 1) / kB);
  if (g.NS == 2)
    k_contract_inv<kA, kB, 2><<<grid, 64, 0, st>>>(g, flmn, nb.base, dl.inv, dl.inv_off, X1);
  else
    k_contract_inv<kA, kB, 1><<<grid, 64, 0, st>>>(g, flmn, nb.base, dl.inv, dl.inv_off, X1);
  SO3_LAUNCH_CHECK();
}

void launch_contract_fwd(const Geo &g, const double2 *Gp, NBase nb, const DlTables &dl,
                         double2 *flmn, hipStream_t st) {
  const int lbn = (g.L + 63) / 64;
  dim3 grid(lbn * ((g.L + kA - 1) / kA), (g.N + kB - 1) / kB);
  if (g.NS == 2)
    k_contract_fwd<kA, kB, 2><<<grid, 64, 0, st>>>(g, Gp, nb.base, dl.fwd, dl.fwd_off, flmn);
  else
    k_contract_fwd<kA, kB, 1><<<grid, 64, 0, st>>>(g, Gp, nb.base, dl.fwd, dl.fwd_off, flmn);
  SO3_LAUNCH_CHECK();
}

}  // namespace so3g
