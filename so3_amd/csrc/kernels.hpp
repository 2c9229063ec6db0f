// kernels.hpp -- geometry of one transform and the kernel entry points.
#pragma once
#include <cstdint>

#include "dl_tables.hpp"
#include "fft_device.hpp"

namespace so3g {

// Shape of one transform as the kernels see it.
//   n slots: complex n in [-(N-1), N-1] at slot n mod (2N-1); real n in [0, N-1]
//   m slots: m in [-(L-1), L-1] at slot m mod nA (MW_SS slot L is the unused
//            Nyquist column)
struct Geo {
  int L, N, L0;
  int nA, nB, nG;  // alpha / beta samples, n slots
  int ldA;         // row stride of the X2 workspace (>= nA)
  int Pb;          // beta DFT length: 2L-1 (MW) or 2L (MW_SS)
  int ss;          // MW_SS sampling
  int real;        // real signal: n >= 0 only, c2r / r2c over gamma
  int NS;          // n signs handled per |n|: 2 complex, 1 real
  int nmodeL;      // keep only l == |n| (direct transforms with N_MODE_L)
  int steer;       // steerable (via SSHT, complex): N gamma samples at pi g / N
  int clmode;      // (2l+1)/(8 pi^2) factor: 1 in the inverse contraction
                   // (synthesis), 2 in the forward one (adjoint of the
                   // inverse), 0 nowhere (forward, adjoint of the forward)
  int corder;      // contraction workgroup order (0 or 1, see contract.hip)
  int geng;        // gamma engine for 2N-1 in [33, 127]: 0 the default rule,
                   // 1 the workgroup engine, 2 the 256-point register FFT
};

// Per-n data: base[n + N - 1] = storage index of (l=0, m=0, n) in the flmn
// layout (index(l,m,n) = base + l*l + l + m in both storages; negative for
// compact storage), or kNoBlock when block n is not part of the transform.
constexpr int64_t kNoBlock = INT64_MIN;
struct NBase {
  const int64_t *base;
};

// Workspace layouts (double2):
//   X1 rows  : [n_slot][m_slot][m'], m' < L          (inverse contraction out)
//   Gp       : [|n|][|m|][2*NS sign combos][m']        (forward contraction in)
//   X2       : rows of ldA entries (a or m slot), one row per (n slot, b) at
//              x2_row below                           (beta / alpha / gamma)

// X2 rows come in blocks of kX2Blk beta samples: block h = b / kX2Blk holds,
// n slot after n slot, its w = min(kX2Blk, nB - h kX2Blk) rows.  A gamma
// column (b, a) then steps w rows per n slot (64 n slots of one load span a
// few 2 MB pages rather than 64 of them), a beta sequence (n slot, m) steps
// one row per b inside a block, and alpha sees nG nB contiguous rows.
constexpr int kX2Blk = 8;
// (row indices stay below nG nB < 2^31: 32-bit arithmetic, b >= 0)
__host__ __device__ inline int x2_w(const Geo &g, int b) {
  const int r = g.nB - (int)((unsigned)b / kX2Blk) * kX2Blk;
  return r < kX2Blk ? r : kX2Blk;
}
__host__ __device__ inline int64_t x2_row(const Geo &g, int nslot, int b) {
  const int h = (int)((unsigned)b / kX2Blk);
  return (int64_t)(h * kX2Blk * g.nG + nslot * x2_w(g, b) + (int)((unsigned)b % kX2Blk));
}
void launch_contract_inv(const Geo &g, const double2 *flmn, NBase nb, const DlTables &dl, double2 *X1,
                         hipStream_t st);

// One item of a grouped (persistent, queue-fed) contraction launch: the
// transform's geometry, its input (flmn inverse, G' forward) and output (X1
// inverse, flmn forward), n bases, Delta table (ti / offi / nmt inverse,
// tf / offf forward), its block count (contract_blocks) and first virtual
// block (items sorted by off, both multiples of 8).
struct CItem {
  Geo g;
  const void *in;
  double2 *out;
  const int64_t *nbase;
  const double *tab;
  const int64_t *off_tab;
  int nmt;
  unsigned nblk, off;
};
unsigned contract_blocks(const Geo &g);
// items_dev: device array of nitems CItem; queues: 8 device counters
// (zeroed by the launch itself); dir +1 inverse, -1 forward; ns 2 complex,
// 1 real.  The inverse's zero rows (|n| >= L) are not written: see
// launch_zero_rows.
void launch_contract_grouped(const CItem *items_dev, int nitems, unsigned total_blocks, int ns, int dir,
                             unsigned *queues, hipStream_t st);
void launch_zero_rows(const Geo &g, double2 *X1, hipStream_t st);
void launch_contract_fwd(const Geo &g, const double2 *Gp, NBase nb, const DlTables &dl,
                         double2 *flmn, hipStream_t st);

void launch_inv_beta(const Geo &g, const double2 *X1, double2 *X2, NBase nb, const double2 *phase,
                     hipStream_t st);
// Complex transforms whose alpha and gamma DFTs both run on the register
// path take the gamma DFT in place on X2 (whose blocked rows keep a column's
// accesses within a few pages) and the alpha DFT between X2 rows and f rows:
//   inverse: launch_gamma_x2(+1) then launch_alpha_x2f(+1)
//   forward: launch_alpha_fx2(-1) then launch_gamma_x2(-1)
bool reg_order(const Geo &g);
void launch_gamma_x2(const Geo &g, double2 *X2, int dir, double scale, hipStream_t st);
void launch_alpha_x2f(const Geo &g, const double2 *X2, double2 *f, int dir, double scale, hipStream_t st);
void launch_alpha_fx2(const Geo &g, const double2 *f, double2 *X2, int dir, double scale, hipStream_t st);
void launch_alpha(const Geo &g, double2 *X2, int dir, double scale, hipStream_t st);
// nrows DFTs of length n (sign dir) on rows ld apart, results times scale
void launch_rows(double2 *X, int64_t nrows, int n, int ld, int dir, double scale, hipStream_t st);
void launch_inv_gamma(const Geo &g, const double2 *X2, void *f, double scale, hipStream_t st);
void launch_fwd_gamma(const Geo &g, const void *f, double2 *X2, double scale, hipStream_t st);
// gamma of short lengths 2N - 1 <= kGammaShortMax as a direct DFT (one
// column per thread, or per pair of lanes for complex lengths above 31;
// gamma_short.hip): dir = +1 X2 -> f (as launch_inv_gamma), -1 f -> X2 (as
// launch_fwd_gamma); false when the length is not short.
// gamma_short_prepare makes its table ahead of a timed or captured call.
constexpr int kGammaShortMax = 63;
bool gamma_short_applies(const Geo &g);
void gamma_short_prepare(const Geo &g);
bool launch_gamma_short(const Geo &g, const void *src, void *dst, int dir, double scale, hipStream_t st);
// the engine launch_inv_gamma / launch_fwd_gamma run for g when 2N-1 is in
// [33, 127] and the length is not short, i.e. in [65, 127] (1 workgroup, 2
// 256-point register FFT; g.geng overrides the default rule), else 0
int gamma_engine(const Geo &g);
// steerable gamma: dir = +1 X2 -> f (N samples kept), dir = -1 f -> X2
void launch_gamma_steer(const Geo &g, const void *src, void *dst, int dir, double scale, hipStream_t st);
// adjoint != 0: zero-fill instead of the beta extension and no weights (the
// beta stage of the adjoint of the inverse).  wr_even (the weight spectrum
// without the two odd taps, natural order) and pole (the folded responses of
// the pole samples, 2L entries) serve the paired kernels; null for the
// adjoint.
void launch_fwd_beta(const Geo &g, const double2 *X2, double2 *Gp, NBase nb, const double2 *phase,
                     const double2 *wr, const double2 *wr_even,
                     const double2 *pole, int Pw, int adjoint, hipStream_t st);
// beta stage of the adjoint of the forward transform (so3_adjoint.c:730-905):
// m' < 0 by symmetry, weight convolution, phase exp(+i m' pi/(2L-1)),
// backward DFT over m' / (2L-1), adjoint of the extension
void launch_adjf_beta(const Geo &g, const double2 *X1, double2 *X2, NBase nb, const double2 *phase,
                      const double2 *wr, int Pw, hipStream_t st);
// zero the flmn entries a forward transform must write but does not compute
void launch_fwd_zero(const Geo &g, double2 *flmn, NBase nb, int direct, int64_t flmn_size,
                     int storage, int order, hipStream_t st);

// LDS budget per workgroup, bytes
int lds_limit();

}  // namespace so3g
