// dl_tables.hpp -- device-resident Delta^l_{m'm} = d^l_{m'm}(pi/2) tables.
//
// Generated on the GPU in FP64 (Risbo or Trapani recursion), once per
// (device, L, method), and kept for the life of the process.  Two layouts of
// the quarter m', m >= 0 serve the two contraction kernels:
//   inv: for each l and column c, a row over m' of length P64(l+1)
//        (zero beyond l):  inv[inv_off[l] + c*P64(l+1) + m'] = Delta^l_{m',c}
//   fwd: for each (m', c), a segment over l from lb = 64*floor(max(m',c)/64)
//        to L-1 (zero below max(m',c)):
//        fwd[fwd_off[m'*L + c] + (l - lb)] = Delta^l_{m',c}
#pragma once
#include <cstdint>

#include "gpu_common.hpp"

namespace so3g {

struct DlTables {
  int L = 0, method = 0;
  double *inv = nullptr;
  int64_t *inv_off = nullptr;  // device, L+1 entries
  double *fwd = nullptr;
  int64_t *fwd_off = nullptr;  // device, L*L entries
  size_t bytes = 0;
  float build_ms = 0.f;
};

// Cached per (device, L, method); builds on first use (synchronously).
const DlTables &dl_tables(int L, int method);

// Drop every cached table (frees device memory).
void dl_tables_clear();

}  // namespace so3g
