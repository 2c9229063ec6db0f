// dl_tables.hpp -- device-resident Delta^l_{m'm} = d^l_{m'm}(pi/2) tables.
//
// Generated on the GPU in FP64 (Risbo or Trapani recursion), once per
// (device, L, method), and kept for the life of the process.  The quarter
// m', c >= 0 is stored twice, each time as ready-made operand tiles of the
// 16x16x4 FP64 matrix instruction (64 doubles = one 512-byte load per
// wavefront, lane i taking element i), zero where l < max(m', c) or l >= L:
//   tf (forward, lane index l, reduction over m'), for column c, l-tile
//      lt >= c/16 (l = 16 lt + (i & 15)) and m'-chunk mc <= 4 lt + 3
//      (m' = 4 mc + (i >> 4)):
//        tf[offf[c] + 256 (T(lt) - T(c >> 4)) + 64 mc + i],  T(k) = k(k+1)/2
//   ti (inverse, lane index m', reduction over l), for column c, m'-tile mt
//      (m' = 16 mt + (i & 15)) and l-chunk lc >= lcmin = max(c, 16 mt) / 4
//      (l = 4 lc + (i >> 4)):
//        ti[offi[c * nmt + mt] + 64 (lc - lcmin) + i]
#pragma once
#include <cstdint>

#include "gpu_common.hpp"

namespace so3g {

struct DlTables {
  int L = 0, method = 0;
  int nmt = 0;                 // ceil(L / 16)
  double *tf = nullptr;
  int64_t *offf = nullptr;     // device, L + 1 entries
  double *ti = nullptr;
  int64_t *offi = nullptr;     // device, L * nmt entries
  size_t bytes = 0;
  float build_ms = 0.f;
};

// Cached per (device, table L, method).  A table built for L' serves every
// L <= L' with bit-identical results (the smallest cached L' >= L is used),
// so a mix of sizes needs one table per method, at its largest L.  Builds on
// first use (synchronously).  The cache keeps at most dl_tables_limit bytes
// of tables (32 GiB by default): building a table that would exceed it first
// waits for the device and drops the least recently used unpinned tables.
const DlTables &dl_tables(int L, int method);
void dl_tables_limit(size_t bytes);
// The table that serves L, or nullptr (never builds, never allocates).
const DlTables *dl_tables_find(int L, int method);
// Pin the table that serves L (returns its table L) / unpin it: pinned
// tables are never evicted (captured graphs hold their addresses).
int dl_tables_pin(int L, int method);
void dl_tables_unpin(int table_L, int method);

// Drop every cached table and zero page (frees device memory).
void dl_tables_clear();

// A device page of kZeroPage zero doubles on the current device: the copy
// source of operand regions that lie outside a transform.
constexpr int kZeroPage = 1024;
const double *zero_page();

}  // namespace so3g
