// The simple-graph check of one CSR row, shared by the host driver
// (tch_check_simple, graph_host.cpp) and the serving loop's admission of
// requests (tc_server_submit).  Header-only, no HIP.
#pragma once
#include <algorithm>
#include <cstdint>

namespace tcs {

// Faults of a row, in the order they are looked for.
enum RowFault { kRowOk = 0, kRowRange = 1, kRowUnsorted = 2, kRowSelfLoop = 3, kRowAsym = 4 };

// First fault of row v of a CSR whose offsets are already known to be in
// [0, nnz] and non-decreasing: an id outside [0, n), ids not strictly
// increasing, a self-loop, or an arc v -> w without w -> v.
inline int row_fault(int32_t n, const int32_t* off, const int32_t* ind, int64_t v) {
  for (int64_t k = off[v]; k < off[v + 1]; ++k) {
    const int32_t w = ind[k];
    if (w < 0 || w >= n) return kRowRange;
    if (k > off[v] && ind[k - 1] >= w) return kRowUnsorted;
    if (w == v) return kRowSelfLoop;
  }
  for (int64_t k = off[v]; k < off[v + 1]; ++k) {
    const int32_t w = ind[k];
    if (!std::binary_search(ind + off[w], ind + off[w + 1], (int32_t)v)) return kRowAsym;
  }
  return kRowOk;
}

// 0 when off/ind form a simple undirected graph on n vertices with nnz
// arcs (offsets from 0 to nnz, non-decreasing; every row fault-free), else
// nonzero.
inline int csr_fault(int32_t n, int64_t nnz, const int32_t* off, const int32_t* ind) {
  if (n < 0 || nnz < 0 || off[0] != 0 || off[n] != nnz) return kRowRange;
  for (int64_t v = 0; v < n; ++v)
    if (off[v + 1] < off[v]) return kRowRange;
  for (int64_t v = 0; v < n; ++v) {
    const int f = row_fault(n, off, ind, v);
    if (f) return f;
  }
  return kRowOk;
}

}  // namespace tcs
