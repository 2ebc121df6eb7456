// The simple-graph check of one CSR row (tch_check_simple and
// tch_check_simple64, graph_host.cpp).  Header-only, no HIP.
#pragma once
#include <algorithm>
#include <cstdint>

namespace tcs {

// Faults of a row, in the order they are looked for.
enum RowFault { kRowOk = 0, kRowRange = 1, kRowUnsorted = 2, kRowSelfLoop = 3, kRowAsym = 4 };

// First fault of row v of a CSR whose offsets are already known to be in
// [0, nnz] and non-decreasing: an id outside [0, n), ids not strictly
// increasing, a self-loop, or an arc v -> w without w -> v.  OT and IT are
// the offset and id types (int32 or int64).
template <class OT, class IT>
inline int row_fault(int64_t n, const OT* off, const IT* ind, int64_t v) {
  for (int64_t k = off[v]; k < (int64_t)off[v + 1]; ++k) {
    const IT w = ind[k];
    if (w < 0 || (int64_t)w >= n) return kRowRange;
    if (k > (int64_t)off[v] && ind[k - 1] >= w) return kRowUnsorted;
    if ((int64_t)w == v) return kRowSelfLoop;
  }
  for (int64_t k = off[v]; k < (int64_t)off[v + 1]; ++k) {
    const IT w = ind[k];
    if (!std::binary_search(ind + off[w], ind + off[w + 1], (IT)v)) return kRowAsym;
  }
  return kRowOk;
}

}  // namespace tcs
