// Internal interface between the C-ABI layer (tc_capi.hip) and the kernels
// (tc_kernels.hip).  Not part of the public API.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace tcd {

// Arguments of the count kernels.  The edge list is built on the device:
// edge e joins global vertices eu[e] < ew[e] (or an oriented arc in the
// total-only mode); wpre is the exclusive prefix of per-edge work
// |N(u)| + |N(w)| with wpre[E] = W; dscal[0] = E, dscal[1] = W.
struct CountArgs {
  const int64_t* wpre;
  const int32_t* eu;
  const int32_t* ew;
  const int64_t* voff;  // global row offsets [nv + 1]
  const int32_t* ind;   // neighbour ids (graph-local in batches)
  const int64_t* dscal;
  int32_t* out;                 // per-vertex counts (nullable in total-only mode)
  unsigned long long* total;    // int64 total (nullable)
  int32_t total_mult;           // total += total_mult * |N(u) ∩ N(w)| per edge
  int32_t small_max;
  int32_t max_ranges;           // level-1 ranges: min(W / min_units, max_ranges)
  int32_t min_units;            // smallest level-1 range
  unsigned int* qctr;           // 8 per-XCD range queue counters, zeroed per count
  int32_t wave_acap;            // streamed kernel: longest shorter-list a single wave stages
  int32_t search_ratio;         // edges with |B| >= search_ratio * |A| go to k_count_search (0: none)
  int32_t u_only;               // credit each edge's count to out[eu] only (vertex-list mode)
  int32_t probe_ratio;          // wave phase: edges with |B| >= probe_ratio * |A| walk A per chunk (0: never)
};

// Kernel variants compiled into the library.
struct Variant {
  int wg_size, items, ring, algo, chunk, depth;
};
int num_variants();
Variant variant(int i);

// Arc positions per work piece of the oriented count kernel: single graphs
// and batches (their many short rows and small graphs run faster in longer
// pieces).  A graph whose pieces at that size would fit the grid's waves
// takes the smallest size (down to kOrientPieceMin) that still gives at
// most one piece per wave, and every wave then counts one piece; prow
// holds kOrientPieceSlack entries beyond nnz / kOrientPiece for that case.
// A piece's match counts must fit 16 bits: 2 * piece < 65536.
constexpr int64_t kOrientPiece = 256;
constexpr int64_t kOrientPieceBatch = 1024;
constexpr int64_t kOrientPieceMin = 16;
constexpr int64_t kOrientPieceSlack = 65536;
static_assert(2 * kOrientPieceBatch < 65536 && kOrientPieceMin <= kOrientPiece && kOrientPiece <= kOrientPieceBatch,
              "piece sizes");
// Ids of N+(r) per wave LDS window of the walked-array oriented kernel
// (k_count_orient): the default, compiled for every output, and the two
// others compiled for int32 counts.  A window holds 6 (window + 1) bytes of
// LDS per wave.
constexpr int kOrientWindow = 851;
constexpr int kOrientWindowSmall = 425;
constexpr int kOrientWindowLarge = 1701;
inline bool orient_window_ok(int w) { return w == kOrientWindow || w == kOrientWindowSmall || w == kOrientWindowLarge; }
// Piece size for nnz arc positions from base (kOrientPiece or
// kOrientPieceBatch) and the grid's waves; a result below base leaves at
// most min(waves, kOrientPieceSlack) pieces.
inline int64_t orient_piece(int64_t nnz, int64_t base, int64_t waves) {
  if (waves > kOrientPieceSlack) waves = kOrientPieceSlack;
  if (nnz > waves * base) return base;
  int64_t p = base;
  while (p > kOrientPieceMin && (nnz + p / 2 - 1) / (p / 2) <= waves) p /= 2;
  return p;
}

// Rows longer than this are listed and oriented first, one row per wave, so
// that a task of 64 vertices never holds several hub rows.
constexpr int64_t kHugeRow = 4096;
// The scalar block: dscal[0..7], then 16 ticket counters from dscal + kOtaskAt.
constexpr int kOtaskAt = 32, kOtaskStride = 32;
constexpr int kScalars = kOtaskAt + 16 * kOtaskStride;

// Scratch sizing and prep/count launchers (stream-ordered, no host sync).
struct Prep {
  int64_t* voff;    // [nv + 1]
  int32_t* vbase;   // [nv] graph v_base per vertex, or nullptr (single graph)
  int32_t* vmark;   // [nv] vertex-list mode: 1 for listed vertices, else nullptr
  int32_t* vcnt;    // [nv] vertex-list mode: per-vertex counts before the gather
  int32_t* ecnt;    // [nv]
  int64_t* eoff;    // [nv + 1]
  int32_t* eu;      // [ecap]
  int32_t* ew;      // [ecap]
  int32_t* ework;   // [ecap]
  int64_t* wpre;    // [ecap + 1]
  int64_t* dscal;   // [kScalars]: E, W, overflow flag, -, the edge build's vertex counter, -,
                    // huge rows listed, huge rows taken; then the ticket counters (otask)
  // 16 ticket counters kOtaskStride int64 apart (otask = dscal + kOtaskAt),
  // so that no two share a 256-byte block: 0-7 hand out the orientation's
  // tasks, 8-15 the count's pieces, ticket n of counter x being item 8 n + x
  unsigned long long* otask;
  unsigned int* qctr;  // [24]: three launches' per-XCD range queues
  int64_t* scan_tmp;
  int64_t ecap;
  // Degree-ordered orientation, in place of the input rows: row v keeps its
  // arc positions voff[v] .. voff[v+1] in oind, N+(v) first (ascending),
  // then the in-neighbours; orow[2v] = voff[v], orow[2v+1] = end of N+(v).
  int64_t* orow;    // [2 nv] or nullptr
  int32_t* oind;    // [nnz]
  // row of the first arc of every piece (orient_piece arc positions)
  int32_t* prow;    // [nnz / kOrientPiece + 2 + kOrientPieceSlack]
  // rows of more than kHugeRow arcs, oriented one per wave before the rest
  int32_t* huge;    // [nnz / kHugeRow + 1]
  int32_t* odeg;    // [nv + 1]: every vertex's degree, read by the orientation
  // rows whose N+ exceeds a wave's LDS window, counted by k_count_long
  int32_t* lrow;    // [nnz / 512 + 2]
};

size_t scan_tmp_elems(int64_t n_cap);

hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
hipError_t launch_fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t s);
hipError_t launch_voff_from_off32(const int32_t* off, int32_t nv, int64_t* voff, hipStream_t s);
hipError_t launch_voff_batch(const int64_t* tab, int32_t G, const int32_t* off_cat, int64_t nv_total,
                             int64_t* voff, int32_t* vbase, hipStream_t s);
// mode 0: undirected edges u < w with both degrees >= 2 (per-vertex counts)
// mode 1: all arcs of an oriented CSR whose endpoints both have out-arcs
// mode 2: every arc v -> w of a listed vertex (p.vmark), w != v, both
//         degrees >= 2 (the vertex-list count)
// mode 3: mode 2 with the arcs whose w > lim twice (shadow vertices)
hipError_t launch_build_edges(const Prep& p, int64_t nv, const int64_t* voff, const int32_t* ind, int mode,
                              int cus, hipStream_t s, int64_t lim = INT64_MAX);
// Orientation in place (k_orient_rows) into p.orow / p.oind of a CSR with
// nnz arcs (the size p was laid out for); ind holds graph-local ids when
// p.vbase is set.  prepped: launch_prep_single has run (degrees and the
// huge-row list are written, p.dscal[0..kScalars) were zeroed before it).
hipError_t launch_orient(const Prep& p, int64_t nv, int64_t nnz, const int32_t* ind, int cus, hipStream_t s,
                         bool prepped = false);
hipError_t launch_orient(const Prep& p, int64_t nv, int64_t nnz, const int64_t* ind, int cus, hipStream_t s,
                         bool prepped = false);
// A single graph's prep in one launch: widened offsets (int32 off; an int64
// CSR's offsets are p.voff itself), degrees, the huge-row list, the piece
// rows for pieces of `piece` arcs (orient_piece_of) and the output cleared
// (out nullable, total nullable).  p.dscal[0..kScalars) must be zero; then
// launch_orient and launch_count_orient run with prepped = true.
hipError_t launch_prep_single(const Prep& p, int64_t nv, int64_t nnz, const int32_t* off, int64_t piece,
                              int32_t* out, int64_t* total, hipStream_t s);
hipError_t launch_prep_single(const Prep& p, int64_t nv, int64_t nnz, const int64_t* off, int64_t piece,
                              int64_t* out, int64_t* total, hipStream_t s);
// Persistent grid: as many workgroups as fit on the device at once.
int resident_wgs(int variant_idx, int cus);
// which = 1: the second kernel of a two-launch variant
int resident_wgs2(int variant_idx, int cus, int which);
// Resident workgroups of the oriented count kernel; kind 0 = the total
// only, 1 = int32 counts, 2 = int64 counts; lookup 0 = the sorted-window
// merge (k_count_orient), else the table lookups (k_count_hash).
// nwg, nwg2, nwg3: grids of the streamed kernel's wave launch, its
// workgroup launch and the lookup kernel (persistent: resident workgroups).
hipError_t launch_count(const CountArgs& a, int algo, int variant_idx, int nwg, int nwg2, int nwg3, hipStream_t s);
int resident_search_wgs(int cus);
hipError_t launch_lcc(const int32_t* off, int32_t nv, const int32_t* out, double* lcc, hipStream_t s);
hipError_t launch_mark_list(const int32_t* list, int64_t nlist, int64_t nv, int32_t* vmark, hipStream_t s);
// Launch shape of the oriented count (launch_count_orient).
//   groups  arcs of a row streamed at a time per wave (1, 2, 4, 8, 16, 32)
//   side    1: rows are the higher vertex of each arc (partners are its
//           in-neighbours); 0: the lower one (partners are out-neighbours);
//           2: 1 unless the orientation of the same call listed rows of more
//           than kHugeRow arcs (p.dscal[6]), else 0
//   piece   base piece size in arc positions (orient_piece picks the one
//           used); at least kOrientPieceTune for layouts sized by it
//   lookup  0: the row window is a sorted LDS array that every lane walks
//           (k_count_orient); 1: an LDS cuckoo table probed twice per
//           element, the partner lists read in wave-coalesced 16-byte
//           vectors (k_count_hash); 2: the same kernel with every window a
//           sorted array searched per element
//   lng     rows whose N+ is longer than the per-wave kernel's LDS window
//           are listed (k_long_rows) and counted by workgroups sharing one
//           table of up to 2496 ids (k_count_long) instead of in windows
//   wg      threads per workgroup of k_count_orient (128 or 256)
//   window  ids of N+(r) per wave LDS window of k_count_orient
//           (kOrientWindow; kOrientWindowSmall or kOrientWindowLarge with
//           int32 counts)
struct OrientLaunch {
  int groups = 4;
  int side = 1;
  int64_t piece = kOrientPiece;
  int lookup = 0;
  bool lng = false;
  int wg = 256;
  int window = kOrientWindow;
};
int resident_orient_wgs(const OrientLaunch& o, int kind, int cus);
// Per-vertex counts (out; nullptr: the total only) over the orientation in
// p.orow / p.oind of nnz arcs.
// prepped: launch_prep_single wrote the piece rows and zeroed the counter.
hipError_t launch_count_orient(const Prep& p, int64_t nv, int64_t nnz, int32_t* out, int64_t* total,
                               const OrientLaunch& o, int cus, hipStream_t s, bool prepped = false);
hipError_t launch_count_orient(const Prep& p, int64_t nv, int64_t nnz, int64_t* out, int64_t* total,
                               const OrientLaunch& o, int cus, hipStream_t s, bool prepped = false);
// The piece size launch_count_orient takes for this launch shape (kind 0 =
// the total only, 1 = int32 counts, 2 = int64 counts).
int64_t orient_piece_of(const OrientLaunch& o, int kind, int cus, int64_t nnz);
hipError_t launch_gather_list(const int32_t* list, int64_t nlist, int64_t nv, const int32_t* vcnt, int32_t* out,
                              int64_t* total, hipStream_t s);
hipError_t launch_validate(const int32_t* off, const int32_t* ind, int32_t nv, int64_t nnz, int32_t* flag,
                           int cus, hipStream_t s);
hipError_t launch_mismatch(const int32_t* a, const int32_t* b, int64_t n, int32_t* cnt, hipStream_t s);
hipError_t launch_copy_i32(int32_t* dst, const int32_t* src, int64_t n, hipStream_t s);
// Per-graph totals of a batch's counts (nv_total vertices): waves of 1024
// vertices adding into totals[g] (zeroed first); keep_marked: totals[g] < 0
// (a graph set aside by launch_validate_batch) is left as it is and the
// others must already be 0.  nv_total < 0: one workgroup per graph.
hipError_t launch_graph_totals(const int64_t* tab, int32_t G, const int32_t* out, int64_t* totals,
                               hipStream_t s, bool keep_marked = false, int64_t nv_total = -1);
// The simple-graph contract of each graph of a packed batch (nnz_total arcs)
// whose offsets are known to run from 0 to nnz_g without decreasing:
// totals[g] = -1 for a graph that breaks it (0 for the others), whose ids
// are then rewritten in place to their row's own local id, so the count
// reads none of its rows.
hipError_t launch_validate_batch(const int64_t* tab, int32_t G, int32_t* off_cat, int32_t* ind_cat, int64_t nnz_total,
                                 int64_t* totals, int cus, hipStream_t s);

}  // namespace tcd
