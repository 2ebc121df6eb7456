// C-ABI of the gfx950 triangle counter (declared in include/tc.h).
//
// The context owns device scratch for the edge list and work prefix, sized
// from nv/nnz on the host so nothing is read back from the device: every
// entry point is stream-ordered and asynchronous except tc_validate_i32.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tc.h"
#include "tc_device.h"

// Scratch of tc_count_noalloc_i32: one copy of this array on every GPU the
// library runs on, held until the process ends.  In host code `scratch` is
// only the host-side handle of the symbol, never a device address: the
// device address is taken once per device (noalloc_setup, from
// tc_ctx_create with that device current) and only that address is passed
// to kernels.
static __device__ __align__(256) unsigned char scratch[TC_NOALLOC_SCRATCH_BYTES];

struct tc_ctx {
  int device = 0;
  int cus = 256;
  std::string err;
  void* pool = nullptr;  // device scratch owned by the context (grown by ensure)
  size_t cap = 0;
  tc_cfg def{};
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool ev_recorded = false;
  const int64_t* last_dscal = nullptr;  // E, W of the last count (device)
};

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

int set_err(tc_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(tc_ctx* c, hipError_t e, const char* where) {
  return set_err(c, TC_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}

tc_cfg builtin_default() {
  tc_cfg c{};
  c.wg_size = 256;
  c.items = 8;
  c.ring = 4096;
  c.small_max = 256;
  c.ranges_per_wg = 32;
  c.chunk = 16;
  c.depth = 2;
  c.algo = 0;
  c.search_ratio = 64;
  c.probe_ratio = 4;
  c.oriented = 3;     // automatic (orient_side)
  c.row_groups = -1;  // automatic (orient_groups)
  c.lookup = 0;       // automatic (orient_lookup)
  c.long_rows = 0;    // automatic (orient_long)
  c.window = tcd::kOrientWindow;
  c.orient_wg = 256;
  c.piece = -1;       // automatic (orient_piece_base)
  return c;
}

// Resolve a configuration (zero fields take the context default) to a
// compiled kernel variant.  Returns the variant index or -1.
int resolve(const tc_ctx* ctx, const tc_cfg* in, tc_cfg* out) {
  tc_cfg c = ctx->def;
  if (in) {
    if (in->wg_size) c.wg_size = in->wg_size;
    if (in->items) c.items = in->items;
    if (in->ring) c.ring = in->ring;
    if (in->small_max) c.small_max = in->small_max;
    if (in->ranges_per_wg) c.ranges_per_wg = in->ranges_per_wg;
    if (in->chunk) c.chunk = in->chunk;
    if (in->depth) c.depth = in->depth;
    if (in->search_ratio) c.search_ratio = in->search_ratio;
    if (in->probe_ratio) c.probe_ratio = in->probe_ratio;
    if (in->oriented) c.oriented = in->oriented > 0 ? (in->oriented > 3 ? 3 : in->oriented) : 0;
    if (in->row_groups) c.row_groups = in->row_groups;
    if (in->lookup) c.lookup = in->lookup;
    if (in->long_rows) c.long_rows = in->long_rows;
    if (in->window) c.window = in->window;
    if (in->orient_wg) c.orient_wg = in->orient_wg;
    if (in->piece) c.piece = in->piece;
    c.algo = in->algo;
  }
  int vi = -1;
  for (int i = 0; i < tcd::num_variants(); ++i) {
    tcd::Variant v = tcd::variant(i);
    const int want = c.algo == 1 ? 0 : c.algo;  // the simple kernel reuses the streamed sizes
    const bool streamed = want == 0 || want == 3;
    const bool stream_ok = !streamed || (v.chunk == c.chunk && v.depth == c.depth);
    if (v.wg_size == c.wg_size && v.items == c.items && v.ring == c.ring && v.algo == want && stream_ok) vi = i;
  }
  if (vi < 0 || c.ranges_per_wg < 1 || c.ranges_per_wg > 4096 || c.small_max < 1 || c.algo < 0 || c.algo > 3) return -1;
  if (c.row_groups > 0 && c.row_groups != 1 && c.row_groups != 2 && c.row_groups != 4 && c.row_groups != 8 &&
      c.row_groups != 16 && c.row_groups != 32)
    return -1;
  if (c.lookup < 0 || c.lookup > 3 || c.long_rows < 0 || c.long_rows > 2) return -1;
  if (!tcd::orient_window_ok(c.window)) return -1;
  if (c.orient_wg != 128 && c.orient_wg != 256) return -1;
  // layouts size the piece rows for pieces of kOrientPiece arcs or more
  if (c.piece != -1 && (c.piece < tcd::kOrientPiece || 2 * (int64_t)c.piece >= 65536)) return -1;
  const int lanecap = c.wg_size * c.items;
  if (c.small_max > lanecap) c.small_max = lanecap;
  if (c.algo != 0) c.oriented = 0;  // the oriented kernel is algo 0's
  *out = c;
  return vi;
}

// Partners a wave streams at once on the oriented scheme when row_groups
// asks for the automatic choice (<= 0).  Larger power-law graphs have
// longer N+ lists, which want fewer groups of more lanes; sparser graphs
// want more.  Measured best: R-MAT s26 1 (1335 ms; 2: 1562), s24 2 (209 ms;
// 4: 222), s20 8, the mixed 4096-graph batch 16, RGG n=4M 16-32.  A batch
// is sized by its mean degree only: its lists are those of its (small)
// graphs, whatever the batch's total.
int orient_lookup(const tc_cfg& cfg, int64_t nv, int64_t nnz, bool batch);
int orient_side(const tc_cfg& cfg, int64_t nv, int64_t nnz, bool batch);
constexpr int64_t kOutArcs = int64_t(1) << 30;
int orient_groups(const tc_cfg& cfg, int64_t nv, int64_t nnz, bool batch = false) {
  if (cfg.row_groups > 0) return cfg.row_groups;
  if (!batch && nnz >= kOutArcs) return 1;
  // the table kernel: with in-neighbour partners eight per wave, the
  // partners of a row ordered by list length (s24: 84.0 ms; 4: 87.1; s22
  // 21.4 / 21.6, s20 5.47 / 5.46; before the ordering 4 was best, s24 88.4
  // ms; 2: 90.3, 8: 90.6, 1: 98.4); with out-neighbour partners one from
  // 2^28 arcs (s24: 143.4 ms; 2: 144.8), two below (s22: 26.6 ms; 1: 27.7)
  if (!batch && orient_lookup(cfg, nv, nnz, batch) == 1)
    return orient_side(cfg, nv, nnz, batch) == 1 ? 8 : nnz >= (int64_t(1) << 28) ? 1 : 2;
  if (!batch && nnz >= (int64_t(1) << 28)) return 2;
  const double dbar = nv > 0 ? (double)nnz / (double)nv : 0.0;
  if (dbar < 10) return 32;
  if (dbar < 20) return 16;
  return 8;
}

// Scratch layout of one call.  The edge-list scheme keeps an undirected
// edge list (at most nnz/2 edges for a symmetric CSR; in the vertex-list
// mode every arc of a listed vertex is one, at most nnz, and with shadow
// vertices at most two); the oriented scheme keeps the orientation in place
// of the input rows.
enum class Scheme { kEdges, kList, kListShadow, kOrient, kOrientList };

Scheme scheme_of(const tc_cfg& cfg) { return cfg.oriented ? Scheme::kOrient : Scheme::kEdges; }

// Partner side of the oriented count (launch_count_orient's side): rows take
// their in-neighbours as partners (streaming the lower end's row of each
// arc) with oriented = 2 and out-neighbours with 1.  The automatic 3 takes
// in-neighbours for batches, for single graphs below 2^24 arcs and for the
// table kernel below 2^30 arcs; otherwise in-neighbours unless the
// prep lists hub rows (input rows of more than 4096 arcs, decided on the
// device).
// Measured with the walked array, in/out count time: the mixed batch
// 33.1/36.6 ms, R-MAT s14 0.09/0.18, s18 1.84/2.33, RGG n=4M 1.73/1.97, ER
// n=4M 1.86/1.98; but R-MAT s20 7.32/6.27 and s24 393/207 ms, where the
// hub-side lists streamed with out-neighbour partners stay in the caches.
// With the table kernel (the build's displacement fix made its many small
// builds cheap): s24 88.4/117.0 ms, s22 22.0/25.0, s21 12.9/13.9, s20
// 6.05/6.76; but s26 (2.1e9 arcs) 1543/1097.
constexpr int64_t kSideArcs = int64_t(1) << 24;
int orient_side(const tc_cfg& cfg, int64_t nv, int64_t nnz, bool batch) {
  if (cfg.oriented == 2) return 1;
  if (cfg.oriented != 3) return 0;
  if (batch || nnz < kSideArcs) return 1;
  if (orient_lookup(cfg, nv, nnz, batch) == 1 && nnz < kOutArcs) return 1;
  return 2;
}

// How the oriented count meets a partner's list with the row's (the
// launcher's lookup argument): 0 = the row window as a sorted LDS array
// walked by each lane (k_count_orient), 1 = an LDS cuckoo table probed
// twice per element of wave-coalesced reads (k_count_hash), 2 = the same
// kernel searching a sorted window per element.  cfg.lookup 1, 2 and 3 ask
// for them; 0 (automatic) takes the table for single graphs from 3 * 2^25
// arcs up, and from 2^24 arcs when the mean degree is at least 16, and the
// walked array otherwise.  Measured count times, walked / table (the
// table with in-neighbour partners, four per wave, below 2^30 arcs): R-MAT
// s24 183.1/88.4 ms, s22 (1.3e8 arcs) 34.5/22.0, s21 (6.4e7) 13.9/12.9, s20
// 6.13/6.05; but s18 1.83/2.24, RGG n=4M 1.64/1.92 (16 per wave) and ER
// n=4M (mean degree 8) 1.58/1.81: on short rows the table's builds do not
// pay for themselves.
constexpr int64_t kTableArcs = int64_t(3) << 25;
constexpr int64_t kTableDenseArcs = int64_t(1) << 24;
int orient_lookup(const tc_cfg& cfg, int64_t nv, int64_t nnz, bool batch) {
  if (cfg.lookup >= 1 && cfg.lookup <= 3) return cfg.lookup - 1;
  if (batch) return 0;
  return nnz >= kTableArcs || (nnz >= kTableDenseArcs && nnz >= 16 * nv) ? 1 : 0;
}
// Rows longer than a wave's LDS window counted by workgroups sharing one
// table (the table kernel's long-row phase, or k_count_long after the walked
// array) rather than in windows: cfg.long_rows 2 = always, 1 = never,
// 0 (automatic) = single graphs from 2^26 arcs up.  Measured at R-MAT s24:
// walked array 209.8 -> 188.7 ms, table 202.5 -> 175.0 ms; at s20 and RGG
// n=4M (few long rows) the two extra launches cost 1-3%, and on the mixed
// 4096-graph batch (small graphs, hardly any long row) 33.2 -> 33.7 ms.
constexpr int64_t kLongArcs = int64_t(1) << 26;
bool orient_long(const tc_cfg& cfg, int64_t nnz, bool batch) {
  return cfg.long_rows == 2 || (cfg.long_rows == 0 && !batch && nnz >= kLongArcs);
}

// The oriented count kernels of one call over the orientation in p, with
// cfg's automatic choices resolved for this size (partners per wave, partner
// side, lookup, long rows) and the piece size of single graphs or batches.
template <typename OT>
hipError_t count_oriented(const tc_ctx* c, const tcd::Prep& p, const tc_cfg& cfg, int64_t nv, int64_t nnz, OT* out,
                          int64_t* total, hipStream_t s, bool batch = false);

struct Layout {
  size_t voff, vbase, ecnt, eoff, eu, ew, ework, wpre, dscal, qctr, scan, orow, oind, prow, huge, odeg, lrow, vmark, vcnt,
      total;
  int64_t ecap;
  Scheme sc;
};

Layout layout(int64_t nv, int64_t nnz, bool batch, Scheme sc) {
  Layout L{};
  size_t p = 0;
  auto take = [&](size_t bytes) {
    size_t at = p;
    p += align_up(bytes);
    return at;
  };
  const bool orient = sc == Scheme::kOrient || sc == Scheme::kOrientList, edges = !orient;
  const bool list = sc == Scheme::kList || sc == Scheme::kListShadow, counts = list || sc == Scheme::kOrientList;
  L.sc = sc;
  L.ecap = edges ? (sc == Scheme::kListShadow ? 2 * nnz + 1 : list ? nnz + 1 : nnz / 2 + 1) : 0;
  L.voff = take(8 * (size_t)(nv + 1));
  L.vbase = take(batch ? 4 * (size_t)(nv + 1) : 0);
  L.dscal = take(8 * tcd::kScalars);
  L.qctr = take(4 * 24);
  L.ecnt = take(edges ? 4 * (size_t)(nv + 1) : 0);
  L.eoff = take(edges ? 8 * (size_t)(nv + 1) : 0);
  L.eu = take(4 * (size_t)L.ecap);
  L.ew = take(4 * (size_t)L.ecap);
  L.ework = take(4 * (size_t)L.ecap);
  L.wpre = take(edges ? 8 * (size_t)(L.ecap + 1) : 0);
  L.scan = take(edges ? 8 * tcd::scan_tmp_elems(std::max<int64_t>(nv + 1, L.ecap)) : 0);
  L.orow = take(orient ? 16 * (size_t)(nv + 1) : 0);
  L.oind = take(orient ? 4 * (size_t)(nnz + 1) : 0);
  L.prow = take(orient ? 4 * (size_t)(nnz / tcd::kOrientPiece + 2 + tcd::kOrientPieceSlack) : 0);
  L.huge = take(orient ? 4 * (size_t)(nnz / tcd::kHugeRow + 1) : 0);
  L.odeg = take(orient ? 4 * (size_t)(nv + 1) : 0);
  L.lrow = take(orient ? 4 * (size_t)(nnz / 512 + 2) : 0);
  L.vmark = take(list ? 4 * (size_t)(nv + 1) : 0);
  L.vcnt = take(counts ? 4 * (size_t)(nv + 1) : 0);
  L.total = p;
  return L;
}

// Scratch for a call whatever the scheme (for callers reserving up front).
size_t layout_bound(int64_t nv, int64_t nnz, bool batch) {
  return std::max(layout(nv, nnz, batch, Scheme::kEdges).total, layout(nv, nnz, batch, Scheme::kOrient).total);
}

int ensure(tc_ctx* c, size_t bytes) {
  if (bytes <= c->cap) return TC_OK;
  if (c->pool) {
    hipError_t e = hipFree(c->pool);
    c->pool = nullptr;
    c->cap = 0;
    c->last_dscal = nullptr;
    if (e != hipSuccess) return hip_fail(c, e, "hipFree");
  }
  size_t want = bytes + bytes / 8;
  hipError_t e = hipMalloc(&c->pool, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    c->pool = nullptr;
    return set_err(c, TC_ENOMEM, "device scratch of " + std::to_string(want) + " bytes: " + hipGetErrorString(e));
  }
  c->cap = want;
  return TC_OK;
}

tcd::Prep make_prep(void* base, const Layout& L, bool batch) {
  char* b = static_cast<char*>(base);
  tcd::Prep p{};
  p.voff = reinterpret_cast<int64_t*>(b + L.voff);
  p.vbase = batch ? reinterpret_cast<int32_t*>(b + L.vbase) : nullptr;
  p.ecnt = reinterpret_cast<int32_t*>(b + L.ecnt);
  p.eoff = reinterpret_cast<int64_t*>(b + L.eoff);
  p.eu = reinterpret_cast<int32_t*>(b + L.eu);
  p.ew = reinterpret_cast<int32_t*>(b + L.ew);
  p.ework = reinterpret_cast<int32_t*>(b + L.ework);
  p.wpre = reinterpret_cast<int64_t*>(b + L.wpre);
  p.dscal = reinterpret_cast<int64_t*>(b + L.dscal);
  p.otask = reinterpret_cast<unsigned long long*>(p.dscal + tcd::kOtaskAt);
  p.qctr = reinterpret_cast<unsigned int*>(b + L.qctr);
  p.scan_tmp = reinterpret_cast<int64_t*>(b + L.scan);
  p.ecap = L.ecap;
  const bool orient = L.sc == Scheme::kOrient || L.sc == Scheme::kOrientList;
  const bool list = L.sc == Scheme::kList || L.sc == Scheme::kListShadow;
  const bool counts = list || L.sc == Scheme::kOrientList;
  p.orow = orient ? reinterpret_cast<int64_t*>(b + L.orow) : nullptr;
  p.oind = orient ? reinterpret_cast<int32_t*>(b + L.oind) : nullptr;
  p.prow = orient ? reinterpret_cast<int32_t*>(b + L.prow) : nullptr;
  p.huge = orient ? reinterpret_cast<int32_t*>(b + L.huge) : nullptr;
  p.odeg = orient ? reinterpret_cast<int32_t*>(b + L.odeg) : nullptr;
  p.lrow = orient ? reinterpret_cast<int32_t*>(b + L.lrow) : nullptr;
  p.vmark = list ? reinterpret_cast<int32_t*>(b + L.vmark) : nullptr;
  p.vcnt = counts ? reinterpret_cast<int32_t*>(b + L.vcnt) : nullptr;
  return p;
}

// The oriented count's launch shape for cfg at this size: partners per
// wave, partner side, piece size (cfg.piece, or 256 arcs for single graphs
// and 1024 for batches), lookup, long rows, workgroup size and window.
tcd::OrientLaunch orient_launch(const tc_cfg& cfg, int64_t nv, int64_t nnz, bool batch) {
  tcd::OrientLaunch o;
  o.groups = orient_groups(cfg, nv, nnz, batch);
  o.side = orient_side(cfg, nv, nnz, batch);
  o.lookup = orient_lookup(cfg, nv, nnz, batch);
  // pieces of 512 arcs for the table kernel with in-neighbour partners from
  // 2^26 arcs (count ms, pieces of 256/512: R-MAT s22 20.9/18.5, s24
  // 82.1/82.1; s20 5.32/5.38 keeps 256)
  const bool wide = !batch && o.lookup >= 1 && o.side == 1 && nnz >= (int64_t(1) << 26);
  o.piece = cfg.piece > 0 ? cfg.piece : batch ? tcd::kOrientPieceBatch : wide ? 2 * tcd::kOrientPiece : tcd::kOrientPiece;
  o.lng = orient_long(cfg, nnz, batch);
  o.wg = cfg.orient_wg;
  o.window = cfg.window;
  return o;
}

template <typename OT>
hipError_t count_oriented(const tc_ctx* c, const tcd::Prep& p, const tc_cfg& cfg, int64_t nv, int64_t nnz, OT* out,
                          int64_t* total, hipStream_t s, bool batch) {
  tcd::OrientLaunch o = orient_launch(cfg, nv, nnz, batch);
  // the other windows are compiled for int32 counts only
  if (!out || sizeof(OT) != sizeof(int32_t)) o.window = tcd::kOrientWindow;
  return tcd::launch_count_orient(p, nv, nnz, out, total, o, c->cus, s);
}

void record(tc_ctx* c, int i, hipStream_t s) {
  if (c->timing && c->ev[i]) (void)hipEventRecord(c->ev[i], s);
}

// A single graph's oriented count in four launches: the counters cleared,
// one prep kernel (widened offsets, degrees, huge-row list, piece rows, the
// output cleared), the orientation, the count.  OffT/IT: the caller's offset
// and id types; out nullable (the total only), total nullable.
template <typename OffT, typename IT, typename OT>
hipError_t oriented_single(tc_ctx* c, const tcd::Prep& p, const tc_cfg& cfg, int64_t nv, int64_t nnz, const OffT* off,
                           const IT* ind, OT* out, int64_t* total, hipStream_t s) {
  hipError_t e = tcd::launch_fill_i64(p.dscal, tcd::kScalars, 0, s);
  if (e != hipSuccess) return e;
  tcd::OrientLaunch o = orient_launch(cfg, nv, nnz, false);
  if (!out || sizeof(OT) != sizeof(int32_t)) o.window = tcd::kOrientWindow;  // other windows: int32 counts only
  const int kind = !out ? 0 : sizeof(OT) == sizeof(int32_t) ? 1 : 2;
  if ((e = tcd::launch_prep_single(p, nv, nnz, off, tcd::orient_piece_of(o, kind, c->cus, nnz), out, total, s)) !=
      hipSuccess)
    return e;
  if ((e = tcd::launch_orient(p, nv, nnz, ind, c->cus, s, true)) != hipSuccess) return e;
  record(c, 1, s);
  return tcd::launch_count_orient(p, nv, nnz, out, total, o, c->cus, s, true);
}

// Persistent grid for the Merge Path kernel (every workgroup resident) and a
// fixed grid for the one-lane-per-edge kernel.
int count_wg(const tc_ctx* c, const tc_cfg& cfg, int vi, int which = 0) {
  if (cfg.algo == 1) return 16 * c->cus;
  int n = tcd::resident_wgs2(vi, c->cus, which);
  if (n >= 64) n &= ~7;
  return std::max(n, 1);
}

// Edge list (already built in p) -> counts.
int run_count(tc_ctx* c, const tcd::Prep& p, const tc_cfg& cfg, int vi, const int32_t* ind, int32_t* out,
              int64_t* total, int32_t mult, hipStream_t s, bool u_only = false) {
  tcd::CountArgs a{};
  a.wpre = p.wpre;
  a.eu = p.eu;
  a.ew = p.ew;
  a.voff = p.voff;
  a.ind = ind;
  a.dscal = p.dscal;
  a.out = out;
  a.total = reinterpret_cast<unsigned long long*>(total);
  a.total_mult = mult;
  a.small_max = cfg.small_max;
  a.min_units = 4 * cfg.wg_size * cfg.items;
  a.qctr = p.qctr;
  a.search_ratio = (cfg.algo == 0 || cfg.algo == 3) && cfg.search_ratio > 0 ? cfg.search_ratio : 0;
  a.u_only = u_only ? 1 : 0;
  a.probe_ratio = cfg.probe_ratio > 0 ? cfg.probe_ratio : 0;
  const int nwg = count_wg(c, cfg, vi);
  a.max_ranges = nwg * cfg.ranges_per_wg;
  int nwg3 = tcd::resident_search_wgs(c->cus);
  if (nwg3 >= 64) nwg3 &= ~7;
  hipError_t e = tcd::launch_count(a, cfg.algo, vi, nwg, count_wg(c, cfg, vi, 1), nwg3 > 0 ? nwg3 : 1, s);
  if (e != hipSuccess) return hip_fail(c, e, "count kernel launch");
  return TC_OK;
}

// Per-vertex count of an int32 CSR (tc_count_i32's launches) with its
// scratch at base, laid out as L for cfg's scheme.  Allocates nothing.
int count32(tc_ctx* c, void* base, const Layout& L, const tc_cfg& cfg, int vi, int32_t nv, int64_t nnz,
            const int32_t* d_off, const int32_t* d_ind, int32_t* d_out, int64_t* d_total, hipStream_t s) {
  hipError_t e;
  int rc = TC_OK;
  tcd::Prep p = make_prep(base, L, false);
  record(c, 0, s);
  if (cfg.oriented) {
    // Degree-ordered orientation: triangle {a, b, c} with (d, id) order
    // a < b < c is found once, at arc a -> b with match c; every one of the
    // three gets 2 (out[v] = 2 T(v)), the total 6 per triangle.
    // No edge list is built: tc_last_edge_work reports (0, 0) for this mode.
    if ((e = oriented_single(c, p, cfg, nv, nnz, d_off, d_ind, d_out, d_total, s)) != hipSuccess)
      return hip_fail(c, e, "oriented count launch");
  } else {
    if ((e = tcd::launch_fill_i32(d_out, nv, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
    if (d_total && (e = tcd::launch_fill_i64(d_total, 1, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
    if ((e = tcd::launch_voff_from_off32(d_off, nv, p.voff, s)) != hipSuccess) return hip_fail(c, e, "voff");
    if ((e = tcd::launch_build_edges(p, nv, p.voff, d_ind, 0, c->cus, s)) != hipSuccess)
      return hip_fail(c, e, "edge list");
    record(c, 1, s);
    rc = run_count(c, p, cfg, vi, d_ind, d_out, d_total, 2, s);
  }
  if (rc) return rc;
  c->last_dscal = p.dscal;
  record(c, 2, s);
  record(c, 3, s);
  c->ev_recorded = c->timing;
  return TC_OK;
}

// The batched count (tc_batch_count_i32's launches) with its scratch in
// the context's pool, or at base (base_bytes long, 256-byte aligned; the
// serving loop's arena blocks) when base is given.  Allocates nothing in
// the second case.
int batch_count(tc_ctx* c, void* base, size_t base_bytes, int32_t G, const int64_t* d_tab, int64_t nv_total,
                int64_t nnz_total, const int32_t* d_off_cat, const int32_t* d_ind_cat, int32_t* d_out_cat,
                int64_t* d_totals, const tc_cfg* cfg_in, void* stream) {
  if (!c) return TC_EINVAL;
  if (G <= 0 || !d_tab || !d_off_cat || nv_total < 0 || nnz_total < 0 || (nnz_total > 0 && !d_ind_cat) ||
      (nv_total > 0 && !d_out_cat))
    return set_err(c, TC_EINVAL, "invalid argument");
  if (nv_total >= INT32_MAX) return set_err(c, TC_ERANGE, "total vertex count beyond int32");
  tc_cfg cfg;
  const int vi = resolve(c, cfg_in, &cfg);
  if (vi < 0) return set_err(c, TC_EINVAL, "unsupported launch configuration");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  const Layout L = layout(nv_total, nnz_total, true, scheme_of(cfg));
  int rc = TC_OK;
  if (base) {
    if (L.total > base_bytes) return set_err(c, TC_ENOMEM, "batch scratch beyond the given region");
  } else if ((rc = ensure(c, L.total)) != TC_OK) {
    return rc;
  }
  tcd::Prep p = make_prep(base ? base : c->pool, L, true);
  record(c, 0, s);
  if ((e = tcd::launch_fill_i32(d_out_cat, nv_total, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
  if ((e = tcd::launch_voff_batch(d_tab, G, d_off_cat, nv_total, p.voff, p.vbase, s)) != hipSuccess)
    return hip_fail(c, e, "voff");
  if (L.sc == Scheme::kOrient) {
    // The batch is one disjoint graph in global ids: the orientation
    // (graph-local ids translated through p.vbase) never joins two graphs,
    // so every triangle stays inside its graph.
    if ((e = tcd::launch_fill_i64(p.dscal, 2, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
    if ((e = tcd::launch_orient(p, nv_total, nnz_total, d_ind_cat, c->cus, s)) != hipSuccess)
      return hip_fail(c, e, "orient");
    record(c, 1, s);
    if ((e = count_oriented(c, p, cfg, nv_total, nnz_total, d_out_cat, static_cast<int64_t*>(nullptr), s, true)) !=
        hipSuccess)
      return hip_fail(c, e, "count kernel launch");
  } else {
    if ((e = tcd::launch_build_edges(p, nv_total, p.voff, d_ind_cat, 0, c->cus, s)) != hipSuccess)
      return hip_fail(c, e, "edge list");
    record(c, 1, s);
    rc = run_count(c, p, cfg, vi, d_ind_cat, d_out_cat, nullptr, 2, s);
    if (rc) return rc;
  }
  c->last_dscal = p.dscal;
  record(c, 2, s);
  if (d_totals && (e = tcd::launch_graph_totals(d_tab, G, d_out_cat, d_totals, s, false, nv_total)) != hipSuccess)
    return hip_fail(c, e, "totals");
  record(c, 3, s);
  c->ev_recorded = c->timing;
  return TC_OK;
}

// Contexts of the kernelCall stand-ins: one per device and calling thread,
// so concurrent callers never share scratch; destroyed when the thread ends.
struct ShimContexts {
  tc_ctx* ctx[64] = {nullptr};
  ~ShimContexts() {
    for (tc_ctx*& c : ctx)
      if (c) tc_ctx_destroy(c);
  }
};
tc_ctx* shim_ctx(int dev) {
  static thread_local ShimContexts held;
  if (dev < 0 || dev >= 64) return nullptr;
  if (!held.ctx[dev] && tc_ctx_create(dev, &held.ctx[dev]) != TC_OK) held.ctx[dev] = nullptr;
  return held.ctx[dev];
}

// Device address of `scratch` per device (nullptr until resolved) and the
// error of a failed lookup.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {nullptr};
hipError_t g_noalloc_err[kMaxDevices] = {};

// Resolves `scratch` on device dev, which must be current; once per device.
void noalloc_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[dev]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    g_noalloc_err[dev] = e != hipSuccess ? e : hipErrorInvalidSymbol;
    return;
  }
  g_noalloc_ptr[dev] = p;
  g_noalloc_err[dev] = hipSuccess;
}

void* noalloc_ptr(int dev, hipError_t* err) {
  if (dev < 0 || dev >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  *err = g_noalloc_err[dev];
  return g_noalloc_ptr[dev];
}

}  // namespace

extern "C" {

int tc_ctx_create(int device, tc_ctx** out) {
  if (!out) return TC_EINVAL;
  *out = nullptr;
  int dev = device;
  hipError_t e;
  if (dev < 0) {
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return TC_EHIP;
  }
  e = hipSetDevice(dev);
  if (e != hipSuccess) return TC_EHIP;
  noalloc_setup(dev);  // tc_count_noalloc_i32's array on this device
  auto* c = new tc_ctx;
  c->device = dev;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0) c->cus = cus;
  c->def = builtin_default();
  for (auto& ev : c->ev) {
    if (hipEventCreate(&ev) != hipSuccess) ev = nullptr;
  }
  *out = c;
  return TC_OK;
}

int tc_ctx_destroy(tc_ctx* c) {
  if (!c) return TC_EINVAL;
  int dev0 = 0;
  (void)hipGetDevice(&dev0);
  (void)hipSetDevice(c->device);
  if (c->pool) (void)hipFree(c->pool);
  for (auto& ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  (void)hipSetDevice(dev0);
  delete c;
  return TC_OK;
}

const char* tc_last_error(const tc_ctx* c) { return c ? c->err.c_str() : "null context"; }

size_t tc_scratch_bytes(int32_t nv, int64_t nnz) { return layout_bound(nv, nnz, false); }
size_t tc_batch_scratch_bytes(int32_t G, int64_t nv_total, int64_t nnz_total) {
  return layout_bound(nv_total, nnz_total, true);
}

int tc_ctx_reserve(tc_ctx* c, size_t bytes) {
  if (!c) return TC_EINVAL;
  return ensure(c, bytes);
}

int tc_ctx_set_default_cfg(tc_ctx* c, const tc_cfg* cfg) {
  if (!c || !cfg) return TC_EINVAL;
  tc_cfg r;
  tc_cfg saved = c->def;
  c->def = builtin_default();
  if (resolve(c, cfg, &r) < 0) {
    c->def = saved;
    return set_err(c, TC_EINVAL, "unsupported configuration");
  }
  c->def = r;
  return TC_OK;
}

int tc_ctx_get_default_cfg(const tc_ctx* c, tc_cfg* cfg) {
  if (!c || !cfg) return TC_EINVAL;
  *cfg = c->def;
  return TC_OK;
}

int tc_ctx_set_timing(tc_ctx* c, int enable) {
  if (!c) return TC_EINVAL;
  c->timing = enable != 0;
  c->ev_recorded = false;
  return TC_OK;
}

int tc_last_timing(tc_ctx* c, float ms[4]) {
  if (!c || !ms) return TC_EINVAL;
  if (!c->timing || !c->ev_recorded) return set_err(c, TC_EINVAL, "no timed call recorded");
  hipError_t e = hipEventSynchronize(c->ev[3]);
  if (e != hipSuccess) return hip_fail(c, e, "hipEventSynchronize");
  if (hipEventElapsedTime(&ms[0], c->ev[0], c->ev[1]) != hipSuccess ||
      hipEventElapsedTime(&ms[1], c->ev[1], c->ev[2]) != hipSuccess ||
      hipEventElapsedTime(&ms[2], c->ev[2], c->ev[3]) != hipSuccess ||
      hipEventElapsedTime(&ms[3], c->ev[0], c->ev[3]) != hipSuccess)
    return set_err(c, TC_EHIP, "hipEventElapsedTime failed");
  return TC_OK;
}

int tc_device_cus(const tc_ctx* c) { return c ? c->cus : 0; }

int tc_copy_i32(tc_ctx* c, int32_t* d_dst, const int32_t* d_src, int64_t n, void* stream) {
  if (!c) return TC_EINVAL;
  if (n < 0 || (n > 0 && (!d_dst || !d_src))) return set_err(c, TC_EINVAL, "invalid argument");
  hipError_t e = hipSetDevice(c->device);
  if (e == hipSuccess) e = tcd::launch_copy_i32(d_dst, d_src, n, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? TC_OK : hip_fail(c, e, "copy");
}

int tc_oriented_plan(tc_ctx* c, int64_t nv, int64_t nnz, int32_t batch, const tc_cfg* cfg_in, int32_t plan[4]) {
  if (!c || !plan || nv < 0 || nnz < 0) return TC_EINVAL;
  tc_cfg cfg;
  if (resolve(c, cfg_in, &cfg) < 0) return set_err(c, TC_EINVAL, "unsupported launch configuration");
  if (!cfg.oriented) {
    plan[0] = plan[1] = -1;
    plan[2] = plan[3] = 0;
    return TC_OK;
  }
  const tcd::OrientLaunch o = orient_launch(cfg, nv, nnz, batch != 0);
  plan[0] = o.lookup;
  plan[1] = o.side;
  plan[2] = o.groups;
  plan[3] = o.lng ? 1 : 0;
  return TC_OK;
}

int tc_last_edge_work(tc_ctx* c, int64_t out[2]) {
  if (!c || !out) return TC_EINVAL;
  if (!c->last_dscal) return set_err(c, TC_EINVAL, "no count has run on this context");
  hipError_t e = hipSetDevice(c->device);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, c->last_dscal, 2 * sizeof(int64_t), hipMemcpyDeviceToHost);
  return e == hipSuccess ? TC_OK : hip_fail(c, e, "tc_last_edge_work");
}

int tc_count_i32(tc_ctx* c, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind, int32_t* d_out,
                 int64_t* d_total, const tc_cfg* cfg_in, void* stream) {
  if (!c) return TC_EINVAL;
  if (nv < 0 || nnz < 0 || !d_off || (nnz > 0 && !d_ind) || (nv > 0 && !d_out))
    return set_err(c, TC_EINVAL, "invalid argument");
  if (nnz > INT32_MAX) return set_err(c, TC_ERANGE, "nnz >= 2^31 on the int32 path");
  tc_cfg cfg;
  const int vi = resolve(c, cfg_in, &cfg);
  if (vi < 0) return set_err(c, TC_EINVAL, "unsupported launch configuration");
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  const Layout L = layout(nv, nnz, false, scheme_of(cfg));
  int rc = ensure(c, L.total);
  if (rc) return rc;
  return count32(c, c->pool, L, cfg, vi, nv, nnz, d_off, d_ind, d_out, d_total, static_cast<hipStream_t>(stream));
}

int tc_count_noalloc_i32(tc_ctx* c, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                         int32_t* d_out, int64_t* d_total, const tc_cfg* cfg_in, void* stream) {
  if (!c) return TC_EINVAL;
  if (nv < 0 || nnz < 0 || !d_off || (nnz > 0 && !d_ind) || (nv > 0 && !d_out))
    return set_err(c, TC_EINVAL, "invalid argument");
  if (nnz > INT32_MAX) return set_err(c, TC_ERANGE, "nnz >= 2^31 on the int32 path");
  tc_cfg cfg;
  const int vi = resolve(c, cfg_in, &cfg);
  if (vi < 0) return set_err(c, TC_EINVAL, "unsupported launch configuration");
  if (!cfg.oriented) return set_err(c, TC_EINVAL, "the no-allocation count runs the oriented scheme only");
  const Layout L = layout(nv, nnz, false, Scheme::kOrient);
  if (L.total > TC_NOALLOC_SCRATCH_BYTES)
    return set_err(c, TC_ERANGE, "graph needs " + std::to_string(L.total) + " bytes of scratch, more than the " +
                                     std::to_string(TC_NOALLOC_SCRATCH_BYTES) + "-byte no-allocation array");
  hipError_t e = hipSuccess;
  void* base = noalloc_ptr(c->device, &e);
  if (!base) return hip_fail(c, e, "no-allocation scratch address");
  if ((e = hipSetDevice(c->device)) != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  return count32(c, base, L, cfg, vi, nv, nnz, d_off, d_ind, d_out, d_total, static_cast<hipStream_t>(stream));
}

namespace {
// The int64 CSR on the oriented scheme: per-vertex int64 counts into d_out,
// or the total only when d_out is null.
int count_i64(tc_ctx* c, int64_t nv, int64_t nnz, const int64_t* d_off, const int64_t* d_ind, int64_t* d_out,
              int64_t* d_total, const tc_cfg* cfg_in, void* stream) {
  if (nv >= INT32_MAX) return set_err(c, TC_ERANGE, "more than 2^31 - 1 vertices");
  tc_cfg cfg;
  if (resolve(c, cfg_in, &cfg) < 0) return set_err(c, TC_EINVAL, "unsupported launch configuration");
  if (!cfg.oriented) return set_err(c, TC_EINVAL, "the int64 path runs the oriented scheme (algo 0, oriented >= 1)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  const Layout L = layout(nv, nnz, false, Scheme::kOrient);
  int rc = ensure(c, L.total);
  if (rc) return rc;
  tcd::Prep p = make_prep(c->pool, L, false);
  p.voff = const_cast<int64_t*>(d_off);  // the caller's offsets are already int64; the kernels only read them
  record(c, 0, s);
  if ((e = oriented_single(c, p, cfg, nv, nnz, d_off, d_ind, d_out, d_total, s)) != hipSuccess)
    return hip_fail(c, e, "oriented count launch");
  c->last_dscal = p.dscal;
  record(c, 2, s);
  record(c, 3, s);
  c->ev_recorded = c->timing;
  return TC_OK;
}

}  // namespace

int tc_count_i64(tc_ctx* c, int64_t nv, int64_t nnz, const int64_t* d_off, const int64_t* d_ind, int64_t* d_out,
                 int64_t* d_total, const tc_cfg* cfg_in, void* stream) {
  if (!c) return TC_EINVAL;
  if (nv < 0 || nnz < 0 || !d_off || (nnz > 0 && !d_ind) || (nv > 0 && !d_out))
    return set_err(c, TC_EINVAL, "invalid argument");
  return count_i64(c, nv, nnz, d_off, d_ind, d_out, d_total, cfg_in, stream);
}

int tc_count_total_i64(tc_ctx* c, int64_t nv, int64_t nnz, const int64_t* d_off, const int64_t* d_ind,
                       int64_t* d_total, void* stream) {
  if (!c) return TC_EINVAL;
  if (nv < 0 || nnz < 0 || !d_off || (nnz > 0 && !d_ind) || !d_total) return set_err(c, TC_EINVAL, "invalid argument");
  return count_i64(c, nv, nnz, d_off, d_ind, nullptr, d_total, nullptr, stream);
}

int tc_count_total_i32(tc_ctx* c, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                       int64_t* d_total, void* stream) {
  if (!c) return TC_EINVAL;
  if (nv < 0 || nnz < 0 || !d_off || (nnz > 0 && !d_ind) || !d_total) return set_err(c, TC_EINVAL, "invalid argument");
  if (nnz > INT32_MAX) return set_err(c, TC_ERANGE, "nnz >= 2^31 on the int32 path");
  tc_cfg cfg;
  if (resolve(c, nullptr, &cfg) < 0) return set_err(c, TC_EINVAL, "unsupported launch configuration");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  const Layout L = layout(nv, nnz, false, Scheme::kOrient);
  int rc = ensure(c, L.total);
  if (rc) return rc;
  tcd::Prep p = make_prep(c->pool, L, false);
  record(c, 0, s);
  // The degree-ordered orientation finds every triangle once; without
  // per-vertex output the count kernel keeps only the total (6 per
  // triangle).  No edge list: tc_last_edge_work reports (0, 0).
  if ((e = oriented_single(c, p, cfg, nv, nnz, d_off, d_ind, static_cast<int32_t*>(nullptr), d_total, s)) !=
      hipSuccess)
    return hip_fail(c, e, "oriented count launch");
  c->last_dscal = p.dscal;
  record(c, 2, s);
  record(c, 3, s);
  c->ev_recorded = c->timing;
  return TC_OK;
}

namespace {
// The vertex-list count (tc_count_vertices_i32) and its shadow-vertex form
// (tc_count_vertices_lim_i32, shadow = true): the listed vertices' arcs as
// an edge list credited to the listed end, or, for long lists without
// shadow vertices, the whole oriented count gathered per position.
int count_vertices(tc_ctx* c, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind, int32_t nlist,
                   const int32_t* d_list, bool shadow, int64_t lim, int32_t* d_out, int64_t* d_total,
                   const tc_cfg* cfg_in, void* stream) {
  if (!c) return TC_EINVAL;
  if (nv < 0 || nnz < 0 || nlist < 0 || !d_off || (nnz > 0 && !d_ind) || (nlist > 0 && (!d_list || !d_out)))
    return set_err(c, TC_EINVAL, "invalid argument");
  if (nnz > INT32_MAX) return set_err(c, TC_ERANGE, "nnz >= 2^31 on the int32 path");
  if (shadow && 2 * nnz + 1 > INT32_MAX)
    return set_err(c, TC_ERANGE, "nnz >= 2^30 with shadow vertices (every arc may be listed twice)");
  tc_cfg cfg;
  const int vi = resolve(c, cfg_in, &cfg);
  if (vi < 0) return set_err(c, TC_EINVAL, "unsupported launch configuration");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  // A list holding an eighth of the vertices or more (repeats counted) is
  // cheaper as the whole oriented count gathered per position: the value of
  // a listed vertex is the same 2 T(v).  Shorter lists intersect only the
  // listed vertices' arcs.
  if (!shadow && cfg.oriented && (int64_t)nlist * 8 >= nv) {
    const Layout L = layout(nv, nnz, false, Scheme::kOrientList);
    int rc = ensure(c, L.total);
    if (rc) return rc;
    tcd::Prep p = make_prep(c->pool, L, false);
    record(c, 0, s);
    if (d_total && (e = tcd::launch_fill_i64(d_total, 1, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
    if ((e = tcd::launch_fill_i32(p.vcnt, nv, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
    if ((e = tcd::launch_fill_i64(p.dscal, 2, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
    if ((e = tcd::launch_voff_from_off32(d_off, nv, p.voff, s)) != hipSuccess) return hip_fail(c, e, "voff");
    if ((e = tcd::launch_orient(p, nv, nnz, d_ind, c->cus, s)) != hipSuccess) return hip_fail(c, e, "orient");
    record(c, 1, s);
    if ((e = count_oriented(c, p, cfg, nv, nnz, p.vcnt, nullptr, s)) != hipSuccess)
      return hip_fail(c, e, "count kernel launch");
    c->last_dscal = p.dscal;
    record(c, 2, s);
    if ((e = tcd::launch_gather_list(d_list, nlist, nv, p.vcnt, d_out, d_total, s)) != hipSuccess)
      return hip_fail(c, e, "gather");
    record(c, 3, s);
    c->ev_recorded = c->timing;
    return TC_OK;
  }
  const Layout L = layout(nv, nnz, false, shadow ? Scheme::kListShadow : Scheme::kList);
  int rc = ensure(c, L.total);
  if (rc) return rc;
  tcd::Prep p = make_prep(c->pool, L, false);
  record(c, 0, s);
  if (d_total && (e = tcd::launch_fill_i64(d_total, 1, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
  if ((e = tcd::launch_fill_i32(p.vmark, nv, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
  if ((e = tcd::launch_fill_i32(p.vcnt, nv, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
  if ((e = tcd::launch_mark_list(d_list, nlist, nv, p.vmark, s)) != hipSuccess) return hip_fail(c, e, "mark");
  if ((e = tcd::launch_voff_from_off32(d_off, nv, p.voff, s)) != hipSuccess) return hip_fail(c, e, "voff");
  if ((e = tcd::launch_build_edges(p, nv, p.voff, d_ind, shadow ? 3 : 2, c->cus, s, shadow ? lim : INT64_MAX)) !=
      hipSuccess)
    return hip_fail(c, e, "edge list");
  record(c, 1, s);
  rc = run_count(c, p, cfg, vi, d_ind, p.vcnt, nullptr, 1, s, true);
  if (rc) return rc;
  c->last_dscal = p.dscal;
  record(c, 2, s);
  if ((e = tcd::launch_gather_list(d_list, nlist, nv, p.vcnt, d_out, d_total, s)) != hipSuccess)
    return hip_fail(c, e, "gather");
  record(c, 3, s);
  c->ev_recorded = c->timing;
  return TC_OK;
}
}  // namespace

int tc_count_vertices_i32(tc_ctx* c, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                          int32_t nlist, const int32_t* d_list, int32_t* d_out, int64_t* d_total, const tc_cfg* cfg_in,
                          void* stream) {
  return count_vertices(c, nv, nnz, d_off, d_ind, nlist, d_list, false, 0, d_out, d_total, cfg_in, stream);
}

int tc_count_vertices_lim_i32(tc_ctx* c, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                              int32_t nlist, const int32_t* d_list, int32_t lim, int32_t* d_out, int64_t* d_total,
                              const tc_cfg* cfg_in, void* stream) {
  return count_vertices(c, nv, nnz, d_off, d_ind, nlist, d_list, true, lim, d_out, d_total, cfg_in, stream);
}

int tc_batch_count_i32(tc_ctx* c, int32_t G, const int64_t* d_tab, int64_t nv_total, int64_t nnz_total,
                       const int32_t* d_off_cat, const int32_t* d_ind_cat, int32_t* d_out_cat, int64_t* d_totals,
                       const tc_cfg* cfg_in, void* stream) {
  return batch_count(c, nullptr, 0, G, d_tab, nv_total, nnz_total, d_off_cat, d_ind_cat, d_out_cat, d_totals, cfg_in,
                     stream);
}

int tc_lcc_f64(tc_ctx* c, int32_t nv, const int32_t* d_off, const int32_t* d_out, double* d_lcc, void* stream) {
  if (!c) return TC_EINVAL;
  if (nv < 0 || !d_off || (nv > 0 && (!d_out || !d_lcc))) return set_err(c, TC_EINVAL, "invalid argument");
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  e = tcd::launch_lcc(d_off, nv, d_out, d_lcc, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? TC_OK : hip_fail(c, e, "lcc");
}

int tc_validate_i32(tc_ctx* c, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind, int* h_status,
                    void* stream) {
  if (!c) return TC_EINVAL;
  if (nv < 0 || nnz < 0 || !d_off || (nnz > 0 && !d_ind) || !h_status) return set_err(c, TC_EINVAL, "invalid argument");
  if (nnz > INT32_MAX) return set_err(c, TC_ERANGE, "nnz >= 2^31 on the int32 path");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  int rc = ensure(c, std::max<size_t>(c->cap, 256));
  if (rc) return rc;
  int32_t* flag = static_cast<int32_t*>(c->pool);
  if ((e = tcd::launch_fill_i32(flag, 1, 0, s)) != hipSuccess) return hip_fail(c, e, "fill");
  if ((e = tcd::launch_validate(d_off, d_ind, nv, nnz, flag, c->cus, s)) != hipSuccess) return hip_fail(c, e, "validate");
  int32_t h = 0;
  if ((e = hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return hip_fail(c, e, "copy");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(c, e, "sync");
  *h_status = h ? TC_EGRAPH : TC_OK;
  if (h) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "not a simple undirected CSR:%s%s%s%s", (h & 1) ? " ids/offsets out of range;" : "",
                  (h & 2) ? " row not strictly increasing;" : "", (h & 4) ? " self-loop;" : "",
                  (h & 8) ? " arc without reverse;" : "");
    c->err = msg;
  }
  return TC_OK;
}

int tc_kernelCall_i32(unsigned numberBlocks, unsigned numberThreads, int32_t nv, const int32_t* d_off,
                      const int32_t* d_ind, int32_t* outPutTriangles, int32_t threads_per_block, int32_t number_blocks,
                      int32_t shifter) {
  static const unsigned kBlockSizes[] = {32, 64, 96, 128, 192, 256, 512, 1024};
  static const int32_t kTpi[] = {1, 2, 4, 8, 16, 32};
  bool bs_ok = std::find(std::begin(kBlockSizes), std::end(kBlockSizes), numberThreads) != std::end(kBlockSizes);
  bool tpi_ok = std::find(std::begin(kTpi), std::end(kTpi), threads_per_block) != std::end(kTpi);
  if (!bs_ok || !tpi_ok || numberBlocks == 0 || number_blocks != (int32_t)numberThreads / threads_per_block ||
      shifter < 0 || shifter > 5 || (1 << shifter) != threads_per_block || nv < 0 || !d_off)
    return TC_EINVAL;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return TC_EHIP;
  tc_ctx* c = shim_ctx(dev);
  if (!c) return TC_EHIP;
  // nnz from the device row offsets, as the reference kernel reads them.
  int32_t nnz = 0;
  hipError_t e = hipMemcpy(&nnz, d_off + nv, sizeof nnz, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return TC_EHIP;
  return tc_count_i32(c, nv, nnz, d_off, d_ind, outPutTriangles, nullptr, nullptr, nullptr);
}


int tc_kernelCall_i64(unsigned numberBlocks, unsigned numberThreads, int64_t nv, const int64_t* d_off,
                      const int64_t* d_ind, int64_t* outPutTriangles, int64_t threads_per_block, int64_t number_blocks,
                      int64_t shifter) {
  static const unsigned kBlockSizes[] = {32, 64, 96, 128, 192, 256, 512, 1024};
  static const int64_t kTpi[] = {1, 2, 4, 8, 16, 32};
  bool bs_ok = std::find(std::begin(kBlockSizes), std::end(kBlockSizes), numberThreads) != std::end(kBlockSizes);
  bool tpi_ok = std::find(std::begin(kTpi), std::end(kTpi), threads_per_block) != std::end(kTpi);
  if (!bs_ok || !tpi_ok || numberBlocks == 0 || number_blocks != (int64_t)numberThreads / threads_per_block ||
      shifter < 0 || shifter > 5 || (int64_t(1) << shifter) != threads_per_block || nv < 0 || !d_off)
    return TC_EINVAL;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return TC_EHIP;
  tc_ctx* c = shim_ctx(dev);
  if (!c) return TC_EHIP;
  int64_t nnz = 0;
  hipError_t e = hipMemcpy(&nnz, d_off + nv, sizeof nnz, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return TC_EHIP;
  return tc_count_i64(c, nv, nnz, d_off, d_ind, outPutTriangles, nullptr, nullptr, nullptr);
}

}  // extern "C"

// ------------------------------------------------------------------ per-graph
// baseline, HIP graphs and the autotuner
struct tc_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
  // The context the call was recorded on: its scratch is the graph's own
  // (allocated before recording, freed with the graph), so later counts on
  // the caller's context, or other graphs, never move or share it.
  tc_ctx* own = nullptr;
};

namespace {

hipStream_t capture_stream(tc_ctx* c) {
  static thread_local hipStream_t s = nullptr;
  static thread_local int dev = -1;
  if (!s || dev != c->device) {
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    dev = c->device;
  }
  return s;
}

// Runs fn(stream) under stream capture and instantiates the result.
template <class F>
int capture(tc_ctx* c, F&& fn, tc_graph** out) {
  hipStream_t s = capture_stream(c);
  if (!s) return set_err(c, TC_EHIP, "cannot create capture stream");
  const bool timing = c->timing;
  c->timing = false;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    c->timing = timing;
    return hip_fail(c, e, "hipStreamBeginCapture");
  }
  const int rc = fn(s);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(s, &g);
  c->timing = timing;
  if (rc != TC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(c, e, "hipGraphInstantiate");
  }
  auto* r = new tc_graph;
  r->graph = g;
  r->exec = x;
  r->device = c->device;
  r->own = c;
  *out = r;
  return TC_OK;
}

// A context for one recorded call: the caller's device, defaults and
// timing flag, and scratch of its own.
int private_ctx(const tc_ctx* c, tc_ctx** out) {
  int rc = tc_ctx_create(c->device, out);
  if (rc != TC_OK) return rc;
  (*out)->def = c->def;
  return TC_OK;
}

// Everything a capture must not do inside (allocation, occupancy queries)
// done up front for a call of this size.
int prepare_capture(tc_ctx* c, const tc_cfg* cfg_in, int64_t nv, int64_t nnz, bool batch) {
  tc_cfg cfg;
  const int vi = resolve(c, cfg_in, &cfg);
  if (vi < 0) return set_err(c, TC_EINVAL, "unsupported launch configuration");
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  (void)tcd::resident_wgs2(vi, c->cus, 1);
  const Scheme sc = scheme_of(cfg);
  if (sc == Scheme::kOrient)
    (void)tcd::resident_orient_wgs(orient_launch(cfg, nv, nnz, batch), 1, c->cus);
  return ensure(c, layout(nv, nnz, batch, sc).total);
}

struct Candidate {
  tc_cfg cfg;
};

// The oriented scheme at every partners-per-wave setting with out- and
// then in-neighbour partners, then the edge
// scheme's streamed kernels (algo 0 and 3) at each compiled size.
// The tuner's search, in stages, each around the fastest configuration so
// far (best): 0 = partner side x partners per wave on the oriented scheme;
// 1 = LDS window x workgroup size of its count kernel; 2 = piece size (the
// partition granularity); 3 = lookup (walked array or table) x long rows,
// and the table at 4, 8 and 16 partners per wave;
// 4 = the edge-list scheme's streamed kernel variants x small-edge
// threshold x level-1 ranges per workgroup.  The reference's counterpart
// is its sweep over 8 block sizes x 6 threads per intersection
// (src/gpurun.cu:151-172, src/param.cu:4-5).
constexpr int kTuneStages = 5;
std::vector<tc_cfg> tune_stage(int stage, const tc_cfg& best) {
  std::vector<tc_cfg> out;
  if (stage == 0) {
    for (int o : {1, 2})
      for (int g : {1, 2, 4, 8, 16, 32}) {
        tc_cfg c{};
        c.algo = 0;
        c.oriented = o;
        c.row_groups = g;
        c.lookup = 1;
        out.push_back(c);
      }
    return out;
  }
  if (stage == 1) {
    for (int w : {tcd::kOrientWindowSmall, tcd::kOrientWindow, tcd::kOrientWindowLarge})
      for (int wg : {128, 256}) {
        tc_cfg c = best;
        c.window = w;
        c.orient_wg = wg;
        out.push_back(c);
      }
    return out;
  }
  if (stage == 2) {
    for (int pc : {256, 512, 1024, 2048, 4096}) {
      tc_cfg c = best;
      c.piece = pc;
      out.push_back(c);
    }
    return out;
  }
  if (stage == 3) {
    for (int lk : {1, 2})
      for (int lr : {1, 2}) {
        tc_cfg c = best;
        c.lookup = lk;
        c.long_rows = lr;
        out.push_back(c);
      }
    // the table kernel at partners-per-wave settings of its own (stage 0
    // searched them with the walked array)
    for (int g : {4, 8, 16}) {
      tc_cfg c = best;
      c.lookup = 2;
      c.long_rows = 2;
      c.row_groups = g;
      out.push_back(c);
    }
    return out;
  }
  for (int i = 0; i < tcd::num_variants(); ++i) {
    const tcd::Variant v = tcd::variant(i);
    if (v.algo != 0 && v.algo != 3) continue;  // the tuner searches the streamed kernels
    const int lanecap = v.wg_size * v.items;
    const int sms[3] = {64, 256, 1024};
    for (int si = 0; si < 3; ++si) {
      if (sms[si] > lanecap) continue;
      for (int rp : {8, 32, 128}) {
        tc_cfg c{};
        c.wg_size = v.wg_size;
        c.items = v.items;
        c.ring = v.ring;
        c.small_max = sms[si];
        c.ranges_per_wg = rp;
        c.algo = v.algo;
        c.chunk = v.chunk;
        c.depth = v.depth;
        c.oriented = -1;
        out.push_back(c);
      }
    }
  }
  return out;
}

// run(cfg, stream) launches one count writing d_out; n = output length.
template <class Run>
int autotune(tc_ctx* c, Run&& run, int32_t* d_out, int64_t n, int reps, int set_default, tc_tune_result* results,
             int n_results, int* n_written, tc_cfg* best) {
  if (reps < 1) reps = 1;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  hipStream_t s = capture_stream(c);
  if (!s) return set_err(c, TC_EHIP, "cannot create stream");
  int32_t* ref = nullptr;
  int32_t* flag = nullptr;
  if ((e = hipMalloc(&ref, 4 * (size_t)std::max<int64_t>(n, 1))) != hipSuccess) return hip_fail(c, e, "hipMalloc");
  if ((e = hipMalloc(&flag, 4)) != hipSuccess) {
    (void)hipFree(ref);
    return hip_fail(c, e, "hipMalloc");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best_ms = 0;
  int written = 0, rc = TC_OK;
  bool any_diff = false, have_best = false, first = true;
  tc_cfg best_cfg{};
  const bool timing = c->timing;
  c->timing = false;
  // one candidate: its counts against the first candidate's, then timed
  auto trial = [&](const tc_cfg& cand) -> int {
    int r = run(cand, s);
    if (r) return r;
    int32_t h = 0;
    if (first) {
      first = false;
      e = hipMemcpyAsync(ref, d_out, 4 * (size_t)n, hipMemcpyDeviceToDevice, s);
    } else {
      e = tcd::launch_fill_i32(flag, 1, 0, s);
      if (e == hipSuccess) e = tcd::launch_mismatch(ref, d_out, n, flag, s);
      if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "autotune check");
    (void)hipEventRecord(e0, s);
    for (int k = 0; k < reps && r == TC_OK; ++k) r = run(cand, s);
    (void)hipEventRecord(e1, s);
    if (r) return r;
    if ((e = hipEventSynchronize(e1)) != hipSuccess) return hip_fail(c, e, "autotune timing");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= reps;
    const bool same = h == 0;
    any_diff |= !same;
    if (results && written < n_results) results[written++] = tc_tune_result{cand, ms, same ? 1 : 0};
    if (same && (!have_best || ms < best_ms)) {
      have_best = true;
      best_cfg = cand;
      best_ms = ms;
    }
    return TC_OK;
  };
  for (int stage = 0; stage < kTuneStages && rc == TC_OK; ++stage)
    for (const tc_cfg& cand : tune_stage(stage, best_cfg))
      if ((rc = trial(cand)) != TC_OK) break;
  c->timing = timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(ref);
  (void)hipFree(flag);
  if (n_written) *n_written = written;
  if (rc) return rc;
  if (have_best) {
    // the winner as it ran: zero fields resolved against the default in
    // force while tuning, so using it later changes nothing the timing saw
    tc_cfg r{};
    if (resolve(c, &best_cfg, &r) < 0) return set_err(c, TC_EINVAL, "tuned configuration does not resolve");
    if (best) *best = r;
    if (set_default) c->def = r;
  }
  if (any_diff) return set_err(c, TC_EGRAPH, "a tuned configuration produced different counts");
  return TC_OK;
}

}  // namespace

extern "C" {

int tc_batch_count_pergraph_i32(tc_ctx* c, int32_t G, const int64_t* h_tab, const int32_t* d_off_cat,
                                const int32_t* d_ind_cat, int32_t* d_out_cat, int64_t* d_totals, const tc_cfg* cfg,
                                void* stream) {
  if (!c) return TC_EINVAL;
  if (G <= 0 || !h_tab || !d_off_cat || !d_out_cat) return set_err(c, TC_EINVAL, "invalid argument");
  // reserve for the largest graph so the loop allocates nothing
  int64_t max_n = 0, max_nnz = 0;
  for (int32_t g = 0; g < G; ++g) {
    max_n = std::max(max_n, h_tab[2 * g + 2] - h_tab[2 * g]);
    max_nnz = std::max(max_nnz, h_tab[2 * g + 3] - h_tab[2 * g + 1]);
  }
  int rc = ensure(c, layout_bound(max_n, max_nnz, false));
  if (rc) return rc;
  for (int32_t g = 0; g < G; ++g) {
    const int64_t vb = h_tab[2 * g], ab = h_tab[2 * g + 1];
    const int64_t n = h_tab[2 * g + 2] - vb, nnz = h_tab[2 * g + 3] - ab;
    rc = tc_count_i32(c, (int32_t)n, nnz, d_off_cat + vb + g, d_ind_cat + ab, d_out_cat + vb,
                      d_totals ? d_totals + g : nullptr, cfg, stream);
    if (rc) return rc;
  }
  return TC_OK;
}

int tc_count_capture_i32(tc_ctx* c, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                         int32_t* d_out, int64_t* d_total, const tc_cfg* cfg, tc_graph** out) {
  if (!c || !out) return TC_EINVAL;
  *out = nullptr;
  tc_ctx* own = nullptr;
  int rc = private_ctx(c, &own);
  if (rc) return set_err(c, rc, "cannot create the graph's context");
  rc = prepare_capture(own, cfg, nv, nnz, false);
  if (!rc)
    rc = capture(
        own, [&](hipStream_t s) { return tc_count_i32(own, nv, nnz, d_off, d_ind, d_out, d_total, cfg, s); }, out);
  if (rc) {
    set_err(c, rc, own->err);
    tc_ctx_destroy(own);
  }
  return rc;
}

int tc_batch_capture_i32(tc_ctx* c, int32_t G, const int64_t* d_tab, int64_t nv_total, int64_t nnz_total,
                         const int32_t* d_off_cat, const int32_t* d_ind_cat, int32_t* d_out_cat, int64_t* d_totals,
                         const tc_cfg* cfg, tc_graph** out) {
  if (!c || !out) return TC_EINVAL;
  *out = nullptr;
  tc_ctx* own = nullptr;
  int rc = private_ctx(c, &own);
  if (rc) return set_err(c, rc, "cannot create the graph's context");
  rc = prepare_capture(own, cfg, nv_total, nnz_total, true);
  if (!rc)
    rc = capture(
        own,
        [&](hipStream_t s) {
          return tc_batch_count_i32(own, G, d_tab, nv_total, nnz_total, d_off_cat, d_ind_cat, d_out_cat, d_totals, cfg,
                                    s);
        },
        out);
  if (rc) {
    set_err(c, rc, own->err);
    tc_ctx_destroy(own);
  }
  return rc;
}

int tc_graph_launch(tc_graph* g, void* stream) {
  if (!g || !g->exec) return TC_EINVAL;
  hipError_t e = hipSetDevice(g->device);
  if (e == hipSuccess) e = hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? TC_OK : TC_EHIP;
}

int tc_graph_destroy(tc_graph* g) {
  if (!g) return TC_EINVAL;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->own) tc_ctx_destroy(g->own);
  delete g;
  return TC_OK;
}

int tc_autotune_i32(tc_ctx* c, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind, int32_t* d_out,
                    int reps, int set_default, tc_tune_result* results, int n_results, int* n_written, tc_cfg* best) {
  if (!c) return TC_EINVAL;
  return autotune(
      c, [&](const tc_cfg& cfg, hipStream_t s) { return tc_count_i32(c, nv, nnz, d_off, d_ind, d_out, nullptr, &cfg, s); },
      d_out, nv, reps, set_default, results, n_results, n_written, best);
}

int tc_autotune_batch_i32(tc_ctx* c, int32_t G, const int64_t* d_tab, int64_t nv_total, int64_t nnz_total,
                          const int32_t* d_off_cat, const int32_t* d_ind_cat, int32_t* d_out_cat, int reps,
                          int set_default, tc_tune_result* results, int n_results, int* n_written, tc_cfg* best) {
  if (!c) return TC_EINVAL;
  return autotune(
      c,
      [&](const tc_cfg& cfg, hipStream_t s) {
        return tc_batch_count_i32(c, G, d_tab, nv_total, nnz_total, d_off_cat, d_ind_cat, d_out_cat, nullptr, &cfg, s);
      },
      d_out_cat, nv_total, reps, set_default, results, n_results, n_written, best);
}

}  // extern "C"

// ------------------------------------------------------------ serving loop
// Internal interface of tc_server.hip (same library; not exported).
namespace tci {
// Scratch bytes of a batched count of nv vertices and nnz arcs on the
// oriented scheme (the serving loop's contexts keep the default scheme).
size_t batch_orient_scratch(int64_t nv, int64_t nnz) { return layout(nv, nnz, true, Scheme::kOrient).total; }
// tc_batch_count_i32 with the scratch at base (base_bytes long).
int batch_count_at(tc_ctx* c, void* base, size_t base_bytes, int32_t G, const int64_t* d_tab, int64_t nv_total,
                   int64_t nnz_total, const int32_t* d_off_cat, const int32_t* d_ind_cat, int32_t* d_out_cat,
                   int64_t* d_totals, void* stream) {
  if (!base) return TC_EINVAL;
  return batch_count(c, base, base_bytes, G, d_tab, nv_total, nnz_total, d_off_cat, d_ind_cat, d_out_cat, d_totals,
                     nullptr, stream);
}
}  // namespace tci
