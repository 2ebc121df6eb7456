// Host-side graph driver: generators, simple-graph CSR builder, reference
// text formats and the batch packer.  CPU only (g++ -fopenmp); the GPU
// counter consumes the CSR this file produces.
//
// Reference counterparts:
//   readGraph<T> (.info)      src/graphRead.cpp:2-26
//   readGraph (METIS)         src/csrdatacreate.cpp:3-49
//   writeGraph (.info)        src/csrdatacreate.cpp:86-114
// The reference builder sorts rows but keeps duplicates and self-loops
// (csrdatacreate.cpp:45); tch_from_edges/tch_normalize produce the simple
// symmetric CSR the per-vertex count is defined on.
#include "tc_host.h"
#include "csr_check.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

struct tch_graph {
  int32_t n = 0;
  std::vector<int32_t> off;
  std::vector<int32_t> ind;
};

struct tch_batch {
  int32_t G = 0;
  std::vector<int64_t> tab;
  std::vector<int32_t> off;
  std::vector<int32_t> ind;
  std::vector<int32_t> kind;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <class T>
T* fail_null(const std::string& msg) {
  g_err = msg;
  return nullptr;
}

// SplitMix64 finaliser; the stream value number i of seed s is
// mix(s + (i + 1) * gamma), so any draw can be computed independently.
inline uint64_t mix64(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
inline uint64_t draw(uint64_t seed, uint64_t i) { return mix64(seed + (i + 1) * kGamma); }
inline uint64_t splitmix64(uint64_t x) { return mix64(x + kGamma); }
// Multiply-shift bounded integer in [0, n): bias below n / 2^64.
inline uint32_t bounded(uint64_t r, uint32_t n) {
  return (uint32_t)(((unsigned __int128)r * n) >> 64);
}
inline double unit_double(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

// Build the simple symmetric CSR from arc endpoints: count, scatter, then
// sort + unique each row.  The result depends only on the arc multiset.
bool build_simple(int32_t n, int64_t m, const int32_t* src, const int32_t* dst, tch_graph* g) {
  std::vector<int64_t> cnt((size_t)n + 1, 0);
  {
    std::vector<std::atomic<int64_t>> acnt((size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; ++v) acnt[v].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; ++e) {
      int32_t u = src[e], v = dst[e];
      if (u == v) continue;
      acnt[u].fetch_add(1, std::memory_order_relaxed);
      acnt[v].fetch_add(1, std::memory_order_relaxed);
    }
    for (int32_t v = 0; v < n; ++v) cnt[v + 1] = cnt[v] + acnt[v].load(std::memory_order_relaxed);
  }
  const int64_t raw = cnt[n];
  std::vector<int32_t> tmp((size_t)raw);
  {
    std::vector<std::atomic<int64_t>> cur((size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; ++v) cur[v].store(cnt[v], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; ++e) {
      int32_t u = src[e], v = dst[e];
      if (u == v) continue;
      tmp[cur[u].fetch_add(1, std::memory_order_relaxed)] = v;
      tmp[cur[v].fetch_add(1, std::memory_order_relaxed)] = u;
    }
  }
  std::vector<int64_t> deg((size_t)n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < n; ++v) {
    int32_t* b = tmp.data() + cnt[v];
    int32_t* e = tmp.data() + cnt[v + 1];
    std::sort(b, e);
    deg[v] = std::unique(b, e) - b;
  }
  int64_t total = 0;
  g->off.assign((size_t)n + 1, 0);
  for (int32_t v = 0; v < n; ++v) {
    total += deg[v];
    if (total > INT32_MAX) return fail(TCH_ERANGE, "arc count exceeds int32") == 0;
    g->off[v + 1] = (int32_t)total;
  }
  g->ind.resize((size_t)total);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < n; ++v)
    std::copy(tmp.data() + cnt[v], tmp.data() + cnt[v] + deg[v], g->ind.data() + g->off[v]);
  g->n = n;
  return true;
}

tch_graph* graph_from_arcs(int32_t n, std::vector<int32_t>& s, std::vector<int32_t>& d) {
  auto* g = new (std::nothrow) tch_graph;
  if (!g) return fail_null<tch_graph>("out of memory");
  if (!build_simple(n, (int64_t)s.size(), s.data(), d.data(), g)) {
    delete g;
    return nullptr;
  }
  return g;
}

// --- generators ---------------------------------------------------------
void gen_er(int32_t n, int64_t m, uint64_t seed, std::vector<int32_t>& s, std::vector<int32_t>& d) {
  s.resize((size_t)m);
  d.resize((size_t)m);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; ++e) {
    s[e] = (int32_t)bounded(draw(seed, 2 * (uint64_t)e), (uint32_t)n);
    d[e] = (int32_t)bounded(draw(seed, 2 * (uint64_t)e + 1), (uint32_t)n);
  }
}

// R-MAT: per level one 32-bit uniform r; row bit = r >= a+b, column bit =
// (a <= r < a+b) or r >= a+b+c (quadrants a | b / c | d).  Levels are taken
// most significant bit first; two levels per 64-bit draw.
void gen_rmat(int32_t scale, int64_t m, double a, double b, double c, uint64_t seed, int32_t n_keep,
              std::vector<int32_t>& s, std::vector<int32_t>& d) {
  const uint64_t t1 = (uint64_t)std::llround(a * 4294967296.0);
  const uint64_t t2 = (uint64_t)std::llround((a + b) * 4294967296.0);
  const uint64_t t3 = (uint64_t)std::llround((a + b + c) * 4294967296.0);
  const int32_t pairs = (scale + 1) / 2;
  s.resize((size_t)m);
  d.resize((size_t)m);
  std::vector<uint8_t> keep((size_t)m, 1);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; ++e) {
    uint32_t u = 0, v = 0;
    for (int32_t p = 0; p < pairs; ++p) {
      uint64_t r64 = draw(seed, (uint64_t)e * pairs + p);
      for (int32_t h = 0; h < 2; ++h) {
        int32_t lvl = 2 * p + h;
        if (lvl >= scale) break;
        uint64_t r = h ? (r64 >> 32) : (r64 & 0xffffffffull);
        uint32_t bu = r >= t2;
        uint32_t bv = (r >= t1 && r < t2) || r >= t3;
        u |= bu << (scale - 1 - lvl);
        v |= bv << (scale - 1 - lvl);
      }
    }
    s[e] = (int32_t)u;
    d[e] = (int32_t)v;
    if (n_keep > 0 && ((int64_t)u >= n_keep || (int64_t)v >= n_keep)) keep[e] = 0;
  }
  if (n_keep > 0) {
    int64_t w = 0;
    for (int64_t e = 0; e < m; ++e)
      if (keep[e]) {
        s[w] = s[e];
        d[w] = d[e];
        ++w;
      }
    s.resize((size_t)w);
    d.resize((size_t)w);
  }
}

// 2D geometric graph on a uniform grid of cells no smaller than r.  The
// predicate dx*dx + dy*dy < r2 is evaluated in double with no contraction
// (built with -ffp-contract=off), so the edge set is machine independent.
void gen_rgg(int32_t n, double mean_deg, uint64_t seed, std::vector<int32_t>& s, std::vector<int32_t>& d) {
  std::vector<double> x((size_t)n), y((size_t)n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    x[i] = unit_double(draw(seed, 2 * (uint64_t)i));
    y[i] = unit_double(draw(seed, 2 * (uint64_t)i + 1));
  }
  const double r2 = mean_deg / (M_PI * (double)n);
  const double r = std::sqrt(r2);
  int64_t gdim = (int64_t)std::floor(1.0 / r) - 1;
  if (gdim < 1) gdim = 1;
  if (gdim > 65536) gdim = 65536;
  const int64_t cells = gdim * gdim;
  auto cell_of = [&](int64_t i) {
    int64_t cx = (int64_t)(x[i] * (double)gdim), cy = (int64_t)(y[i] * (double)gdim);
    cx = std::min<int64_t>(std::max<int64_t>(cx, 0), gdim - 1);
    cy = std::min<int64_t>(std::max<int64_t>(cy, 0), gdim - 1);
    return cy * gdim + cx;
  };
  std::vector<int64_t> cstart((size_t)cells + 1, 0);
  std::vector<int64_t> pc((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    pc[i] = cell_of(i);
    cstart[pc[i] + 1]++;
  }
  for (int64_t c = 0; c < cells; ++c) cstart[c + 1] += cstart[c];
  std::vector<int32_t> pts((size_t)n);
  {
    std::vector<int64_t> cur(cstart.begin(), cstart.end() - 1);
    for (int64_t i = 0; i < n; ++i) pts[cur[pc[i]]++] = (int32_t)i;
  }
  // Each thread collects pairs (i<j) for its rows of cells.
  int nt = omp_get_max_threads();
  std::vector<std::vector<int32_t>> ls(nt), ld(nt);
#pragma omp parallel
  {
    int t = omp_get_thread_num();
    auto& S = ls[t];
    auto& D = ld[t];
#pragma omp for schedule(dynamic, 16)
    for (int64_t cy = 0; cy < gdim; ++cy)
      for (int64_t cx = 0; cx < gdim; ++cx) {
        int64_t c = cy * gdim + cx;
        for (int64_t a = cstart[c]; a < cstart[c + 1]; ++a) {
          int32_t i = pts[a];
          for (int64_t dy = -1; dy <= 1; ++dy)
            for (int64_t dx = -1; dx <= 1; ++dx) {
              int64_t ny = cy + dy, nx = cx + dx;
              if (ny < 0 || nx < 0 || ny >= gdim || nx >= gdim) continue;
              int64_t nc = ny * gdim + nx;
              for (int64_t q = cstart[nc]; q < cstart[nc + 1]; ++q) {
                int32_t j = pts[q];
                if (j <= i) continue;
                double ddx = x[i] - x[j];
                double ddy = y[i] - y[j];
                double dd = ddx * ddx;
                double ee = ddy * ddy;
                if (dd + ee < r2) {
                  S.push_back(i);
                  D.push_back(j);
                }
              }
            }
        }
      }
  }
  size_t tot = 0;
  for (auto& v : ls) tot += v.size();
  s.clear();
  d.clear();
  s.reserve(tot);
  d.reserve(tot);
  for (int t = 0; t < nt; ++t) {
    s.insert(s.end(), ls[t].begin(), ls[t].end());
    d.insert(d.end(), ld[t].begin(), ld[t].end());
  }
}

// --- text parsing ---------------------------------------------------------
bool slurp(const char* path, std::string& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize(sz > 0 ? (size_t)sz : 0);
  size_t got = sz > 0 ? std::fread(&buf[0], 1, (size_t)sz, f) : 0;
  std::fclose(f);
  return got == buf.size();
}

struct Scanner {
  const char* p;
  const char* e;
  // Next integer token; false at end of input or on a non-numeric token.
  bool next(int64_t& v) {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    if (p >= e) return false;
    bool neg = false;
    if (*p == '-' || *p == '+') {
      neg = *p == '-';
      ++p;
    }
    if (p >= e || *p < '0' || *p > '9') return false;
    int64_t x = 0;
    while (p < e && *p >= '0' && *p <= '9') {
      x = x * 10 + (*p - '0');
      if (x > ((int64_t)1 << 40)) return false;
      ++p;
    }
    v = neg ? -x : x;
    return true;
  }
};

}  // namespace

// ------------------------------------------------------------------ C-ABI
extern "C" {

const char* tch_last_error(void) { return g_err.c_str(); }
int32_t tch_graph_n(const tch_graph* g) { return g ? g->n : 0; }
int64_t tch_graph_nnz(const tch_graph* g) { return g ? (int64_t)g->ind.size() : 0; }
const int32_t* tch_graph_off(const tch_graph* g) { return g ? g->off.data() : nullptr; }
const int32_t* tch_graph_ind(const tch_graph* g) { return g ? g->ind.data() : nullptr; }
void tch_graph_free(tch_graph* g) { delete g; }

tch_graph* tch_from_edges(int32_t n, int64_t m, const int32_t* src, const int32_t* dst) {
  if (n < 0 || m < 0 || (m > 0 && (!src || !dst))) return fail_null<tch_graph>("invalid argument");
  for (int64_t e = 0; e < m; ++e)
    if (src[e] < 0 || src[e] >= n || dst[e] < 0 || dst[e] >= n)
      return fail_null<tch_graph>("edge " + std::to_string(e) + " has an endpoint outside [0,n)");
  auto* g = new (std::nothrow) tch_graph;
  if (!g) return fail_null<tch_graph>("out of memory");
  if (!build_simple(n, m, src, dst, g)) {
    delete g;
    return nullptr;
  }
  return g;
}

tch_graph* tch_from_csr(int32_t n, int64_t nnz, const int32_t* off, const int32_t* ind) {
  if (n < 0 || nnz < 0 || nnz > INT32_MAX || !off || (nnz > 0 && !ind))
    return fail_null<tch_graph>("invalid argument");
  auto* g = new (std::nothrow) tch_graph;
  if (!g) return fail_null<tch_graph>("out of memory");
  g->n = n;
  g->off.assign(off, off + (size_t)n + 1);
  g->ind.assign(ind, ind + nnz);
  return g;
}

tch_graph* tch_normalize(const tch_graph* g) {
  if (!g) return fail_null<tch_graph>("invalid argument");
  const int64_t nnz = (int64_t)g->ind.size();
  std::vector<int32_t> s((size_t)nnz), d((size_t)nnz);
  for (int32_t v = 0; v < g->n; ++v)
    for (int32_t k = g->off[v]; k < g->off[v + 1]; ++k) {
      s[k] = v;
      d[k] = g->ind[k];
    }
  return tch_from_edges(g->n, nnz, s.data(), d.data());
}

tch_graph* tch_gen_er(int32_t n, int64_t m, uint64_t seed) {
  if (n <= 0 || m < 0) return fail_null<tch_graph>("invalid argument");
  std::vector<int32_t> s, d;
  gen_er(n, m, seed, s, d);
  return graph_from_arcs(n, s, d);
}

tch_graph* tch_gen_rmat(int32_t scale, int64_t m, double a, double b, double c, uint64_t seed,
                        int32_t n_keep) {
  if (scale < 1 || scale > 30 || m < 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return fail_null<tch_graph>("invalid argument");
  int64_t full = (int64_t)1 << scale;
  if (n_keep > full) return fail_null<tch_graph>("n_keep exceeds 2^scale");
  std::vector<int32_t> s, d;
  gen_rmat(scale, m, a, b, c, seed, n_keep, s, d);
  return graph_from_arcs(n_keep > 0 ? n_keep : (int32_t)full, s, d);
}

tch_graph* tch_gen_rgg(int32_t n, double mean_deg, uint64_t seed) {
  if (n <= 0 || !(mean_deg > 0)) return fail_null<tch_graph>("invalid argument");
  std::vector<int32_t> s, d;
  gen_rgg(n, mean_deg, seed, s, d);
  return graph_from_arcs(n, s, d);
}

namespace {
// TCH_E* code of a row fault (tcs::row_fault).
int row_code(int f) {
  return f == tcs::kRowRange ? TCH_ERANGE : f == tcs::kRowUnsorted ? TCH_EUNSORTED
       : f == tcs::kRowSelfLoop ? TCH_ESELFLOOP : f == tcs::kRowAsym ? TCH_EASYM : 0;
}
}  // namespace

extern "C++" {
namespace {
// tch_check_simple for either offset/id type.
template <class OT, class IT>
int check_simple(int64_t n, int64_t nnz, const OT* off, const IT* ind) {
  if (n < 0 || nnz < 0 || !off || (nnz > 0 && !ind)) return fail(TCH_EINVAL, "invalid argument");
  if ((int64_t)off[0] != 0 || (int64_t)off[n] != nnz) return fail(TCH_ERANGE, "off[0] != 0 or off[n] != nnz");
  for (int64_t v = 0; v < n; ++v)
    if (off[v + 1] < off[v]) return fail(TCH_ERANGE, "offsets decrease at vertex " + std::to_string(v));
  std::atomic<int> code{0};
  std::atomic<int64_t> where{INT64_MAX};
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < n; ++v) {
    const int c = row_code(tcs::row_fault(n, off, ind, v));
    if (c) {
      int64_t prev = where.load();
      while (v < prev && !where.compare_exchange_weak(prev, v)) {
      }
      int expected = 0;
      code.compare_exchange_strong(expected, c);
    }
  }
  if (code.load()) {
    // Re-derive the code of the lowest offending vertex for a stable message.
    int64_t v = where.load();
    const int c = row_code(tcs::row_fault(n, off, ind, v));
    const char* what = c == TCH_ERANGE ? "id out of range" : c == TCH_EUNSORTED ? "row not strictly increasing"
                       : c == TCH_ESELFLOOP ? "self-loop" : "arc without reverse";
    return fail(c, std::string(what) + " at vertex " + std::to_string(v));
  }
  return TCH_OK;
}
}  // namespace
}  // extern "C++"

int tch_check_simple(int32_t n, int64_t nnz, const int32_t* off, const int32_t* ind) {
  if (nnz > INT32_MAX) return fail(TCH_ERANGE, "nnz exceeds int32");
  return check_simple(n, nnz, off, ind);
}

tch_graph* tch_read_info(const char* path_no_ext) {
  if (!path_no_ext) return fail_null<tch_graph>("invalid argument");
  std::string path = std::string(path_no_ext) + ".info";
  std::string buf;
  if (!slurp(path.c_str(), buf)) return (g_err = "cannot read " + path, nullptr);
  Scanner sc{buf.data(), buf.data() + buf.size()};
  int64_t nv1, nnz;
  if (!sc.next(nv1) || !sc.next(nnz) || nv1 < 1 || nnz < 0)
    return fail_null<tch_graph>(path + ": bad header");
  if (nv1 - 1 > INT32_MAX || nnz > INT32_MAX) return fail_null<tch_graph>(path + ": sizes exceed int32");
  if ((uint64_t)nv1 + (uint64_t)nnz > buf.size() / 2 + 1) return fail_null<tch_graph>(path + ": truncated indices");
  auto* g = new (std::nothrow) tch_graph;
  if (!g) return fail_null<tch_graph>("out of memory");
  g->n = (int32_t)(nv1 - 1);
  g->off.resize((size_t)nv1);
  g->ind.resize((size_t)nnz);
  int64_t v;
  for (int64_t i = 0; i < nv1; ++i) {
    if (!sc.next(v)) {
      delete g;
      return fail_null<tch_graph>(path + ": truncated offsets");
    }
    g->off[i] = (int32_t)v;
  }
  for (int64_t i = 0; i < nnz; ++i) {
    if (!sc.next(v)) {
      delete g;
      return fail_null<tch_graph>(path + ": truncated indices");
    }
    g->ind[i] = (int32_t)v;
  }
  return g;
}

int tch_read_partition(const char* path_no_ext, int32_t K, int32_t* n, int32_t** vals) {
  if (!path_no_ext || K < 1 || !n || !vals) return fail(TCH_EINVAL, "invalid argument");
  *n = 0;
  *vals = nullptr;
  const std::string path = std::string(path_no_ext) + ".part." + std::to_string(K);
  std::string buf;
  if (!slurp(path.c_str(), buf)) return fail(TCH_EIO, "cannot read " + path);
  Scanner sc{buf.data(), buf.data() + buf.size()};
  int64_t cnt, v;
  if (!sc.next(cnt) || cnt < 0 || cnt > INT32_MAX) return fail(TCH_EIO, path + ": bad header");
  int32_t* a = static_cast<int32_t*>(std::malloc(4 * (size_t)(cnt > 0 ? cnt : 1)));
  if (!a) return fail(TCH_ENOMEM, "out of memory");
  for (int64_t i = 0; i < cnt; ++i) {
    if (!sc.next(v) || v < INT32_MIN || v > INT32_MAX) {
      std::free(a);
      return fail(TCH_EIO, path + ": truncated or out of range");
    }
    a[i] = (int32_t)v;
  }
  *n = (int32_t)cnt;
  *vals = a;
  return TCH_OK;
}

int tch_partition_members(int32_t n, const int32_t* part, int32_t p, int32_t* count, int32_t** ids) {
  if (n < 0 || (n > 0 && !part) || !count || !ids) return fail(TCH_EINVAL, "invalid argument");
  int32_t c = 0;
  for (int32_t v = 0; v < n; ++v) c += part[v] == p;
  int32_t* a = static_cast<int32_t*>(std::malloc(4 * (size_t)(c > 0 ? c : 1)));
  if (!a) return fail(TCH_ENOMEM, "out of memory");
  int32_t k = 0;
  for (int32_t v = 0; v < n; ++v)
    if (part[v] == p) a[k++] = v;
  *count = c;
  *ids = a;
  return TCH_OK;
}

void tch_free(void* p) { std::free(p); }

int tch_write_info(const tch_graph* g, const char* path) {
  if (!g || !path) return fail(TCH_EINVAL, "invalid argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(TCH_EIO, std::string("cannot open ") + path);
  std::fprintf(f, "%zu\t%zu\n", g->off.size(), g->ind.size());
  for (size_t i = 0; i < g->off.size(); ++i) std::fprintf(f, i ? "\t%d" : "%d", g->off[i]);
  std::fputc('\n', f);
  for (size_t i = 0; i < g->ind.size(); ++i) std::fprintf(f, i ? "\t%d" : "%d", g->ind[i]);
  bool ok = std::fclose(f) == 0;
  return ok ? TCH_OK : fail(TCH_EIO, std::string("write failed: ") + path);
}

tch_graph* tch_read_metis(const char* path) {
  if (!path) return fail_null<tch_graph>("invalid argument");
  std::string buf;
  if (!slurp(path, buf)) return fail_null<tch_graph>(std::string("cannot read ") + path);
  size_t pos = 0;
  auto getline = [&](std::string& line) {
    if (pos >= buf.size()) return false;
    size_t nl = buf.find('\n', pos);
    if (nl == std::string::npos) nl = buf.size();
    line.assign(buf, pos, nl - pos);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    pos = nl + 1;
    return true;
  };
  std::string line;
  do {
    if (!getline(line)) return fail_null<tch_graph>(std::string(path) + ": no header");
  } while (!line.empty() && line[0] == '%');
  Scanner hs{line.data(), line.data() + line.size()};
  int64_t n, m;
  if (!hs.next(n) || !hs.next(m) || n < 0 || m < 0 || n > INT32_MAX)
    return fail_null<tch_graph>(std::string(path) + ": bad header");
  auto* g = new (std::nothrow) tch_graph;
  if (!g) return fail_null<tch_graph>("out of memory");
  g->n = (int32_t)n;
  g->off.reserve((size_t)n + 1);
  g->off.push_back(0);
  g->ind.reserve((size_t)std::min<int64_t>(2 * m, INT32_MAX));
  int64_t rows = 0;
  while (rows < n && getline(line)) {
    ++rows;
    Scanner rs{line.data(), line.data() + line.size()};
    size_t start = g->ind.size();
    int64_t v;
    while (rs.next(v)) {
      if (v < 1 || v > n) {
        delete g;
        return fail_null<tch_graph>(std::string(path) + ": neighbour id out of range in row " +
                                    std::to_string(rows));
      }
      g->ind.push_back((int32_t)(v - 1));
    }
    std::sort(g->ind.begin() + start, g->ind.end());
    if (g->ind.size() > (size_t)INT32_MAX) {
      delete g;
      return fail_null<tch_graph>(std::string(path) + ": arc count exceeds int32");
    }
    g->off.push_back((int32_t)g->ind.size());
  }
  if (rows < n) {
    delete g;
    return fail_null<tch_graph>(std::string(path) + ": truncated (fewer rows than the header says)");
  }
  return g;
}

int tch_write_bin(const tch_graph* g, const char* path) {
  if (!g || !path) return fail(TCH_EINVAL, "invalid argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(TCH_EIO, std::string("cannot open ") + path);
  char magic[8] = {'T', 'C', 'C', 'S', 'R', '0', '0', '1'};
  int64_t hdr[3] = {g->n, (int64_t)g->ind.size(), 0};
  bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hdr, 8, 3, f) == 3 &&
            std::fwrite(g->off.data(), 4, g->off.size(), f) == g->off.size() &&
            std::fwrite(g->ind.data(), 4, g->ind.size(), f) == g->ind.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? TCH_OK : fail(TCH_EIO, std::string("write failed: ") + path);
}

tch_graph* tch_read_bin(const char* path) {
  if (!path) return fail_null<tch_graph>("invalid argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail_null<tch_graph>(std::string("cannot read ") + path);
  char magic[8];
  int64_t hdr[3];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "TCCSR001", 8) != 0 || std::fread(hdr, 8, 3, f) != 3 ||
      hdr[0] < 0 || hdr[0] > INT32_MAX || hdr[1] < 0 || hdr[1] > INT32_MAX) {
    std::fclose(f);
    return fail_null<tch_graph>(std::string(path) + ": bad header");
  }
  auto* g = new (std::nothrow) tch_graph;
  if (!g) {
    std::fclose(f);
    return fail_null<tch_graph>("out of memory");
  }
  g->n = (int32_t)hdr[0];
  g->off.resize((size_t)hdr[0] + 1);
  g->ind.resize((size_t)hdr[1]);
  bool ok = std::fread(g->off.data(), 4, g->off.size(), f) == g->off.size() &&
            std::fread(g->ind.data(), 4, g->ind.size(), f) == g->ind.size();
  std::fclose(f);
  if (!ok) {
    delete g;
    return fail_null<tch_graph>(std::string(path) + ": truncated");
  }
  return g;
}

// ---------------------------------------------------------------- int64 CSR
// The reference binary counts with T = int64 (src/main.cu:36 runs
// run<int64_t>; readGraph<int64_t>, src/graphRead.cpp:2-26, reads the
// .info text into int64 arrays).  tch_graph64 holds that layout with no
// int32 cap on the arc count; ids stay below 2^31, the device counter's
// limit for them.
struct tch_graph64 {
  int64_t n = 0;
  std::vector<int64_t> off;
  std::vector<int64_t> ind;
};

int64_t tch_graph64_n(const tch_graph64* g) { return g ? g->n : 0; }
int64_t tch_graph64_nnz(const tch_graph64* g) { return g ? (int64_t)g->ind.size() : 0; }
const int64_t* tch_graph64_off(const tch_graph64* g) { return g ? g->off.data() : nullptr; }
const int64_t* tch_graph64_ind(const tch_graph64* g) { return g ? g->ind.data() : nullptr; }
void tch_graph64_free(tch_graph64* g) { delete g; }

tch_graph64* tch_widen(const tch_graph* g) {
  if (!g) return fail_null<tch_graph64>("invalid argument");
  auto* h = new (std::nothrow) tch_graph64;
  if (!h) return fail_null<tch_graph64>("out of memory");
  h->n = g->n;
  h->off.assign(g->off.begin(), g->off.end());
  h->ind.assign(g->ind.begin(), g->ind.end());
  return h;
}

int tch_check_simple64(int64_t n, int64_t nnz, const int64_t* off, const int64_t* ind) {
  if (n > INT32_MAX) return fail(TCH_ERANGE, "vertex count exceeds int32");
  return check_simple(n, nnz, off, ind);
}

int tch_peek_sizes(const char* path, int64_t* nv, int64_t* nnz) {
  if (!path || !nv || !nnz) return fail(TCH_EINVAL, "invalid argument");
  const std::string p(path);
  const bool bin = p.size() >= 4 && p.compare(p.size() - 4, 4, ".bin") == 0;
  const std::string file = bin ? p : p + ".info";
  FILE* f = std::fopen(file.c_str(), "rb");
  if (!f) return fail(TCH_EIO, "cannot read " + file);
  char head[160] = {};
  const size_t got = std::fread(head, 1, sizeof head - 1, f);
  std::fclose(f);
  if (bin) {
    int64_t hdr[2];
    if (got < 24 || (std::memcmp(head, "TCCSR001", 8) != 0 && std::memcmp(head, "TCCSR064", 8) != 0))
      return fail(TCH_EIO, file + ": bad header");
    std::memcpy(hdr, head + 8, 16);
    *nv = hdr[0];
    *nnz = hdr[1];
  } else {
    Scanner sc{head, head + got};
    int64_t nv1;
    if (!sc.next(nv1) || !sc.next(*nnz) || nv1 < 1) return fail(TCH_EIO, file + ": bad header");
    *nv = nv1 - 1;
  }
  if (*nv < 0 || *nnz < 0) return fail(TCH_EIO, file + ": bad header");
  return TCH_OK;
}

tch_graph64* tch_read_info64(const char* path_no_ext) {
  if (!path_no_ext) return fail_null<tch_graph64>("invalid argument");
  std::string path = std::string(path_no_ext) + ".info";
  std::string buf;
  if (!slurp(path.c_str(), buf)) return (g_err = "cannot read " + path, nullptr);
  Scanner sc{buf.data(), buf.data() + buf.size()};
  int64_t nv1, nnz;
  if (!sc.next(nv1) || !sc.next(nnz) || nv1 < 1 || nnz < 0) return fail_null<tch_graph64>(path + ": bad header");
  if (nv1 - 1 > INT32_MAX) return fail_null<tch_graph64>(path + ": vertex count exceeds int32");
  // every number takes at least a digit and a separator: a header asking for
  // more than the file can hold is refused before anything is allocated
  if ((uint64_t)nv1 + (uint64_t)nnz > buf.size() / 2 + 1) return fail_null<tch_graph64>(path + ": truncated indices");
  auto* g = new (std::nothrow) tch_graph64;
  if (!g) return fail_null<tch_graph64>("out of memory");
  g->n = nv1 - 1;
  try {
    g->off.resize((size_t)nv1);
    g->ind.resize((size_t)nnz);
  } catch (...) {
    delete g;
    return fail_null<tch_graph64>("out of memory");
  }
  for (int64_t i = 0; i < nv1; ++i)
    if (!sc.next(g->off[i])) {
      delete g;
      return fail_null<tch_graph64>(path + ": truncated offsets");
    }
  for (int64_t i = 0; i < nnz; ++i)
    if (!sc.next(g->ind[i])) {
      delete g;
      return fail_null<tch_graph64>(path + ": truncated indices");
    }
  return g;
}

int tch_write_bin64(const tch_graph64* g, const char* path) {
  if (!g || !path) return fail(TCH_EINVAL, "invalid argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(TCH_EIO, std::string("cannot open ") + path);
  char magic[8] = {'T', 'C', 'C', 'S', 'R', '0', '6', '4'};
  int64_t hdr[3] = {g->n, (int64_t)g->ind.size(), 0};
  bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hdr, 8, 3, f) == 3 &&
            std::fwrite(g->off.data(), 8, g->off.size(), f) == g->off.size() &&
            std::fwrite(g->ind.data(), 8, g->ind.size(), f) == g->ind.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? TCH_OK : fail(TCH_EIO, std::string("write failed: ") + path);
}

tch_graph64* tch_read_bin64(const char* path) {
  if (!path) return fail_null<tch_graph64>("invalid argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail_null<tch_graph64>(std::string("cannot read ") + path);
  char magic[8];
  int64_t hdr[3];
  const bool ok_hdr = std::fread(magic, 1, 8, f) == 8 && std::fread(hdr, 8, 3, f) == 3;
  const bool wide = ok_hdr && std::memcmp(magic, "TCCSR064", 8) == 0;
  const bool narrow = ok_hdr && std::memcmp(magic, "TCCSR001", 8) == 0;
  if (!(wide || narrow) || hdr[0] < 0 || hdr[0] > INT32_MAX || hdr[1] < 0 || (narrow && hdr[1] > INT32_MAX)) {
    std::fclose(f);
    return fail_null<tch_graph64>(std::string(path) + ": bad header");
  }
  // the header may not ask for more than the file holds
  const long body = std::ftell(f);
  std::fseek(f, 0, SEEK_END);
  const long fsize = std::ftell(f);
  std::fseek(f, body, SEEK_SET);
  const uint64_t width = wide ? 8 : 4;
  if (fsize < body || (uint64_t)(fsize - body) < width * ((uint64_t)hdr[0] + 1 + (uint64_t)hdr[1])) {
    std::fclose(f);
    return fail_null<tch_graph64>(std::string(path) + ": truncated");
  }
  auto* g = new (std::nothrow) tch_graph64;
  if (!g) {
    std::fclose(f);
    return fail_null<tch_graph64>("out of memory");
  }
  g->n = hdr[0];
  bool ok = true;
  try {
    g->off.resize((size_t)hdr[0] + 1);
    g->ind.resize((size_t)hdr[1]);
  } catch (...) {
    ok = false;
  }
  if (ok && wide) {
    ok = std::fread(g->off.data(), 8, g->off.size(), f) == g->off.size() &&
         std::fread(g->ind.data(), 8, g->ind.size(), f) == g->ind.size();
  } else if (ok) {
    // the int32 cache, widened in blocks
    std::vector<int32_t> blk(1 << 20);
    for (int pass = 0; pass < 2 && ok; ++pass) {
      std::vector<int64_t>& dst = pass ? g->ind : g->off;
      for (size_t at = 0; at < dst.size() && ok;) {
        const size_t k = std::min(blk.size(), dst.size() - at);
        ok = std::fread(blk.data(), 4, k, f) == k;
        for (size_t i = 0; ok && i < k; ++i) dst[at + i] = blk[i];
        at += k;
      }
    }
  }
  std::fclose(f);
  if (!ok) {
    delete g;
    return fail_null<tch_graph64>(std::string(path) + ": truncated or out of memory");
  }
  return g;
}

// ---------------------------------------------------------------- batches
static tch_batch* pack(std::vector<tch_graph*>& gs, std::vector<int32_t>& kinds) {
  auto* b = new (std::nothrow) tch_batch;
  if (!b) return fail_null<tch_batch>("out of memory");
  const int32_t G = (int32_t)gs.size();
  b->G = G;
  b->tab.resize(2 * ((size_t)G + 1));
  int64_t vb = 0, ab = 0;
  for (int32_t g = 0; g < G; ++g) {
    b->tab[2 * g] = vb;
    b->tab[2 * g + 1] = ab;
    vb += gs[g]->n;
    ab += (int64_t)gs[g]->ind.size();
  }
  b->tab[2 * G] = vb;
  b->tab[2 * G + 1] = ab;
  b->off.resize((size_t)(vb + G));
  b->ind.resize((size_t)ab);
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t g = 0; g < G; ++g) {
    std::copy(gs[g]->off.begin(), gs[g]->off.end(), b->off.begin() + b->tab[2 * g] + g);
    std::copy(gs[g]->ind.begin(), gs[g]->ind.end(), b->ind.begin() + b->tab[2 * g + 1]);
  }
  b->kind = kinds;
  return b;
}

tch_batch* tch_gen_batch(int32_t G, int32_t n_lo, int32_t n_hi, int32_t ef_lo, int32_t ef_hi,
                         uint32_t kinds_mask, uint64_t seed) {
  if (G <= 0 || n_lo < 2 || n_hi < n_lo || ef_lo < 1 || ef_hi < ef_lo || (kinds_mask & 7u) == 0)
    return fail_null<tch_batch>("invalid argument");
  std::vector<int32_t> allowed;
  for (int k = 0; k < 3; ++k)
    if (kinds_mask & (1u << k)) allowed.push_back(k);
  std::vector<tch_graph*> gs((size_t)G, nullptr);
  std::vector<int32_t> kinds((size_t)G);
  std::atomic<bool> bad{false};
  const double llo = std::log((double)n_lo), lhi = std::log((double)n_hi);
  int outer = omp_get_max_threads();
  // Graphs are generated concurrently, each single-threaded.
#pragma omp parallel for schedule(dynamic, 1) num_threads(outer)
  for (int32_t g = 0; g < G; ++g) {
    uint64_t gs_seed = splitmix64(seed ^ (uint64_t)g);
    uint64_t r0 = draw(gs_seed, 0), r1 = draw(gs_seed, 1), r2 = draw(gs_seed, 2);
    double lu = llo + (lhi - llo) * unit_double(r0);
    int32_t n = (int32_t)std::floor(std::exp(lu));
    n = std::min(std::max(n, n_lo), n_hi);
    int32_t kind = allowed[bounded(r1, (uint32_t)allowed.size())];
    kinds[g] = kind;
    uint64_t body = draw(gs_seed, 3);
    std::vector<int32_t> s, d;
    if (kind == 0) {
      gen_er(n, 4 * (int64_t)n, body, s, d);
    } else if (kind == 1) {
      int32_t ef = ef_lo + (int32_t)bounded(r2, (uint32_t)(ef_hi - ef_lo + 1));
      int32_t scale = 1;
      while (((int64_t)1 << scale) < n) ++scale;
      gen_rmat(scale, (int64_t)ef * n, 0.57, 0.19, 0.19, body, n, s, d);
    } else {
      gen_rgg(n, 6.0, body, s, d);
    }
    auto* gr = new (std::nothrow) tch_graph;
    if (!gr || !build_simple(n, (int64_t)s.size(), s.data(), d.data(), gr)) {
      delete gr;
      bad = true;
      continue;
    }
    gs[g] = gr;
  }
  tch_batch* b = bad ? nullptr : pack(gs, kinds);
  for (auto* gr : gs) delete gr;
  if (bad) return fail_null<tch_batch>("graph generation failed");
  return b;
}

tch_batch* tch_pack(int32_t G, const tch_graph* const* graphs) {
  if (G <= 0 || !graphs) return fail_null<tch_batch>("invalid argument");
  std::vector<tch_graph*> gs((size_t)G);
  std::vector<int32_t> kinds((size_t)G, -1);
  for (int32_t g = 0; g < G; ++g) {
    if (!graphs[g]) return fail_null<tch_batch>("null graph in pack");
    gs[g] = const_cast<tch_graph*>(graphs[g]);
  }
  return pack(gs, kinds);
}

int32_t tch_batch_count(const tch_batch* b) { return b ? b->G : 0; }
const int64_t* tch_batch_tab(const tch_batch* b) { return b ? b->tab.data() : nullptr; }
const int32_t* tch_batch_off(const tch_batch* b) { return b ? b->off.data() : nullptr; }
const int32_t* tch_batch_ind(const tch_batch* b) { return b ? b->ind.data() : nullptr; }
const int32_t* tch_batch_kind(const tch_batch* b) { return b ? b->kind.data() : nullptr; }

tch_graph* tch_batch_graph(const tch_batch* b, int32_t g) {
  if (!b || g < 0 || g >= b->G) return fail_null<tch_graph>("invalid argument");
  auto* gr = new (std::nothrow) tch_graph;
  if (!gr) return fail_null<tch_graph>("out of memory");
  int64_t vb = b->tab[2 * g], ab = b->tab[2 * g + 1];
  int64_t n = b->tab[2 * g + 2] - vb, nnz = b->tab[2 * g + 3] - ab;
  gr->n = (int32_t)n;
  gr->off.assign(b->off.begin() + vb + g, b->off.begin() + vb + g + n + 1);
  gr->ind.assign(b->ind.begin() + ab, b->ind.begin() + ab + nnz);
  return gr;
}

void tch_batch_free(tch_batch* b) { delete b; }

}  // extern "C"

// ------------------------------------------------------------ admission sim
#include "admission.h"

extern "C" {

int64_t tch_request_bytes(int64_t nv, int64_t nnz) {
  if (nv < 0 || nnz < 0) return fail(TCH_EINVAL, "invalid argument");
  return (int64_t)tcs::request_bytes(nv, nnz);
}

int64_t tch_batch_block_bytes(int32_t G, int64_t nv_total, int64_t nnz_total) {
  if (G < 1 || nv_total < 0 || nnz_total < 0) return fail(TCH_EINVAL, "invalid argument");
  return (int64_t)tcs::batch_layout(G, nv_total, nnz_total).bytes;
}

int64_t tch_pack_block(int32_t G, const int32_t* nv, const int64_t* nnz, const int32_t* const* off,
                       const int32_t* const* ind, void* block, int64_t block_bytes) {
  if (G < 1 || !nv || !nnz || !off || !ind || !block) return fail(TCH_EINVAL, "invalid argument");
  int64_t v = 0, a = 0;
  for (int32_t g = 0; g < G; ++g) {
    if (nv[g] < 0 || nnz[g] < 0 || !off[g] || (nnz[g] > 0 && !ind[g])) return fail(TCH_EINVAL, "invalid graph");
    v += nv[g];
    a += nnz[g];
  }
  if ((int64_t)tcs::batch_layout(G, v, a).upload_end > block_bytes) return fail(TCH_EINVAL, "block too small");
  return (int64_t)tcs::pack_block(G, nv, nnz, off, ind, static_cast<char*>(block)).upload_end;
}

int tch_copy_pool_run(int32_t threads, int32_t callers, int32_t rounds, int32_t n, void* const* dst,
                      const void* const* src, const int64_t* bytes) {
  if (threads < 0 || callers < 1 || rounds < 0 || n < 0 || (n > 0 && (!dst || !src || !bytes)))
    return fail(TCH_EINVAL, "invalid argument");
  tcs::CopyPool pool(threads);
  std::vector<std::thread> cs;
  for (int32_t c = 0; c < callers; ++c)
    cs.emplace_back([&, c] {
      std::vector<tcs::CopyTask> t;
      for (int32_t i = 0; i < n; ++i)
        t.push_back({static_cast<char*>(dst[(size_t)c * n + i]), static_cast<const char*>(src[i]), (size_t)bytes[i]});
      for (int32_t r = 0; r < rounds; ++r) pool.run(t);
    });
  for (auto& t : cs) t.join();
  return TCH_OK;
}

int tch_admission_sim(int64_t budget, int32_t max_batch, int32_t nsteps, const int64_t* script,
                      int32_t* admitted_at, int32_t* batch_of, int64_t* peak_used) {
  if (budget <= 0 || max_batch < 1 || nsteps < 0 || (nsteps > 0 && (!script || !admitted_at || !batch_of)))
    return fail(TCH_EINVAL, "invalid argument");
  tcs::Admission adm((size_t)budget, tcs::kBatchFixed);  // the policy alone: shares are given in bytes
  struct Blk {
    size_t off, bytes;
    int32_t live;
  };
  std::vector<Blk> blks;
  std::vector<int32_t> state;  // per request: 0 waiting, 1 admitted, 2 completed
  std::vector<int64_t> ids;
  int admitted = 0;
  for (int32_t i = 0; i < nsteps; ++i) {
    const int64_t x = script[i];
    if (x > 0) {
      if (!adm.fits_ever((size_t)x)) return fail(TCH_ENOMEM, "request larger than the budget");
      const int32_t r = (int32_t)state.size();
      adm.push(r, (size_t)x);
      state.push_back(0);
      admitted_at[r] = -1;
      batch_of[r] = -1;
    } else if (x < 0) {
      const int64_t r = -x - 1;
      if (r >= (int64_t)state.size() || state[r] != 1) return fail(TCH_EINVAL, "completion of a request not in flight");
      state[r] = 2;
      Blk& b = blks[batch_of[r]];
      if (--b.live == 0) adm.arena().release(b.off, b.bytes);
    }
    for (;;) {
      size_t bytes = 0;
      const size_t off = adm.admit_batch(max_batch, ids, bytes);
      if (off == SIZE_MAX) break;
      for (int64_t r : ids) {
        state[r] = 1;
        admitted_at[r] = i;
        batch_of[r] = (int32_t)blks.size();
      }
      blks.push_back(Blk{off, bytes, (int32_t)ids.size()});
      admitted += (int)ids.size();
    }
  }
  if (peak_used) *peak_used = (int64_t)adm.arena().peak();
  return admitted;
}

}  // extern "C"
