// tc_main — command-line driver of the gfx950 triangle counter.
//
// Counterpart of the reference binary bin/ccgpunew (src/main.cu:1-38,
// src/gpurun.cu:2-56, 115-176, src/param.cu:23-153):
//
//   tc_main [--int64] [--variants] <graph> <run> <blocks> [<blockSize> <threadsPerIntersection> [<partitions>]]
//
// <graph> is a path without extension whose "<graph>.info" text CSR is read
// (graphRead.cpp:2-26), or one of
//   gen:er:<n>:<m>:<seed>   gen:rmat:<scale>:<edgefactor>:<seed>
//   gen:rgg:<n>:<meandeg>:<seed>   <file>.bin (binary CSR)   <file>.graph (METIS)
// <run> is accepted and ignored as in the reference (param.cu:54-66).
// With a block size and threads-per-intersection the single run prints, like
// gpurun.cu:49-51,
//   nv <tab> total <tab> memAllocMs <tab> tCountMs <tab> execMs
// where total is the int64 sum of the per-vertex counts (6·T).  The legacy
// launch shape is validated exactly as param.cu:225-247 does; the default
// count (degree-oriented) runs.  Without a block size it writes the
// reference's sweep file "<graph>.o.<blocks>" (gpurun.cu:146-172): one row
// "bs tpi allocMs execMs sumMs total" per legacy shape, block sizes
// {32..1024} outer, threads per intersection {1..32} inner (param.cu:4-5),
// each shape run as its nearest configuration of the oriented count: the
// workgroup of bs threads (128 for bs <= 128, else 256) and tpi lanes per
// intersected list (row_groups = 64 / max(tpi, 2) lists per wave).
// --variants writes instead one row per oriented partners-per-wave setting
// and per compiled edge-scheme variant: scheme algo wg_size items ring chunk
// depth row_groups allocMs execMs sumMs total.
// The reference binary counts in int64 (main.cu:36, run<int64_t>).  Here the
// int32 CSR is the default; --int64, a .info/.bin file with 2^31 arcs or
// more, or an int64 binary cache, reads the graph as int64 (tch_read_info64
// / tch_read_bin64) and counts with tc_count_i64.
// int32 inputs that are not simple undirected graphs are normalised
// (self-loops and duplicate arcs dropped, reverse arcs added) with a
// warning; int64 ones are refused.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tc.h"
#include "tc_host.h"

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

std::vector<std::string> split(const std::string& s, char c) {
  std::vector<std::string> out;
  size_t p = 0;
  for (;;) {
    size_t q = s.find(c, p);
    out.push_back(s.substr(p, q == std::string::npos ? std::string::npos : q - p));
    if (q == std::string::npos) break;
    p = q + 1;
  }
  return out;
}

bool ends_with(const std::string& s, const char* suf) {
  size_t n = std::strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

tch_graph* load(const std::string& spec) {
  if (spec.rfind("gen:", 0) == 0) {
    auto f = split(spec, ':');
    if (f.size() == 5 && f[1] == "er")
      return tch_gen_er(std::atoi(f[2].c_str()), std::atoll(f[3].c_str()), std::strtoull(f[4].c_str(), nullptr, 10));
    if (f.size() == 5 && f[1] == "rmat") {
      int scale = std::atoi(f[2].c_str());
      long long ef = std::atoll(f[3].c_str());
      return tch_gen_rmat(scale, ef << scale, 0.57, 0.19, 0.19, std::strtoull(f[4].c_str(), nullptr, 10), 0);
    }
    if (f.size() == 5 && f[1] == "rgg")
      return tch_gen_rgg(std::atoi(f[2].c_str()), std::atof(f[3].c_str()), std::strtoull(f[4].c_str(), nullptr, 10));
    std::fprintf(stderr, "unknown generator spec %s\n", spec.c_str());
    return nullptr;
  }
  if (ends_with(spec, ".bin")) return tch_read_bin(spec.c_str());
  if (ends_with(spec, ".graph") || ends_with(spec, ".metis")) return tch_read_metis(spec.c_str());
  return tch_read_info(spec.c_str());
}

#define HIPCHECK(x)                                                                         \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) {                                                                 \
      std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return 1;                                                                             \
    }                                                                                       \
  } while (0)

void usage() {
  std::fprintf(stderr,
               "usage: tc_main [--int64] [--variants] <graph> <run> <blocks> [<blockSize> <threadsPerIntersection>\n"
               "               [<partitions>]]\n"
               "  <graph>: path without .info | gen:er:n:m:seed | gen:rmat:scale:ef:seed |\n"
               "           gen:rgg:n:deg:seed | file.bin | file.graph\n");
}

}  // namespace

// tc_main --copy <MiB>: two plain device copies of that size (the second
// is the one to read from a counter profile), for byte-counter calibration.
int copy_mode(long long mib) {
  tc_ctx* ctx = nullptr;
  if (mib <= 0 || tc_ctx_create(0, &ctx) != TC_OK) return 2;
  const size_t n = (size_t)mib * 1024 * 1024 / 4;
  int32_t *a = nullptr, *b = nullptr;
  HIPCHECK(hipMalloc(&a, 4 * n));
  HIPCHECK(hipMalloc(&b, 4 * n));
  HIPCHECK(hipMemset(a, 1, 4 * n));
  for (int r = 0; r < 2; ++r)
    if (tc_copy_i32(ctx, b, a, (int64_t)n, nullptr) != TC_OK) return 1;
  HIPCHECK(hipDeviceSynchronize());
  std::printf("copied %lld MiB twice\n", mib);
  (void)hipFree(a);
  (void)hipFree(b);
  tc_ctx_destroy(ctx);
  return 0;
}

namespace {

// The count entry point of each CSR width.
int count(tc_ctx* c, int64_t nv, int64_t nnz, const int32_t* off, const int32_t* ind, int32_t* out, int64_t* total,
          const tc_cfg* cfg) {
  return tc_count_i32(c, (int32_t)nv, nnz, off, ind, out, total, cfg, nullptr);
}
int count(tc_ctx* c, int64_t nv, int64_t nnz, const int64_t* off, const int64_t* ind, int64_t* out, int64_t* total,
          const tc_cfg* cfg) {
  return tc_count_i64(c, nv, nnz, off, ind, out, total, cfg, nullptr);
}

struct Run {
  std::string spec;
  long long blocks = 0, bs = 0, tpi = 0;
  int parts = 0;  // <partitions> argument, 0 = none
  bool single = false, variants = false;
};

// <graph>.o.<blocks> beside the graph file (gpurun.cu:146-149); a
// generator spec names a file in the working directory.
std::string sweep_name(const Run& r) {
  std::string base = r.spec;
  if (r.spec.rfind("gen:", 0) == 0)
    for (char& ch : base)
      if (ch == ':' || ch == '/') ch = '_';
  return base + ".o." + std::to_string(r.blocks);
}

// Uploads the CSR, then runs what the arguments ask for: the single run,
// the partition run (int32 only), the reference sweep or the variant sweep.
template <class T>
int run_graph(const Run& r, int64_t nv, int64_t nnz, const T* h_off, const T* h_ind) {
  tc_ctx* ctx = nullptr;
  if (tc_ctx_create(0, &ctx) != TC_OK) {
    std::fprintf(stderr, "tc_ctx_create failed\n");
    return 1;
  }
  HIPCHECK(hipFree(nullptr));  // context warm-up, as gpurun.cu:10-12

  auto t0 = clk::now();
  T *d_off = nullptr, *d_ind = nullptr, *d_out = nullptr;
  int64_t* d_total = nullptr;
  HIPCHECK(hipMalloc(&d_off, sizeof(T) * ((size_t)nv + 1)));
  HIPCHECK(hipMalloc(&d_ind, sizeof(T) * std::max<size_t>((size_t)nnz, 1)));
  HIPCHECK(hipMalloc(&d_out, sizeof(T) * std::max<size_t>((size_t)nv, 1)));
  HIPCHECK(hipMalloc(&d_total, 8));
  HIPCHECK(hipMemcpy(d_off, h_off, sizeof(T) * ((size_t)nv + 1), hipMemcpyHostToDevice));
  if (nnz) HIPCHECK(hipMemcpy(d_ind, h_ind, sizeof(T) * (size_t)nnz, hipMemcpyHostToDevice));
  tc_ctx_reserve(ctx, tc_scratch_bytes((int32_t)nv, nnz));
  HIPCHECK(hipDeviceSynchronize());
  const double alloc_ms = ms_since(t0);

  int rc = 0;
  if (r.single && r.parts > 0) {
    if (sizeof(T) != sizeof(int32_t)) {
      std::fprintf(stderr, "partition runs take the int32 CSR (fewer than 2^31 arcs, no --int64)\n");
      return 1;
    }
    // Partition run (param.cuh PAR_PARTITION): <graph>.part.K holds a METIS
    // partition vector; each partition's vertices are counted as one vertex
    // list (the reference's vertex-list kernelCall) and the totals summed.
    const int32_t K = r.parts;
    int32_t npart = 0, *part = nullptr;
    if (tch_read_partition(r.spec.c_str(), K, &npart, &part) != TCH_OK || npart != nv) {
      std::fprintf(stderr, "partition file: %s\n", npart != nv && part ? "length differs from nv" : tch_last_error());
      return 1;
    }
    int32_t* d_list = nullptr;
    HIPCHECK(hipMalloc(&d_list, 4 * std::max<size_t>((size_t)nv, 1)));
    int64_t* d_pt = nullptr;
    HIPCHECK(hipMalloc(&d_pt, 8));
    auto t1 = clk::now();
    int64_t total = 0;
    for (int32_t p = 0; p < K && rc == TC_OK; ++p) {
      int32_t cnt = 0, *ids = nullptr;
      if (tch_partition_members((int32_t)nv, part, p, &cnt, &ids) != TCH_OK) return 1;
      if (cnt) HIPCHECK(hipMemcpy(d_list, ids, 4 * (size_t)cnt, hipMemcpyHostToDevice));
      rc = tc_count_vertices_i32(ctx, (int32_t)nv, nnz, reinterpret_cast<const int32_t*>(d_off),
                                 reinterpret_cast<const int32_t*>(d_ind), cnt, d_list,
                                 reinterpret_cast<int32_t*>(d_out), d_pt, nullptr, nullptr);
      int64_t pt = 0;
      HIPCHECK(hipMemcpy(&pt, d_pt, 8, hipMemcpyDeviceToHost));
      total += pt;
      tch_free(ids);
    }
    const double count_ms = ms_since(t1);
    tch_free(part);
    (void)hipFree(d_list);
    (void)hipFree(d_pt);
    if (rc != TC_OK) {
      std::fprintf(stderr, "count failed: %s\n", tc_last_error(ctx));
      return 1;
    }
    std::printf("%lld\t%lld\t%g\t%g\t%g\n", (long long)nv, (long long)total, alloc_ms, count_ms, ms_since(t0));
  } else if (r.single) {
    auto t1 = clk::now();
    rc = count(ctx, nv, nnz, d_off, d_ind, d_out, d_total, nullptr);
    HIPCHECK(hipDeviceSynchronize());
    const double count_ms = ms_since(t1);
    if (rc != TC_OK) {
      std::fprintf(stderr, "count failed: %s\n", tc_last_error(ctx));
      return 1;
    }
    int64_t total = 0;
    HIPCHECK(hipMemcpy(&total, d_total, 8, hipMemcpyDeviceToHost));
    std::printf("%lld\t%lld\t%g\t%g\t%g\n", (long long)nv, (long long)total, alloc_ms, count_ms, ms_since(t0));
  } else {
    const std::string fname = sweep_name(r);
    FILE* f = std::fopen(fname.c_str(), "w");
    if (!f) {
      std::fprintf(stderr, "cannot write %s\n", fname.c_str());
      return 1;
    }
    // One timed count: launch, then the total back on the host (the
    // reference times kernelCall + thrust::reduce, gpurun.cu:157-168).
    auto timed = [&](const tc_cfg& cfg, double* ms, int64_t* total) {
      auto t1 = clk::now();
      int e = count(ctx, nv, nnz, d_off, d_ind, d_out, d_total, &cfg);
      if (e == TC_OK && hipMemcpy(total, d_total, 8, hipMemcpyDeviceToHost) != hipSuccess) e = TC_EHIP;
      *ms = ms_since(t1);
      return e;
    };
    if (!r.variants) {
      // The reference's 48 legacy shapes (param.cu:4-5), in its order.
      static const int kBs[] = {32, 64, 96, 128, 192, 256, 512, 1024};
      static const int kTpi[] = {1, 2, 4, 8, 16, 32};
      for (int b : kBs) {
        for (int t : kTpi) {
          tc_cfg cfg{};
          cfg.orient_wg = b <= 128 ? 128 : 256;
          cfg.row_groups = 64 / std::max(t, 2);
          double ex = 0;
          int64_t total = 0;
          if ((rc = timed(cfg, &ex, &total)) != TC_OK) break;
          std::fprintf(f, "%d\t%d\t%g\t%g\t%g\t%lld\n", b, t, alloc_ms, ex, ex + alloc_ms, (long long)total);
        }
        if (rc != TC_OK) break;
      }
    } else {
      // {oriented, algo, wg_size, items, ring, chunk, depth, row_groups}; 0 = the context default
      static const int kVar[][8] = {
          {1, 0, 0, 0, 0, 0, 0, 2},          {1, 0, 0, 0, 0, 0, 0, 4},          {1, 0, 0, 0, 0, 0, 0, 8},
          {1, 0, 0, 0, 0, 0, 0, 16},         {1, 0, 0, 0, 0, 0, 0, 32},         {-1, 0, 256, 8, 4096, 16, 2, 0},
          {-1, 0, 256, 8, 4096, 16, 4, 0},   {-1, 0, 128, 16, 4096, 16, 2, 0},  {-1, 0, 512, 4, 4096, 16, 2, 0},
          {-1, 3, 256, 8, 4096, 16, 2, 0},   {-1, 3, 256, 8, 4096, 16, 4, 0},   {-1, 3, 512, 4, 4096, 16, 2, 0},
          {-1, 3, 128, 16, 4096, 16, 2, 0},  {-1, 2, 256, 8, 4096, 0, 0, 0}};
      for (const auto& v : kVar) {
        if (sizeof(T) != sizeof(int32_t) && v[0] < 0) continue;  // the int64 path is oriented only
        tc_cfg cfg{};
        cfg.oriented = v[0];
        cfg.algo = v[1];
        cfg.wg_size = v[2];
        cfg.items = v[3];
        cfg.ring = v[4];
        cfg.chunk = v[5];
        cfg.depth = v[6];
        cfg.row_groups = v[7];
        double ex = 0;
        int64_t total = 0;
        if ((rc = timed(cfg, &ex, &total)) != TC_OK) break;
        std::fprintf(f, "%s\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%g\t%g\t%g\t%lld\n", v[0] > 0 ? "oriented" : "edges", v[1],
                     v[2], v[3], v[4], v[5], v[6], v[7], alloc_ms, ex, ex + alloc_ms, (long long)total);
      }
    }
    std::fclose(f);
    if (rc != TC_OK) {
      std::fprintf(stderr, "count failed: %s\n", tc_last_error(ctx));
      return 1;
    }
    std::printf("wrote %s\n", fname.c_str());
  }
  (void)hipFree(d_off);
  (void)hipFree(d_ind);
  (void)hipFree(d_out);
  (void)hipFree(d_total);
  tc_ctx_destroy(ctx);
  return rc == TC_OK ? 0 : 1;
}

// The int64 driver: the graph read as int64 (files) or widened (generators,
// METIS), refused unless simple.
int run_wide(const Run& r) {
  tch_graph64* g = nullptr;
  const std::string& spec = r.spec;
  if (ends_with(spec, ".bin")) {
    g = tch_read_bin64(spec.c_str());
  } else if (spec.rfind("gen:", 0) == 0 || ends_with(spec, ".graph") || ends_with(spec, ".metis")) {
    tch_graph* h = load(spec);
    if (h) {
      g = tch_widen(h);
      tch_graph_free(h);
    }
  } else {
    g = tch_read_info64(spec.c_str());
  }
  if (!g) {
    std::fprintf(stderr, "cannot load %s: %s\n", spec.c_str(), tch_last_error());
    return 1;
  }
  const int64_t nv = tch_graph64_n(g), nnz = tch_graph64_nnz(g);
  if (tch_check_simple64(nv, nnz, tch_graph64_off(g), tch_graph64_ind(g)) != TCH_OK) {
    std::fprintf(stderr, "not a simple undirected CSR: %s\n", tch_last_error());
    tch_graph64_free(g);
    return 1;
  }
  const int rc = run_graph<int64_t>(r, nv, nnz, tch_graph64_off(g), tch_graph64_ind(g));
  tch_graph64_free(g);
  return rc;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 3 && std::strcmp(argv[1], "--copy") == 0) return copy_mode(std::atoll(argv[2]));
  bool wide = false, variants = false;
  int a = 1;
  for (; a < argc && argv[a][0] == '-' && argv[a][1] == '-'; ++a) {
    if (std::strcmp(argv[a], "--int64") == 0) {
      wide = true;
    } else if (std::strcmp(argv[a], "--variants") == 0) {
      variants = true;
    } else {
      usage();
      return 2;
    }
  }
  argv += a - 1;
  argc -= a - 1;
  if (argc < 4) {
    usage();
    return 2;
  }
  Run r;
  r.spec = argv[1];
  r.blocks = std::atoll(argv[3]);
  r.single = argc > 4;
  r.variants = variants;
  if (r.blocks <= 0) {
    std::fprintf(stderr, "Number of blocks not provided for GPU(1) run.\n");
    return 2;
  }
  if (r.single) {
    r.bs = std::atoll(argv[4]);
    static const long long kBs[] = {32, 64, 96, 128, 192, 256, 512, 1024};
    static const long long kTpi[] = {1, 2, 4, 8, 16, 32};
    if (std::find(std::begin(kBs), std::end(kBs), r.bs) == std::end(kBs)) {
      std::fprintf(stderr, "Invalid block size parameter(32,64,96,128,192,256,512,1024).\n");
      return 2;
    }
    if (argc < 6) {
      std::fprintf(stderr, "Threads per intersection not provided for single GPU run.\n");
      return 2;
    }
    r.tpi = std::atoll(argv[5]);
    if (std::find(std::begin(kTpi), std::end(kTpi), r.tpi) == std::end(kTpi)) {
      std::fprintf(stderr, "Invalid threads per intersection(1,2,4,8,16,32).\n");
      return 2;
    }
    if (argc > 6) {
      r.parts = std::atoi(argv[6]);
      if (r.parts < 1) {
        std::fprintf(stderr, "partition file: bad partition count\n");
        return 1;
      }
    }
  }

  // Files of 2^31 arcs or more, and int64 binary caches, take the int64 CSR.
  if (!wide && r.spec.rfind("gen:", 0) != 0 && !ends_with(r.spec, ".graph") && !ends_with(r.spec, ".metis")) {
    int64_t pnv = 0, pnnz = 0;
    if (tch_peek_sizes(r.spec.c_str(), &pnv, &pnnz) == TCH_OK && pnnz > INT32_MAX) wide = true;
    if (ends_with(r.spec, ".bin")) {
      char magic[8] = {};
      if (FILE* f = std::fopen(r.spec.c_str(), "rb")) {
        wide = wide || (std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "TCCSR064", 8) == 0);
        std::fclose(f);
      }
    }
  }
  if (wide) return run_wide(r);

  tch_graph* g = load(r.spec);
  if (!g) {
    std::fprintf(stderr, "cannot load %s: %s\n", r.spec.c_str(), tch_last_error());
    return 1;
  }
  if (tch_check_simple(tch_graph_n(g), tch_graph_nnz(g), tch_graph_off(g), tch_graph_ind(g)) != TCH_OK) {
    std::fprintf(stderr, "warning: %s; normalising to a simple undirected graph\n", tch_last_error());
    tch_graph* h = tch_normalize(g);
    tch_graph_free(g);
    if (!h) {
      std::fprintf(stderr, "normalise failed: %s\n", tch_last_error());
      return 1;
    }
    g = h;
  }
  const int rc = run_graph<int32_t>(r, tch_graph_n(g), tch_graph_nnz(g), tch_graph_off(g), tch_graph_ind(g));
  tch_graph_free(g);
  return rc;
}
