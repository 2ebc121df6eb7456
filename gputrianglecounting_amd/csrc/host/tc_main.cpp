// tc_main — command-line driver of the gfx950 triangle counter.
//
// Counterpart of the reference binary bin/ccgpunew (src/main.cu:1-38,
// src/gpurun.cu:2-56, 115-176, src/param.cu:23-153):
//
//   tc_main <graph> <run> <blocks> [<blockSize> <threadsPerIntersection>]
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
// count (degree-oriented) runs.  Without a block size it sweeps the
// oriented scheme's partners per wave and the compiled edge-scheme kernel
// variants instead of the reference's 48 legacy shapes and writes
// "<graph>.o.<blocks>" rows: scheme algo wg_size items ring chunk depth
// row_groups allocMs execMs sumMs total.
// Inputs that are not simple undirected graphs are normalised (self-loops
// and duplicate arcs dropped, reverse arcs added) with a warning.
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
               "usage: tc_main <graph> <run> <blocks> [<blockSize> <threadsPerIntersection> [<partitions>]]\n"
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

int main(int argc, char** argv) {
  if (argc == 3 && std::strcmp(argv[1], "--copy") == 0) return copy_mode(std::atoll(argv[2]));
  if (argc < 4) {
    usage();
    return 2;
  }
  const std::string spec = argv[1];
  const long long blocks = std::atoll(argv[3]);
  const bool single = argc > 4;
  long long bs = 0, tpi = 0;
  if (blocks <= 0) {
    std::fprintf(stderr, "Number of blocks not provided for GPU(1) run.\n");
    return 2;
  }
  if (single) {
    bs = std::atoll(argv[4]);
    static const long long kBs[] = {32, 64, 96, 128, 192, 256, 512, 1024};
    static const long long kTpi[] = {1, 2, 4, 8, 16, 32};
    if (std::find(std::begin(kBs), std::end(kBs), bs) == std::end(kBs)) {
      std::fprintf(stderr, "Invalid block size parameter(32,64,96,128,192,256,512,1024).\n");
      return 2;
    }
    if (argc < 6) {
      std::fprintf(stderr, "Threads per intersection not provided for single GPU run.\n");
      return 2;
    }
    tpi = std::atoll(argv[5]);
    if (std::find(std::begin(kTpi), std::end(kTpi), tpi) == std::end(kTpi)) {
      std::fprintf(stderr, "Invalid threads per intersection(1,2,4,8,16,32).\n");
      return 2;
    }
  }

  tch_graph* g = load(spec);
  if (!g) {
    std::fprintf(stderr, "cannot load %s: %s\n", spec.c_str(), tch_last_error());
    return 1;
  }
  int32_t nv = tch_graph_n(g);
  int64_t nnz = tch_graph_nnz(g);
  if (tch_check_simple(nv, nnz, tch_graph_off(g), tch_graph_ind(g)) != TCH_OK) {
    std::fprintf(stderr, "warning: %s; normalising to a simple undirected graph\n", tch_last_error());
    tch_graph* h = tch_normalize(g);
    tch_graph_free(g);
    if (!h) {
      std::fprintf(stderr, "normalise failed: %s\n", tch_last_error());
      return 1;
    }
    g = h;
    nv = tch_graph_n(g);
    nnz = tch_graph_nnz(g);
  }

  tc_ctx* ctx = nullptr;
  if (tc_ctx_create(0, &ctx) != TC_OK) {
    std::fprintf(stderr, "tc_ctx_create failed\n");
    return 1;
  }
  HIPCHECK(hipFree(nullptr));  // context warm-up, as gpurun.cu:10-12

  auto t0 = clk::now();
  int32_t *d_off = nullptr, *d_ind = nullptr, *d_out = nullptr;
  int64_t* d_total = nullptr;
  HIPCHECK(hipMalloc(&d_off, 4 * ((size_t)nv + 1)));
  HIPCHECK(hipMalloc(&d_ind, 4 * std::max<size_t>((size_t)nnz, 1)));
  HIPCHECK(hipMalloc(&d_out, 4 * std::max<size_t>((size_t)nv, 1)));
  HIPCHECK(hipMalloc(&d_total, 8));
  HIPCHECK(hipMemcpy(d_off, tch_graph_off(g), 4 * ((size_t)nv + 1), hipMemcpyHostToDevice));
  if (nnz) HIPCHECK(hipMemcpy(d_ind, tch_graph_ind(g), 4 * (size_t)nnz, hipMemcpyHostToDevice));
  tc_ctx_reserve(ctx, tc_scratch_bytes(nv, nnz));
  HIPCHECK(hipDeviceSynchronize());
  const double alloc_ms = ms_since(t0);

  int rc = 0;
  if (single && argc > 6) {
    // Partition run (param.cuh PAR_PARTITION): <graph>.part.K holds a METIS
    // partition vector; each partition's vertices are counted as one vertex
    // list (the reference's vertex-list kernelCall) and the totals summed.
    const int32_t K = std::atoi(argv[6]);
    int32_t npart = 0, *part = nullptr;
    if (K < 1 || tch_read_partition(spec.c_str(), K, &npart, &part) != TCH_OK || npart != nv) {
      std::fprintf(stderr, "partition file: %s\n", K < 1 ? "bad partition count" : npart != nv ? "length differs from nv" : tch_last_error());
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
      if (tch_partition_members(nv, part, p, &cnt, &ids) != TCH_OK) return 1;
      if (cnt) HIPCHECK(hipMemcpy(d_list, ids, 4 * (size_t)cnt, hipMemcpyHostToDevice));
      rc = tc_count_vertices_i32(ctx, nv, nnz, d_off, d_ind, cnt, d_list, d_out, d_pt, nullptr, nullptr);
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
    std::printf("%d\t%lld\t%g\t%g\t%g\n", nv, (long long)total, alloc_ms, count_ms, ms_since(t0));
  } else if (single) {
    auto t1 = clk::now();
    rc = tc_count_i32(ctx, nv, nnz, d_off, d_ind, d_out, d_total, nullptr, nullptr);
    HIPCHECK(hipDeviceSynchronize());
    const double count_ms = ms_since(t1);
    if (rc != TC_OK) {
      std::fprintf(stderr, "count failed: %s\n", tc_last_error(ctx));
      return 1;
    }
    int64_t total = 0;
    HIPCHECK(hipMemcpy(&total, d_total, 8, hipMemcpyDeviceToHost));
    std::printf("%d\t%lld\t%g\t%g\t%g\n", nv, (long long)total, alloc_ms, count_ms, ms_since(t0));
  } else {
    // next to the graph file, as gpurun.cu:146-149; a generator spec names a file in the working directory
    std::string base = spec;
    if (spec.rfind("gen:", 0) == 0)
      for (char& ch : base)
        if (ch == ':' || ch == '/') ch = '_';
    std::string fname = base + ".o." + std::to_string(blocks);
    FILE* f = std::fopen(fname.c_str(), "w");
    if (!f) {
      std::fprintf(stderr, "cannot write %s\n", fname.c_str());
      return 1;
    }
    // {oriented, algo, wg_size, items, ring, chunk, depth, row_groups}; 0 = the context default
    static const int kVar[][8] = {
        {1, 0, 0, 0, 0, 0, 0, 2},          {1, 0, 0, 0, 0, 0, 0, 4},          {1, 0, 0, 0, 0, 0, 0, 8},
        {1, 0, 0, 0, 0, 0, 0, 16},         {1, 0, 0, 0, 0, 0, 0, 32},         {-1, 0, 256, 8, 4096, 16, 2, 0},
        {-1, 0, 256, 8, 4096, 16, 4, 0},   {-1, 0, 128, 16, 4096, 16, 2, 0},  {-1, 0, 512, 4, 4096, 16, 2, 0},
        {-1, 3, 256, 8, 4096, 16, 2, 0},   {-1, 3, 256, 8, 4096, 16, 4, 0},   {-1, 3, 512, 4, 4096, 16, 2, 0},
        {-1, 3, 128, 16, 4096, 16, 2, 0},  {-1, 2, 256, 8, 4096, 0, 0, 0}};
    for (const auto& v : kVar) {
      tc_cfg cfg{};
      cfg.oriented = v[0];
      cfg.algo = v[1];
      cfg.wg_size = v[2];
      cfg.items = v[3];
      cfg.ring = v[4];
      cfg.chunk = v[5];
      cfg.depth = v[6];
      cfg.row_groups = v[7];
      auto t1 = clk::now();
      rc = tc_count_i32(ctx, nv, nnz, d_off, d_ind, d_out, d_total, &cfg, nullptr);
      int64_t total = 0;
      HIPCHECK(hipMemcpy(&total, d_total, 8, hipMemcpyDeviceToHost));
      const double ex = ms_since(t1);
      if (rc != TC_OK) {
        std::fprintf(stderr, "count failed: %s\n", tc_last_error(ctx));
        break;
      }
      std::fprintf(f, "%s\t%d\t%d\t%d\t%d\t%d\t%d\t%d\t%g\t%g\t%g\t%lld\n", v[0] > 0 ? "oriented" : "edges", v[1], v[2],
                   v[3], v[4], v[5], v[6], v[7], alloc_ms, ex, ex + alloc_ms, (long long)total);
    }
    std::fclose(f);
    std::printf("wrote %s\n", fname.c_str());
  }
  (void)hipFree(d_off);
  (void)hipFree(d_ind);
  (void)hipFree(d_out);
  (void)hipFree(d_total);
  tc_ctx_destroy(ctx);
  tch_graph_free(g);
  return rc == TC_OK ? 0 : 1;
}
