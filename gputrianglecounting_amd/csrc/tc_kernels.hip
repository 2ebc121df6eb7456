// Device kernels of the gfx950 per-vertex triangle counter.
//
// Pipeline for one graph (all stream-ordered, no host synchronisation):
//   1. voff      int64 row offsets (single graph: widened off; batch: local
//                offsets of every graph rebased onto one vertex space)
//   2. edges     for every undirected edge u < w whose endpoints both have
//                degree >= 2 (the reference skips the others too,
//                clusteringCount.cu:339), write (u, w) and its Merge Path
//                length |N(u)| + |N(w)|; exclusive int64 prefix -> wpre
//   3. count     two-level Merge Path (k_count_mp): level 1 cuts [0, W)
//                into equal ranges, one per workgroup; level 2 splits every
//                LDS tile of a range into 64-lane wave diagonals
//   4. finalise  LCC / per-graph totals when asked
//
// The reference kernel (clusteringCount.cu:314-359) gives each block an
// equal number of vertices (calcWorkPerBlock, 248-264) and intersects both
// arcs of every edge; here each undirected edge is intersected once and its
// count added to both endpoints, and the work split is by cumulative
// intersection length, so a hub's edges spread over many workgroups.
#include <array>
#include <climits>
#include <map>
#include <mutex>
#include <type_traits>

#include "tc_device.h"

namespace tcd {
namespace {

constexpr int kScanBS = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBS * kScanItems;

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int32_t wave_sum32(int32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// The same sum by DPP: row shifts within rows of 16 lanes, then row
// broadcasts (lane 15 into row 1 and 3, lane 31 into rows 2 and 3), six
// adds with no LDS round trip; lane 63 ends with the total, read into every
// lane.  Every lane of the wave must be active.
__device__ __forceinline__ int32_t wave_sum32_dpp(int32_t v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return __builtin_amdgcn_readlane(v, 63);
}
// Sum over aligned groups of LG lanes, every lane of a group getting its
// group's sum; LG = 64 takes the DPP form (every lane active).
template <int LG>
__device__ __forceinline__ int32_t group_sum32(int32_t v) {
  if constexpr (LG == 64) {
    return wave_sum32_dpp(v);
  } else {
#pragma unroll
    for (int o = LG / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
  }
}

// Block-wide sum; every thread gets the result.  red must hold BS/64 slots.
template <int BS>
__device__ __forceinline__ int64_t block_sum64(int64_t v, int64_t* red) {
  constexpr int NW = BS / 64;
  v = wave_sum64(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  int64_t s = 0;
#pragma unroll
  for (int q = 0; q < NW; ++q) s += red[q];
  __syncthreads();
  return s;
}

// A value every lane of the wave holds, moved to scalar registers.
__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Inclusive scan over a wave of 64 lanes.
__device__ __forceinline__ int64_t wave_incl_scan64(int64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int64_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// ---------------------------------------------------------------- fills
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_i64(int64_t* p, int64_t n, int64_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_voff_from_off32(const int32_t* __restrict__ off, int64_t n1, int64_t* __restrict__ voff) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n1; i += (int64_t)gridDim.x * blockDim.x)
    voff[i] = off[i];
}

// Batch: global vertex v of graph g (v_base[g] <= v < v_base[g+1]) gets
// voff[v] = arc_base[g] + off_cat[v_base[g] + g + (v - v_base[g])].  The
// graphs' arcs are contiguous (arc_base[g+1] = arc_base[g] + nnz_g), so
// voff[v+1] ends every row, including the last row of each graph.
__global__ void k_voff_batch(const int64_t* __restrict__ tab, int32_t G, const int32_t* __restrict__ off_cat,
                             int64_t nv, int64_t* __restrict__ voff, int32_t* __restrict__ vbase) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v <= nv; v += (int64_t)gridDim.x * blockDim.x) {
    if (v == nv) {
      voff[nv] = tab[2 * G + 1];
      continue;
    }
    int32_t lo = 0, hi = G;  // last g with v_base[g] <= v
    while (hi - lo > 1) {
      int32_t mid = (lo + hi) >> 1;
      if (tab[2 * mid] <= v) lo = mid;
      else hi = mid;
    }
    const int64_t vb = tab[2 * lo];
    voff[v] = tab[2 * lo + 1] + off_cat[vb + lo + (v - vb)];
    vbase[v] = (int32_t)vb;
  }
}

// ---------------------------------------------------------------- scans
// Exclusive prefix sum of n elements (n = *n_dev when n_dev != nullptr,
// else n_host) into out[0..n], out[n] = total.  Three passes: per-tile sums,
// scan of the tile sums (recursively), per-tile scan with the tile prefix.
template <typename Tin>
__global__ __launch_bounds__(kScanBS) void k_scan_reduce(const Tin* __restrict__ in, const int64_t* n_dev,
                                                         int64_t n_host, int64_t* __restrict__ bsum) {
  __shared__ int64_t red[kScanBS / 64];
  const int64_t n = n_dev ? *n_dev : n_host;
  const int64_t t0 = (int64_t)blockIdx.x * kScanTile;
  int64_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = t0 + (int64_t)k * kScanBS + threadIdx.x;
    if (i < n) s += (int64_t)in[i];
  }
  s = block_sum64<kScanBS>(s, red);
  if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

template <typename Tin>
__global__ __launch_bounds__(kScanBS) void k_scan_apply(const Tin* __restrict__ in, const int64_t* n_dev,
                                                        int64_t n_host, const int64_t* __restrict__ bpre,
                                                        int64_t* __restrict__ out) {
  __shared__ int64_t tile[kScanTile];
  __shared__ int64_t wsum[kScanBS / 64];
  const int64_t n = n_dev ? *n_dev : n_host;
  const int64_t t0 = (int64_t)blockIdx.x * kScanTile;
  if (t0 >= n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;  // n == 0
    return;
  }
  for (int k = 0; k < kScanItems; ++k) {
    int idx = k * kScanBS + threadIdx.x;
    int64_t i = t0 + idx;
    tile[idx] = i < n ? (int64_t)in[i] : 0;
  }
  __syncthreads();
  int64_t run = 0;
  const int base = threadIdx.x * kScanItems;
  for (int k = 0; k < kScanItems; ++k) run += tile[base + k];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int64_t inc = wave_incl_scan64(run);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int64_t wpre = 0;
  for (int q = 0; q < wid; ++q) wpre += wsum[q];
  int64_t acc = (bpre ? bpre[blockIdx.x] : 0) + wpre + inc - run;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t v = tile[base + k];
    int64_t i = t0 + base + k;
    if (i < n) out[i] = acc;
    if (i == n - 1) out[n] = acc + v;
    acc += v;
  }
}

template <typename Tin>
hipError_t scan_excl(const Tin* in, const int64_t* n_dev, int64_t n_cap, int64_t* out, int64_t* tmp,
                     hipStream_t s) {
  int64_t nb = (n_cap + kScanTile - 1) / kScanTile;
  if (nb < 1) nb = 1;
  int64_t* bsum = tmp;
  int64_t* bpre = tmp + nb;
  int64_t* rest = bpre + nb + 1;
  if (nb > 1) {
    hipLaunchKernelGGL(k_scan_reduce<Tin>, dim3((unsigned)nb), dim3(kScanBS), 0, s, in, n_dev, n_cap, bsum);
    hipError_t e = scan_excl<int64_t>(bsum, nullptr, nb, bpre, rest, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_apply<Tin>, dim3((unsigned)nb), dim3(kScanBS), 0, s, in, n_dev, n_cap,
                       (const int64_t*)bpre, out);
  } else {
    hipLaunchKernelGGL(k_scan_apply<Tin>, dim3(1), dim3(kScanBS), 0, s, in, n_dev, n_cap,
                       (const int64_t*)nullptr, out);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------- edge list
// One wave per vertex.  mode 0: count neighbours w > v (graph-local ids)
// with d(v) >= 2 and d(w) >= 2.  mode 1 (oriented CSR): count neighbours w
// with d+(v) >= 1 and d+(w) >= 1.  mode 2 (vertex list): every arc v -> w of
// a marked v with w != v, d(v) >= 2 and d(w) >= 2 (the reference's per-arc
// rule, clusteringCount.cu:295-298).  mode 3 (vertex list with shadow
// vertices): the arcs of mode 2, those with w > lim twice (edge_copies).
__device__ __forceinline__ bool edge_keep(int mode, int32_t vl, int32_t wl, int64_t dv, int64_t dw) {
  if (mode == 0) return wl > vl && dv >= 2 && dw >= 2;
  if (mode == 2 || mode == 3) return wl != vl && dv >= 2 && dw >= 2;
  return dv >= 1 && dw >= 1;
}
// Copies of a kept arc v -> w in the edge list.  In the shadow-vertex
// count (clusteringCount.cu:150-177) a match x counts 1 + [x > lim]; on a
// simple graph the pairs (w, x) of N(v) with w ~ x are the pairs (x, w), so
// summing 1 + [x > lim] over them equals counting every arc v -> x once
// plus once more when x > lim.
__device__ __forceinline__ int edge_copies(int mode, int32_t wl, int64_t lim) { return mode == 3 && wl > lim ? 2 : 1; }
__device__ __forceinline__ int64_t edge_min_deg(int mode) { return mode == 1 ? 1 : 2; }

// Vertices are handed to waves in chunks of kVChunk from a counter (ctr),
// so hub rows do not pile up on one wave; each lane keeps 4 arc loads in
// flight.
constexpr int kVChunk = 16;

__device__ __forceinline__ int64_t grab_vertices(unsigned long long* ctr) {
  int64_t v0 = 0;
  if ((threadIdx.x & 63) == 0) v0 = (int64_t)atomicAdd(ctr, (unsigned long long)kVChunk);
  return uniform64(__shfl(v0, 0, 64));
}

__global__ __launch_bounds__(256) void k_edge_count(const int64_t* __restrict__ voff, const int32_t* __restrict__ ind,
                                                    const int32_t* __restrict__ vbase, int64_t nv, int mode,
                                                    const int32_t* __restrict__ vmark, int32_t* __restrict__ ecnt,
                                                    unsigned long long* ctr, int64_t lim) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    const int64_t v0 = grab_vertices(ctr);
    if (v0 >= nv) break;
    const int64_t v1 = v0 + kVChunk < nv ? v0 + kVChunk : nv;
    for (int64_t v = v0; v < v1; ++v) {
      const int64_t b = voff[v], e = voff[v + 1], dv = e - b;
      const int32_t vb = vbase ? vbase[v] : 0;
      const int32_t vl = (int32_t)(v - vb);
      int32_t c = 0;
      if (dv >= edge_min_deg(mode) && ((mode != 2 && mode != 3) || vmark[v])) {
        for (int64_t k0 = b; k0 < e; k0 += 256) {
          int32_t wl[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t k = k0 + lane + 64 * r;
            wl[r] = k < e ? ind[k] : -1;
          }
          int64_t lo[4], hi[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t wg = (int64_t)vb + (wl[r] < 0 ? 0 : wl[r]);
            lo[r] = voff[wg];
            hi[r] = voff[wg + 1];
          }
#pragma unroll
          for (int r = 0; r < 4; ++r)
            c += wl[r] >= 0 && edge_keep(mode, vl, wl[r], dv, hi[r] - lo[r]) ? edge_copies(mode, wl[r], lim) : 0;
        }
      }
      c = wave_sum32(c);
      if (lane == 0) ecnt[v] = c;
    }
  }
}

__global__ __launch_bounds__(256) void k_edge_fill(const int64_t* __restrict__ voff, const int32_t* __restrict__ ind,
                                                   const int32_t* __restrict__ vbase, int64_t nv, int mode,
                                                   const int32_t* __restrict__ vmark, const int64_t* __restrict__ eoff,
                                                   int64_t ecap,
                                                   int32_t* __restrict__ eu, int32_t* __restrict__ ew,
                                                   int32_t* __restrict__ ework, unsigned long long* ctr, int64_t lim) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (;;) {
    const int64_t v0 = grab_vertices(ctr);
    if (v0 >= nv) break;
    const int64_t v1 = v0 + kVChunk < nv ? v0 + kVChunk : nv;
    for (int64_t v = v0; v < v1; ++v) {
      const int64_t b = voff[v], e = voff[v + 1], dv = e - b;
      if (dv < edge_min_deg(mode) || ((mode == 2 || mode == 3) && !vmark[v])) continue;
      const int32_t vb = vbase ? vbase[v] : 0;
      const int32_t vl = (int32_t)(v - vb);
      int64_t pos = eoff[v];
      for (int64_t k0 = b; k0 < e; k0 += 256) {
        int32_t wl[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t k = k0 + lane + 64 * r;
          wl[r] = k < e ? ind[k] : -1;
        }
        int64_t lo[4], hi[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t wg = (int64_t)vb + (wl[r] < 0 ? 0 : wl[r]);
          lo[r] = voff[wg];
          hi[r] = voff[wg + 1];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t dw = hi[r] - lo[r];
          const bool keep = wl[r] >= 0 && edge_keep(mode, vl, wl[r], dv, dw);
          for (int copy = 0; copy < 2; ++copy) {  // the second copy only in the shadow-vertex mode
            const bool put = keep && copy < edge_copies(mode, wl[r], lim);
            const uint64_t m = __ballot(put);
            if (put) {
              const int64_t idx = pos + __popcll(m & lt);
              if (idx < ecap) {
                eu[idx] = (int32_t)v;
                ew[idx] = (int32_t)((int64_t)vb + wl[r]);
                ework[idx] = (int32_t)(dv + dw);
              }
            }
            pos += __popcll(m);
            if (mode != 3) break;
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------ orientation
// Arc v -> w is kept iff (d(v), v) < (d(w), w) (ids graph-local in a batch,
// where the order only ever compares vertices of one graph).
__device__ __forceinline__ bool orient_keep(int64_t dv, int64_t v, int64_t dw, int64_t w) {
  return dv < dw || (dv == dw && v < w);
}

// Orientation in place.  The arcs of row v are written back to oind at the
// row's own positions voff[v] .. voff[v+1], the kept ones N+(v) from the
// front in ascending order, the others (v's in-neighbours) from the back;
// orow[2v] = voff[v] and orow[2v+1] = voff[v] + |N+(v)|.  One pass, no
// scan, and a row never spills into another whatever the input (a CSR that
// is not symmetric only orients differently).
//
// Waves take tasks of 64 vertices from eight counters, one per XCD (the
// grid has at most one wave per 64 vertices, so a small graph does not
// crowd the counters with idle waves).  A row of at most kShortRow arcs is
// done by its own lane, four
// arcs per step (sparse graphs have rows of a few arcs, which would leave
// most of a wave idle); the longer rows of the 64 are then done one at a
// time by the whole wave, 512 arcs per step with the next step's ids
// already loading, and ballots placing the kept and the other arcs.  The
// hub rows (hundreds of thousands of arcs, one wave each) are what the
// pass waits for last, so their steps are made long and overlapped.
constexpr int64_t kShortRow = 32;

template <typename IT>
__device__ __forceinline__ void orient_row_wave(const int32_t* __restrict__ deg, const IT* __restrict__ ind,
                                                int32_t* __restrict__ oind, int64_t* __restrict__ orow, int64_t v,
                                                int64_t b, int64_t e, int32_t vb, int lane, uint64_t lt) {
  const int64_t dv = e - b;
  const int32_t vl = (int32_t)(v - vb);
  int64_t pf = b, pb = e;  // next front position, one past the next back one
  int32_t wn[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int64_t k = b + lane + 64 * r;
    wn[r] = k < e ? (int32_t)ind[k] : -1;
  }
  for (int64_t k0 = b; k0 < e; k0 += 512) {
    int32_t wl[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) wl[r] = wn[r];
    int32_t dg[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) dg[r] = deg[(int64_t)vb + (wl[r] < 0 ? 0 : wl[r])];
#pragma unroll
    for (int r = 0; r < 8; ++r) {  // the next step's ids, in flight behind the degree reads
      const int64_t k = k0 + 512 + lane + 64 * r;
      wn[r] = k < e ? (int32_t)ind[k] : -1;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const bool valid = wl[r] >= 0;
      const bool keep = valid && orient_keep(dv, vl, dg[r], wl[r]);
      const uint64_t mk = __ballot(keep), mi = __ballot(valid && !keep);
      const int32_t wg = (int32_t)((int64_t)vb + wl[r]);
      if (keep) oind[pf + __popcll(mk & lt)] = wg;
      else if (valid) oind[pb - 1 - __popcll(mi & lt)] = wg;
      pf += __popcll(mk);
      pb -= __popcll(mi);
    }
  }
  if (lane == 0) {
    orow[2 * v] = b;
    orow[2 * v + 1] = pf;
  }
}

// Every vertex's degree into deg (4 bytes: the orientation's random
// reads then touch half the footprint of the 16-byte offset pairs), and
// the rows of more than kHugeRow arcs into huge[0 .. *cnt), one atomic per
// wave; a CSR whose rows sum to more than the list was sized for (nnz /
// kHugeRow + 1 = cap entries) lists only the first cap.
__global__ void k_huge_rows(const int64_t* __restrict__ voff, int64_t nv, int32_t* __restrict__ huge, int64_t cap,
                            unsigned long long* cnt, int32_t* __restrict__ deg) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < nv; base += step) {  // uniform per workgroup
    const int64_t u = base + threadIdx.x;
    const int64_t d = u < nv ? voff[u + 1] - voff[u] : 0;
    if (u < nv) deg[u] = (int32_t)d;  // a simple graph's degree is below nv < 2^31
    const bool h = d > kHugeRow;
    const uint64_t m = __ballot(h);
    if (!m) continue;
    unsigned long long at = 0;
    if (lane == 0) at = atomicAdd(cnt, (unsigned long long)__popcll(m));
    const int64_t i = (int64_t)__shfl(at, 0, 64) + __popcll(m & lt);
    if (h && i < cap) huge[i] = (int32_t)u;
  }
}

// Inclusive max over the wave's lanes by DPP row shifts and broadcasts
// (values >= 0; every lane active).
__device__ __forceinline__ int32_t wave_incl_max_dpp(int32_t v) {
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return v;
}

__device__ __forceinline__ int64_t readlane64(int64_t x, int l) {
  return (int64_t)(uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)(uint64_t)x, l) |
         ((int64_t)__builtin_amdgcn_readlane((int32_t)((uint64_t)x >> 32), l) << 32);
}

// LDS writes of one wave made visible to its own later LDS reads (and kept
// in order by the compiler).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rows of a task of up to kWalkRow arcs are oriented together, as one arc
// range walked in chunks of 64 consecutive arcs (one per lane).
constexpr int64_t kWalkRow = 1024;
// Chunks of a task whose ids and degrees are loaded at once.
constexpr int kWalkAhead = 4;
// The next task (64 vertices from 64 t) for a wave of a workgroup on XCD
// xcd, or -1 when every task is taken.  Task t belongs to counter t & 7
// (otask[x * kOtaskStride]), whose n-th ticket is task 8 n + x; a wave takes
// tickets from its own XCD's counter, then from the others in turn, moving
// on once a counter has run past the last task (k counts the counters
// left behind).  Every lane returns the same value.
__device__ __forceinline__ int64_t next_task(unsigned long long* otask, int64_t T, int xcd, int& k) {
  const int lane = threadIdx.x & 63;
  for (; k < 8; ++k) {
    const int x = (xcd + k) & 7;
    if (x >= T) continue;  // this counter holds no task
    int64_t g = 0;
    if (lane == 0) g = (int64_t)atomicAdd(otask + (int64_t)x * kOtaskStride, 1ull);
    const int64_t t = uniform64(__shfl(g, 0, 64)) * 8 + x;
    if (t < T) return t;
  }
  return -1;
}

// IT: the caller's id type (int32 or int64 CSR); ids are below 2^31.
// First the listed huge rows, one per wave from hctr (they start together
// and run beside everything else), then tasks of 64 vertices from the 8
// counters otask (next_task), whose long-row pass skips the listed rows.
// One counter for every task serialised its atomics, about 9 ns each: 0.5 ms
// of the 1.1 ms orientation of a 4M-vertex geometric graph (65,536 tasks).
// Counters owning contiguous eighths of the vertex range instead of
// interleaved tasks left an R-MAT graph's heavy low-id tasks to one XCD.
//
// WALK (the default): a task's rows of at most kWalkRow arcs are one arc
// range [B, E) of the CSR, walked in chunks of 64 consecutive arcs, lane p
// taking arc c0 + p, so every id load covers whole lines.  The owner lane of
// each row that starts in the chunk marks the row's first position; an
// inclusive max-scan over the marks gives each lane its row; each row's
// kept arcs go to the front of its row and the others to the back, ranked
// by ballots within the row's run of lanes, from front and back cursors
// kept in LDS per row; the last lane of a row's run advances them.  Rows
// longer than kWalkRow (and the listed huge rows) are stepped over and
// oriented one at a time by the whole wave.  Without WALK, each row of at
// most kShortRow arcs is oriented by its own lane.
template <typename IT, bool WALK>
__global__ __launch_bounds__(256) void k_orient_rows(const int64_t* __restrict__ voff, const IT* __restrict__ ind,
                                                     const int32_t* __restrict__ vbase, int64_t nv,
                                                     int32_t* __restrict__ oind, int64_t* __restrict__ orow,
                                                     unsigned long long* otask, const int32_t* __restrict__ huge,
                                                     int64_t cap, const unsigned long long* nhuge,
                                                     unsigned long long* hctr, const int32_t* __restrict__ deg) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  int64_t nh = uniform64((int64_t)*nhuge);
  if (nh > cap) nh = cap;
  // The ticket is taken by all lanes at once (lane 0 adds 1, the others 0;
  // lane 0's result is the ticket), with no branch on the lane: a lane-0
  // branch here was merged by the compiler with the lane-0 store at the end
  // of orient_row_wave, leaving lanes 1-63 looping without a new ticket.
  if (nh > 0) {
    for (;;) {
      const int64_t i = uniform64((int64_t)atomicAdd(hctr, lane == 0 ? 1ull : 0ull));
      if (i >= nh) break;
      const int64_t v = uniform64(huge[i]);
      orient_row_wave(deg, ind, oind, orow, v, voff[v], voff[v + 1], vbase ? vbase[v] : 0, lane, lt);
    }
  }
  __shared__ int4 s_wrow[WALK ? 256 : 1];   // per row of a task: start - B, degree, local id, graph base
  __shared__ int2 s_wcur[WALK ? 256 : 1];   // per row: kept arcs placed, others placed
  __shared__ int32_t s_wst[WALK ? 256 : 1];  // per chunk position: 1 + the lane whose row starts there
  const int64_t T = (nv + 63) >> 6;
  const int xcd = blockIdx.x & 7;
  int k = 0;
  for (;;) {
    const int64_t t = next_task(otask, T, xcd, k);
    if (t < 0) break;
    const int64_t v0 = t * 64;
    const int64_t v = v0 + lane;
    const bool mine = v < nv;
    int64_t b = 0, e = 0;
    int32_t vb = 0;
    if (mine) {
      b = voff[v];
      e = voff[v + 1];
      vb = vbase ? vbase[v] : 0;
    }
    if (WALK) {
      int4* wrow = s_wrow + (threadIdx.x & ~63);
      int2* wcur = s_wcur + (threadIdx.x & ~63);
      int32_t* wst = s_wst + (threadIdx.x & ~63);
      const int64_t d = e - b;
      const bool walk = mine && d <= kWalkRow;
      const bool obst = mine && d > kWalkRow;  // listed (done) or long (done below)
      const int64_t B = uniform64(__shfl(b, 0, 64));  // lane 0 is always mine
      const int last = __popcll(__ballot(mine)) - 1;
      const int64_t E = readlane64(e, last);
      wrow[lane] = make_int4((int32_t)(b - B), (int32_t)d, (int32_t)(v - vb), vb);
      wcur[lane] = make_int2(0, 0);
      int64_t c0 = B;
      int32_t carry = 0;  // 1 + the lane whose row holds arc c0 when it started before c0
      // ids (and, when the task's rows are all of one graph, their degrees)
      // of kWalkAhead chunks loaded together from pc on, so that a chunk's
      // two dependent loads are not waited for one chunk at a time
      const int32_t vb0 = __builtin_amdgcn_readlane(vb, 0);
      const bool one = vb0 == __builtin_amdgcn_readlane(vb, last);
      bool have = false;
      int64_t pc = 0;
      int32_t pw[kWalkAhead], pd[kWalkAhead];
      while (c0 < E) {
        const uint64_t om = __ballot(obst && e > c0);  // the next row not walked
        int64_t os = INT64_MAX, oe = 0;
        if (om) {
          const int l = __builtin_ctzll(om);
          os = readlane64(b, l);
          oe = readlane64(e, l);
        }
        if (os <= c0) {  // step over it
          c0 = oe;
          carry = 0;
          continue;
        }
        int64_t cend = c0 + 64 < E ? c0 + 64 : E;
        if (os < cend) cend = os;
        const int32_t n = (int32_t)(cend - c0);  // 1 .. 64 arcs
        const int32_t c0r = (int32_t)(c0 - B);
        const bool valid = lane < n;
        const int64_t ahead = c0 - pc;
        if (!have || ahead < 0 || ahead >= 64 * kWalkAhead || (ahead & 63)) {
          have = true;
          pc = c0;
          const int64_t lim = os < E ? os : E;  // up to the next row stepped over
#pragma unroll
          for (int j = 0; j < kWalkAhead; ++j) {
            const int64_t q = c0 + 64 * j + lane;
            pw[j] = q < lim ? (int32_t)ind[q] : 0;
          }
#pragma unroll
          for (int j = 0; j < kWalkAhead; ++j) {
            const int64_t q = c0 + 64 * j + lane;
            pd[j] = one && q < lim ? deg[(int64_t)vb0 + pw[j]] : 0;
          }
        }
        const int jj = (int)((c0 - pc) >> 6);  // the same in every lane
        int32_t w = pw[0], dq = pd[0];
#pragma unroll
        for (int j = 1; j < kWalkAhead; ++j)
          if (jj == j) {
            w = pw[j];
            dq = pd[j];
          }
        if (!valid) w = 0;
        wst[lane] = lane == 0 ? carry : 0;
        wave_lds_sync();
        const int32_t rel = (int32_t)(b - B) - c0r;  // this lane's row start within the chunk
        if (walk && d > 0 && rel >= 0 && rel < n) wst[rel] = lane + 1;
        wave_lds_sync();
        const int32_t r1 = wave_incl_max_dpp(wst[lane]);  // 1 + the lane whose row holds arc c0 + lane
        const int r = r1 > 0 ? r1 - 1 : 0;
        const int4 ri = wrow[r];
        const int2 cur = wcur[r];
        const int32_t dw = valid ? (one ? dq : deg[(int64_t)ri.w + w]) : 0;
        const bool keep = valid && orient_keep(ri.y, ri.z, dw, w);
        const bool other = valid && !keep;
        const uint64_t mk = __ballot(keep), mi = __ballot(other);
        const int32_t sg = ri.x - c0r > 0 ? ri.x - c0r : 0;  // first lane of this row's run
        const uint64_t run = lt & ~((1ull << sg) - 1);       // lanes of the run before this one
        const int32_t rk = __popcll(mk & run), rn = __popcll(mi & run);
        const int64_t rowb = B + ri.x;
        if (keep) oind[rowb + cur.x + rk] = ri.w + w;
        if (other) oind[rowb + ri.y - 1 - cur.y - rn] = ri.w + w;
        const int32_t se = ri.x + ri.y - c0r < n ? ri.x + ri.y - c0r : n;  // end of this row's run
        wave_lds_sync();  // every lane has read its row's cursors
        if (valid && lane == se - 1) wcur[r] = make_int2(cur.x + rk + (keep ? 1 : 0), cur.y + rn + (other ? 1 : 0));
        wave_lds_sync();
        carry = __builtin_amdgcn_readlane(r1, n - 1);
        c0 = cend;
      }
      if (walk) {
        orow[2 * v] = b;
        orow[2 * v + 1] = b + wcur[lane].x;
      }
    }
    if (!WALK && mine && e - b <= kShortRow) {
      const int64_t dv = e - b;
      const int32_t vl = (int32_t)(v - vb);
      int64_t pf = b, pb = e;
      for (int64_t k0 = b; k0 < e; k0 += 4) {
        int32_t wl[4];
        int64_t dw[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) wl[r] = k0 + r < e ? (int32_t)ind[k0 + r] : -1;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t wg = (int64_t)vb + (wl[r] < 0 ? 0 : wl[r]);
          dw[r] = deg[wg];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (wl[r] < 0) continue;
          const int32_t wg = (int32_t)((int64_t)vb + wl[r]);
          if (orient_keep(dv, vl, dw[r], wl[r])) oind[pf++] = wg;
          else oind[--pb] = wg;
        }
      }
      orow[2 * v] = b;
      orow[2 * v + 1] = pf;
    }
    // the long rows among the 64, one at a time by the whole wave
    uint64_t lm = __ballot(mine && e - b > (WALK ? kWalkRow : kShortRow) && e - b <= kHugeRow);
    while (lm) {
      const int l = __builtin_ctzll(lm);
      lm &= lm - 1;
      const int64_t bl = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)(uint64_t)b, l) |
                         ((int64_t)__builtin_amdgcn_readlane((int32_t)((uint64_t)b >> 32), l) << 32);
      const int64_t el = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)(uint64_t)e, l) |
                         ((int64_t)__builtin_amdgcn_readlane((int32_t)((uint64_t)e >> 32), l) << 32);
      orient_row_wave(deg, ind, oind, orow, v0 + l, bl, el, __builtin_amdgcn_readlane(vb, l), lane, lt);
    }
  }
}

__global__ void k_edge_finish(const int64_t* __restrict__ eoff, int64_t nv, int64_t ecap, int64_t* __restrict__ dscal) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    int64_t E = eoff[nv];
    dscal[2] = E > ecap ? 1 : 0;  // overflow flag: input was not symmetric
    dscal[0] = E > ecap ? ecap : E;
  }
}
__global__ void k_work_finish(const int64_t* __restrict__ wpre, int64_t* __restrict__ dscal) {
  if (blockIdx.x == 0 && threadIdx.x == 0) dscal[1] = wpre[dscal[0]];
}

// ---------------------------------------------------------------- count
// Level-1 range of logical workgroup lg out of nr: [R0, R1) of [0, W).
__device__ __forceinline__ void level1_range(int64_t W, int nr, int lg, int64_t& R0, int64_t& R1) {
  const int64_t q = W / nr, r = W % nr;
  R0 = q * lg + (lg < r ? lg : r);
  R1 = R0 + q + (lg < r ? 1 : 0);
}

__device__ __forceinline__ int level1_parts(int64_t W, int max_ranges, int min_units) {
  int64_t nr = W / (min_units > 0 ? min_units : 1);
  if (nr < 1) nr = 1;
  if (nr > max_ranges) nr = max_ranges;
  return (int)nr;
}

// Next level-1 range for a workgroup on XCD xcd.  The nr ranges are cut into
// 8 contiguous blocks, one queue per XCD (workgroup i runs on XCD i % 8), so
// each XCD's L2 sees a contiguous part of the edge list; a workgroup whose
// queue is empty takes ranges from the other XCDs' queues.  -1 when all
// ranges are taken.
__device__ __forceinline__ int next_range(unsigned int* q, int nr, int xcd) {
  for (int k = 0; k < 8; ++k) {
    const int x = (xcd + k) & 7;
    const int b0 = (int)(((int64_t)nr * x) >> 3), b1 = (int)(((int64_t)nr * (x + 1)) >> 3);
    if (b1 <= b0) continue;
    if (__hip_atomic_load(&q[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (unsigned)(b1 - b0)) continue;
    const unsigned r = atomicAdd(&q[x], 1u);
    if (r < (unsigned)(b1 - b0)) return b0 + (int)r;
  }
  return -1;
}

#include "tc_mp_kernel.h"

// Count the elements of chunks [c_lo, c_hi) of a CH*4-byte aligned array
// base (chunk c = elements [c*CH, c*CH + CH)), restricted to element positions
// [g0, g1), that occur in the sorted list A[0, lA) held in LDS with
// A[lA] = INT_MAX.  D chunks are in flight per lane: chunk c + D is loaded
// into the register slot chunk c is compared from.
// With XC, a match at A position p also adds 2 to the 16-bit counter p of
// X (LDS, two counters per word, p & 1 selecting the half).
__device__ __forceinline__ void x_credit(int32_t* X, int32_t pos) {
  atomicAdd(&X[pos >> 1], 2 << ((pos & 1) << 4));
}
__device__ __forceinline__ int32_t x_read(const int32_t* X, int32_t pos) {
  return (X[pos >> 1] >> ((pos & 1) << 4)) & 0xffff;
}
template <int CH, int D, bool XC = false>
__device__ __forceinline__ int32_t stream_chunks(const int32_t* __restrict__ base, int64_t c_lo, int64_t c_hi,
                                                 int64_t g0, int64_t g1, const int32_t* A, int32_t lA,
                                                 int32_t* X = nullptr) {
  constexpr int Q = CH / 4;
  if (c_lo >= c_hi) return 0;
  const int4* b4 = reinterpret_cast<const int4*>(base);
  int4 slot[D][Q];
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (c_lo + k < c_hi) {
#pragma unroll
      for (int q = 0; q < Q; ++q) slot[k][q] = b4[(c_lo + k) * Q + q];
    }
  const int64_t gf = c_lo * CH > g0 ? c_lo * CH : g0;
  const int32_t bf = base[gf];
  int32_t lo = 0, hi = lA;  // lower bound of bf in A
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (A[mid] < bf) lo = mid + 1;
    else hi = mid;
  }
  int32_t pos = lo;
  int32_t acur = A[pos];
  int32_t cnt = 0;
  for (int64_t c0 = c_lo; c0 < c_hi; c0 += D) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const int64_t c = c0 + k;
      if (c < c_hi) {
        int32_t b[CH];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          b[4 * q] = slot[k][q].x;
          b[4 * q + 1] = slot[k][q].y;
          b[4 * q + 2] = slot[k][q].z;
          b[4 * q + 3] = slot[k][q].w;
        }
        if (c + D < c_hi) {
#pragma unroll
          for (int q = 0; q < Q; ++q) slot[k][q] = b4[(c + D) * Q + q];
        }
        const int64_t cbase = c * CH;
        if (cbase < g0 || cbase + CH > g1) {
#pragma unroll
          for (int j = 0; j < CH; ++j) {
            const int64_t g = cbase + j;
            if (g < g0 || g >= g1) b[j] = -1;  // ids are >= 0: never matches, never advances
          }
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) {
          const int32_t bv = b[j];
          while (acur < bv) acur = A[++pos];
          if constexpr (XC) {
            if (acur == bv) {
              ++cnt;
              x_credit(X, pos);
            }
          } else {
            cnt += (acur == bv);
          }
        }
      }
    }
  }
  return cnt;
}

// First position p in [lo, hi) with X[p] >= key (hi if there is none),
// found by the whole wave together: each round probes 64 evenly spaced
// positions with one load per lane, and the ballot of X < key (a prefix of
// the lanes, X being sorted) leaves the gap after the last smaller probe.
// The answer stays in [lo, hi] throughout; every load is inside [lo, hi).
// Wave-uniform arguments and result.
__device__ __forceinline__ int64_t wave_lower_bound(const int32_t* __restrict__ X, int64_t lo, int64_t hi, int32_t key,
                                                    int lane) {
  while (hi - lo > 64) {
    const int64_t step = (hi - lo + 63) >> 6;
    const int64_t q = lo + (int64_t)lane * step;
    const bool less = q < hi && X[q] < key;
    const int c = __popcll(__ballot(less));
    const int64_t qc = lo + (int64_t)c * step;  // first probe not smaller (or past hi)
    const int64_t nlo = c ? lo + (int64_t)(c - 1) * step + 1 : lo;
    const int64_t nhi = c ? (qc < hi ? qc : hi) : lo;
    lo = uniform64(nlo);
    hi = uniform64(nhi);
  }
  const int64_t q = lo + lane;
  const bool less = q < hi && X[q] < key;
  return lo + __popcll(__ballot(less));
}

// Edges whose longer list B is at least `ratio` times the shorter list A
// (and at least kSearchMinB long) are counted by k_count_search, looking each
// element of A up in B, instead of streaming B.  Small whole edges (work at
// most small_max, wholly inside one range) stay with the tiles whatever their
// shape; every count kernel applies the same two tests in this order.
constexpr int32_t kSearchMinB = 512;
__device__ __forceinline__ bool edge_search(int32_t lA, int32_t lB, int32_t ratio) {
  return ratio > 0 && lB >= kSearchMinB && (int64_t)lB >= (int64_t)ratio * lA;
}

// The same count as stream_chunks, driven by A instead of B: for each chunk
// the lane walks the A values up to the chunk's largest element (each A
// value is visited once over the lane's run, the next one already loading)
// and tests each against the chunk's CH elements held in registers.  Costs
// about CH compares per A value and no branch per B element, so it wins
// when A is much shorter than B (few A values per chunk).  Needs
// A[lA] = INT_MAX (a value above every chunk element).
template <int CH, int D, bool XC = false>
__device__ __forceinline__ int32_t stream_chunks_a(const int32_t* __restrict__ base, int64_t c_lo, int64_t c_hi,
                                                   int64_t g0, int64_t g1, const int32_t* A, int32_t lA,
                                                   int32_t* X = nullptr) {
  constexpr int Q = CH / 4;
  if (c_lo >= c_hi) return 0;
  const int4* b4 = reinterpret_cast<const int4*>(base);
  int4 slot[D][Q];
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (c_lo + k < c_hi) {
#pragma unroll
      for (int q = 0; q < Q; ++q) slot[k][q] = b4[(c_lo + k) * Q + q];
    }
  const int64_t gf = c_lo * CH > g0 ? c_lo * CH : g0;
  const int32_t bf = base[gf];
  int32_t lo = 0, hi = lA;  // lower bound of bf in A
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (A[mid] < bf) lo = mid + 1;
    else hi = mid;
  }
  int32_t pos = lo;
  int32_t acur = A[pos];
  int32_t cnt = 0;
  for (int64_t c0 = c_lo; c0 < c_hi; c0 += D) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const int64_t c = c0 + k;
      if (c < c_hi) {
        int32_t b[CH];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          b[4 * q] = slot[k][q].x;
          b[4 * q + 1] = slot[k][q].y;
          b[4 * q + 2] = slot[k][q].z;
          b[4 * q + 3] = slot[k][q].w;
        }
        if (c + D < c_hi) {
#pragma unroll
          for (int q = 0; q < Q; ++q) slot[k][q] = b4[(c + D) * Q + q];
        }
        const int64_t cbase = c * CH;
        if (cbase < g0 || cbase + CH > g1) {
#pragma unroll
          for (int j = 0; j < CH; ++j) {
            const int64_t g = cbase + j;
            if (g < g0 || g >= g1) b[j] = -1;  // ids are >= 0: matches nothing
          }
        }
        int32_t bmax = b[0];
#pragma unroll
        for (int j = 1; j < CH; ++j) bmax = b[j] > bmax ? b[j] : bmax;
        // A values in (previous chunk's max, bmax]; acur <= bmax < INT_MAX
        // keeps pos + 1 <= lA, so the prefetch stays on the staged list
        while (acur <= bmax) {
          const int32_t anext = A[pos + 1];
          bool hit = false;
#pragma unroll
          for (int j = 0; j < CH; ++j) hit |= b[j] == acur;
          cnt += hit;
          if constexpr (XC) {
            if (hit) x_credit(X, pos);
          }
          ++pos;
          acur = anext;
        }
      }
    }
  }
  return cnt;
}

#include "tc_bs_kernel.h"
#include "tc_orient_kernel.h"
#include "tc_hash_kernel.h"
#include "tc_search_kernel.h"

// One lane per edge, two-pointer merge from global memory.  Simple and
// unbalanced; kept as an independent cross-check of k_count_mp.
__global__ __launch_bounds__(256) void k_count_simple(CountArgs a) {
  __shared__ int64_t red[4];
  const int64_t E = a.dscal[0];
  int64_t my = 0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t u = a.eu[e], w = a.ew[e];
    const int32_t* A = a.ind + a.voff[u];
    const int32_t* B = a.ind + a.voff[w];
    const int64_t la = a.voff[u + 1] - a.voff[u], lb = a.voff[w + 1] - a.voff[w];
    int64_t i = 0, j = 0;
    int32_t c = 0;
    while (i < la && j < lb) {
      const int32_t x = A[i], y = B[j];
      c += x == y;
      i += x <= y;
      j += y <= x;
    }
    if (c) {
      if (a.out) {
        atomicAdd(&a.out[u], c);
        if (!a.u_only) atomicAdd(&a.out[w], c);
      }
      my += c;
    }
  }
  const int64_t s = block_sum64<256>(my, red);
  if (threadIdx.x == 0 && a.total && s) atomicAdd(a.total, (unsigned long long)(s * a.total_mult));
}

// ---------------------------------------------------------------- finalise
__global__ void k_lcc(const int32_t* __restrict__ off, int32_t nv, const int32_t* __restrict__ out,
                      double* __restrict__ lcc) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t d = (int64_t)off[v + 1] - off[v];
    lcc[v] = d < 2 ? 0.0 : (double)out[v] / (double)(d * (d - 1));
  }
}

// Plain copy, 16 bytes per lane per iteration (counter calibration).
__global__ __launch_bounds__(256) void k_copy_i32(int4* __restrict__ dst, const int4* __restrict__ src, int64_t n4,
                                                  int32_t* __restrict__ dtail, const int32_t* __restrict__ stail,
                                                  int32_t ntail) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t i = g; i < n4; i += (int64_t)gridDim.x * blockDim.x) dst[i] = src[i];
  if (g < ntail) dtail[g] = stail[g];
}

// Number of positions where a and b differ (autotuner output check).
__global__ __launch_bounds__(256) void k_mismatch(const int32_t* __restrict__ a, const int32_t* __restrict__ b,
                                                  int64_t n, int32_t* __restrict__ cnt) {
  int32_t c = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += a[i] != b[i];
  c = wave_sum32(c);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(cnt, c);
}

// Per-graph totals of a batch: one block per graph sums its vertex range.
__global__ __launch_bounds__(256) void k_graph_totals(const int64_t* __restrict__ tab, int32_t G,
                                                      const int32_t* __restrict__ out, int64_t* __restrict__ totals) {
  __shared__ int64_t red[4];
  for (int32_t g = blockIdx.x; g < G; g += gridDim.x) {
    const int64_t b = tab[2 * g], e = tab[2 * g + 2];
    int64_t s = 0;
    for (int64_t v = b + threadIdx.x; v < e; v += blockDim.x) s += out[v];
    s = block_sum64<256>(s, red);
    if (threadIdx.x == 0) totals[g] = s;
  }
}

// Graph of global vertex v of a batch: the last g with v_base[g] <= v.
__device__ __forceinline__ int32_t graph_of_vertex(const int64_t* __restrict__ tab, int32_t G, int64_t v) {
  int32_t lo = 0, hi = G;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (tab[2 * mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Per-graph totals of a batch, added into totals[] (zeroed before; with
// KEEP, a graph whose total is negative, set aside by k_validate_batch, is
// left as it is).  Each wave sums 1024 consecutive vertices, 16 per lane:
// one atomic per wave when they share a graph (the usual case), else one
// per lane and graph.  Even a batch of a few large graphs spreads over the
// whole device.
template <bool KEEP>
__global__ __launch_bounds__(256) void k_graph_sums(const int64_t* __restrict__ tab, int32_t G,
                                                    const int32_t* __restrict__ out, int64_t nv,
                                                    unsigned long long* __restrict__ totals) {
  constexpr int PER = 16;
  const int lane = threadIdx.x & 63;
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t w0 = wave * 64 * PER; w0 < nv; w0 += nwaves * 64 * PER) {
    const int64_t w1 = w0 + 64 * PER < nv ? w0 + 64 * PER : nv;
    const int32_t g0 = graph_of_vertex(tab, G, w0), g1 = graph_of_vertex(tab, G, w1 - 1);
    const int64_t v0 = w0 + (int64_t)lane * PER;
    if (g0 == g1) {
      int64_t sum = 0;
#pragma unroll
      for (int q = 0; q < PER; ++q)
        if (v0 + q < w1) sum += out[v0 + q];
      sum = wave_sum64(sum);
      if (lane == 0 && sum && (!KEEP || (long long)totals[g0] >= 0)) atomicAdd(totals + g0, (unsigned long long)sum);
    } else if (v0 < w1) {
      int32_t g = graph_of_vertex(tab, G, v0);
      int64_t next = g + 1 < G ? tab[2 * g + 2] : nv, sum = 0;
      for (int q = 0; q < PER && v0 + q < w1; ++q) {
        const int64_t v = v0 + q;
        while (v >= next) {  // the lane's run crosses into the next graph
          if (sum && (!KEEP || (long long)totals[g] >= 0)) atomicAdd(totals + g, (unsigned long long)sum);
          sum = 0;
          ++g;
          next = g + 1 < G ? tab[2 * g + 2] : nv;
        }
        sum += out[v];
      }
      if (sum && (!KEEP || (long long)totals[g] >= 0)) atomicAdd(totals + g, (unsigned long long)sum);
    }
  }
}

// The simple-graph contract of every graph of a packed batch (layout of
// include/tc_host.h: graph g's local offsets at off_cat[v_base + g], its
// local ids at ind_cat[arc_base]), one thread per arc: a graph with an id
// outside [0, n_g), a row not strictly increasing, a self-loop or an arc
// without its reverse gets totals[g] = -1 (cleared to 0 before).  Every
// graph's offsets must already run from 0 to nnz_g without decreasing (the
// caller's check); ids are tested against n_g before they index anything.
// The arc's graph and row come from binary searches over the table and the
// graph's offsets, which neighbouring threads share in cache; no row is
// walked by one wave, so a hub row costs no more than its arcs.
__global__ __launch_bounds__(256) void k_validate_batch(const int64_t* __restrict__ tab, int32_t G,
                                                        const int32_t* __restrict__ off_cat,
                                                        const int32_t* __restrict__ ind_cat, int64_t nnz,
                                                        int64_t* __restrict__ totals) {
  for (int64_t K = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; K < nnz; K += (int64_t)gridDim.x * blockDim.x) {
    int32_t lo = 0, hi = G;  // last g with arc_base[g] <= K: the graph whose arcs hold K
    while (hi - lo > 1) {
      const int32_t mid = (lo + hi) >> 1;
      if (tab[2 * mid + 1] <= K) lo = mid;
      else hi = mid;
    }
    const int64_t vb = tab[2 * lo], ab = tab[2 * lo + 1];
    const int64_t ng = tab[2 * lo + 2] - vb;
    const int32_t* off = off_cat + vb + lo;  // graph-local offsets, ng + 1 of them
    const int32_t* ind = ind_cat + ab;       // graph-local ids, tab[2 lo + 3] - ab of them
    const int64_t k = K - ab;                // 0 <= k < the graph's arc count
    int64_t x = 0, xh = ng;                  // row of k: the last x with off[x] <= k
    while (xh - x > 1) {
      const int64_t mid = (x + xh) >> 1;
      if (off[mid] <= k) x = mid;
      else xh = mid;
    }
    const int32_t w = ind[k];
    bool bad = w < 0 || w >= ng || w == x || (k > off[x] && ind[k - 1] >= w);
    if (!bad) {  // x in row w
      int64_t l = off[w], h = off[w + 1];
      const int64_t hend = h;
      while (l < h) {
        const int64_t mid = (l + h) >> 1;
        if (ind[mid] < x) l = mid + 1;
        else h = mid;
      }
      bad = l >= hend || ind[l] != x;
    }
    if (bad) totals[lo] = -1;
  }
}

// Sets aside the graphs k_validate_batch marked: every id of every row of
// such a graph becomes the row's own local id, so its orientation keeps no
// arc (N+ is empty everywhere) and the count reads none of its rows.  The
// offsets are left as they are, so every row still ends where the next
// begins.  One workgroup per graph at a time.
__global__ __launch_bounds__(256) void k_drop_marked(const int64_t* __restrict__ tab, int32_t G,
                                                     const int32_t* __restrict__ off_cat, int32_t* __restrict__ ind_cat,
                                                     const int64_t* __restrict__ totals) {
  for (int32_t g = blockIdx.x; g < G; g += gridDim.x) {
    if (totals[g] >= 0) continue;  // uniform per block
    const int64_t vb = tab[2 * g], ng = tab[2 * g + 2] - vb;
    const int32_t* off = off_cat + vb + g;
    int32_t* ind = ind_cat + tab[2 * g + 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int64_t x = wv; x < ng; x += nw)
      for (int64_t k = off[x] + lane; k < off[x + 1]; k += 64) ind[k] = (int32_t)x;
  }
}

// Simple-graph contract check, one wave per vertex; flag bits:
// 1 range/offsets, 2 unsorted or duplicate, 4 self-loop, 8 missing reverse.
__global__ __launch_bounds__(256) void k_validate(const int32_t* __restrict__ off, const int32_t* __restrict__ ind,
                                                  int32_t nv, int64_t nnz, int32_t* __restrict__ flag) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  if (wave == 0 && lane == 0 && (off[0] != 0 || off[nv] != nnz)) atomicOr(flag, 1);
  for (int64_t v = wave; v < nv; v += nwaves) {
    const int64_t b = off[v], e = off[v + 1];
    if (e < b || b < 0 || e > nnz) {
      if (lane == 0) atomicOr(flag, 1);
      continue;
    }
    int32_t f = 0;
    for (int64_t k = b + lane; k < e; k += 64) {
      const int32_t w = ind[k];
      if (w < 0 || w >= nv) {
        f |= 1;
        continue;
      }
      if (k > b && ind[k - 1] >= w) f |= 2;
      if (w == v) f |= 4;
      int64_t lo = off[w], hi = off[w + 1];
      if (lo < 0 || hi > nnz || hi < lo) {
        f |= 1;
        continue;
      }
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (ind[mid] < v) lo = mid + 1;
        else hi = mid;
      }
      if (lo >= off[w + 1] || ind[lo] != v) f |= 8;
    }
    if (f) atomicOr(flag, f);
  }
}

unsigned grid_for(int64_t n, int bs, int64_t cap = 65536) {
  int64_t g = (n + bs - 1) / bs;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// ---------------------------------------------------------------- vertex list
// vmark[v] = 1 for every listed v in [0, nv) (entries outside are skipped).
__global__ void k_mark_list(const int32_t* __restrict__ list, int64_t nlist, int64_t nv, int32_t* __restrict__ vmark) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nlist; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = list[i];
    if (v >= 0 && v < nv) vmark[v] = 1;
  }
}

// out[i] = vcnt[list[i]] (-1 for an entry outside [0, nv)), and the total
// of the listed counts added to *total.
__global__ __launch_bounds__(256) void k_gather_list(const int32_t* __restrict__ list, int64_t nlist, int64_t nv,
                                                     const int32_t* __restrict__ vcnt, int32_t* __restrict__ out,
                                                     unsigned long long* total) {
  __shared__ int64_t red[4];
  int64_t my = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nlist; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = list[i];
    const int32_t c = (v >= 0 && v < nv) ? vcnt[v] : -1;
    out[i] = c;
    my += c > 0 ? c : 0;
  }
  const int64_t s = block_sum64<256>(my, red);
  if (threadIdx.x == 0 && total && s) atomicAdd(total, (unsigned long long)s);
}

// Compiled count kernels: algo 0 = streamed (k_count_bs) with every long
// edge on one wave (A in windows), launched as that wave kernel fn then a
// small-tile kernel fn2; 3 = streamed with the longest edges left to a
// workgroup/ring phase in fn2; 2 = Merge Path rings only (k_count_mp).
// Algos 0 and 3 also launch k_count_search.
// {algo, BS, K, P, M, chunk, depth, wave A cap}.
struct VariantDef {
  int algo, bs, k, p, m, ch, sd, awt;
  void (*fn)(CountArgs);
  void (*fn2)(CountArgs);
};
#define TC_BW(BS, CH, P, M, K, SD, AWT)                                                        \
  {0, BS, K, P, M, CH, SD, AWT, k_count_bs<BS, CH, P, M, K, SD, 9, AWT>,                    \
   k_count_bs<BS, CH, P, M, K, SD, 4, AWT>}
#define TC_BS(BS, CH, P, M, K, SD, AWT)                                                        \
  {3, BS, K, P, M, CH, SD, AWT, k_count_bs<BS, CH, P, M, K, SD, 1, AWT>,                    \
   k_count_bs<BS, CH, P, M, K, SD, 6, AWT>}
#define TC_MP(BS, K, P, M) {2, BS, K, P, M, 0, 0, 0, k_count_mp<BS, K, P, M>, nullptr}
const VariantDef kVariants[] = {
    TC_BW(256, 16, 4096, 128, 8, 2, 2039),
    TC_BW(256, 16, 4096, 128, 8, 4, 2039),
    TC_BW(128, 16, 4096, 64, 16, 2, 4087),
    TC_BW(512, 16, 4096, 128, 4, 2, 1019),
    TC_BS(256, 16, 4096, 128, 8, 2, 2039),
    TC_BS(256, 16, 4096, 128, 8, 4, 2039),
    TC_BS(512, 16, 4096, 128, 4, 2, 1019),
    TC_BS(128, 16, 4096, 64, 16, 2, 4087),
    TC_MP(256, 8, 4096, 128),
    TC_MP(256, 16, 8192, 128),
    TC_MP(512, 8, 8192, 128),
    TC_MP(128, 16, 4096, 64),
    TC_MP(256, 4, 2048, 64),
    TC_MP(512, 4, 4096, 128),
};
#undef TC_BS
#undef TC_BW
#undef TC_MP
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

}  // namespace

int num_variants() { return kNumVariants; }
Variant variant(int i) {
  const VariantDef& d = kVariants[i];
  return Variant{d.bs, d.k, d.p, d.algo, d.ch, d.sd};
}

size_t scan_tmp_elems(int64_t n_cap) {
  size_t total = 0;
  int64_t n = n_cap;
  do {
    int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb < 1) nb = 1;
    total += 2 * (size_t)nb + 1;
    n = nb;
  } while (n > 1);
  return total + 8;
}

hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_i64(int64_t* p, int64_t n, int64_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t launch_voff_from_off32(const int32_t* off, int32_t nv, int64_t* voff, hipStream_t s) {
  hipLaunchKernelGGL(k_voff_from_off32, dim3(grid_for((int64_t)nv + 1, 256, 16384)), dim3(256), 0, s, off,
                     (int64_t)nv + 1, voff);
  return hipGetLastError();
}
hipError_t launch_voff_batch(const int64_t* tab, int32_t G, const int32_t* off_cat, int64_t nv_total, int64_t* voff,
                             int32_t* vbase, hipStream_t s) {
  hipLaunchKernelGGL(k_voff_batch, dim3(grid_for(nv_total + 1, 256, 16384)), dim3(256), 0, s, tab, G, off_cat,
                     nv_total, voff, vbase);
  return hipGetLastError();
}

hipError_t launch_build_edges(const Prep& p, int64_t nv, const int64_t* voff, const int32_t* ind, int mode, int cus,
                              hipStream_t s, int64_t lim) {
  const unsigned wg = grid_for(nv * 64, 256, (int64_t)cus * 16);
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(p.dscal + 4);
  hipError_t e = launch_fill_i64(p.dscal + 4, 2, 0, s);
  if (e != hipSuccess) return e;
  if (nv > 0)
    hipLaunchKernelGGL(k_edge_count, dim3(wg), dim3(256), 0, s, voff, ind, (const int32_t*)p.vbase, nv, mode,
                       (const int32_t*)p.vmark, p.ecnt, ctr, lim);
  e = scan_excl<int32_t>(p.ecnt, nullptr, nv, p.eoff, p.scan_tmp, s);
  if (e != hipSuccess) return e;
  if (nv > 0)
    hipLaunchKernelGGL(k_edge_fill, dim3(wg), dim3(256), 0, s, voff, ind, (const int32_t*)p.vbase, nv, mode,
                       (const int32_t*)p.vmark, (const int64_t*)p.eoff, p.ecap, p.eu, p.ew, p.ework, ctr + 1, lim);
  hipLaunchKernelGGL(k_edge_finish, dim3(1), dim3(64), 0, s, (const int64_t*)p.eoff, nv, p.ecap, p.dscal);
  e = scan_excl<int32_t>(p.ework, p.dscal, p.ecap, p.wpre, p.scan_tmp, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_work_finish, dim3(1), dim3(64), 0, s, (const int64_t*)p.wpre, p.dscal);
  return hipGetLastError();
}

// prow[m] = the row r with toff[r] <= m * P < toff[r + 1], for every piece
// m of P partner arcs.
__global__ void k_piece_rows(const int64_t* __restrict__ toff, int64_t nv, int64_t P, int32_t* __restrict__ prow) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nv; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t lo = toff[r], hi = toff[r + 1];
    for (int64_t m = (lo + P - 1) / P; m * P < hi; ++m) prow[m] = (int32_t)r;
  }
}

// Everything a single graph needs before its orientation, in one pass over
// the rows (in place of the output and total fills, k_voff_from_off32,
// k_huge_rows and k_piece_rows): the widened offsets (int32 input; an
// int64 CSR's offsets are read in place, voff == nullptr), every vertex's
// degree, the rows of more than kHugeRow arcs listed (hcnt zeroed before),
// the row of the first arc position of each piece of P, and the output
// cleared (out[v] = 0, *total = 0).
template <typename OffT, typename OT>
__global__ __launch_bounds__(256) void k_prep_single(const OffT* __restrict__ off, int64_t nv,
                                                     int64_t* __restrict__ voff, int32_t* __restrict__ deg,
                                                     int32_t* __restrict__ huge, int64_t cap,
                                                     unsigned long long* hcnt, int64_t P, int32_t* __restrict__ prow,
                                                     OT* __restrict__ out, int64_t* __restrict__ total) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (total) *total = 0;
    if (voff) voff[nv] = (int64_t)off[nv];
  }
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < nv; base += step) {  // uniform per workgroup
    const int64_t u = base + threadIdx.x;
    int64_t b = 0, e = 0;
    if (u < nv) {
      b = (int64_t)off[u];
      e = (int64_t)off[u + 1];
      if (voff) voff[u] = b;
      deg[u] = (int32_t)(e - b);  // a simple graph's degree is below nv < 2^31
      if (out) out[u] = 0;
      for (int64_t m = (b + P - 1) / P; m * P < e; ++m) prow[m] = (int32_t)u;
    }
    const bool h = e - b > kHugeRow;
    const uint64_t m = __ballot(h);
    if (!m) continue;
    unsigned long long at = 0;
    if (lane == 0) at = atomicAdd(hcnt, (unsigned long long)__popcll(m));
    const int64_t i = (int64_t)__shfl(at, 0, 64) + __popcll(m & lt);
    if (h && i < cap) huge[i] = (int32_t)u;
  }
}

// Next-chunk prefetch and waves per SIMD of the oriented count kernel (its
// workgroup size is a launch argument, OrientLaunch).
constexpr bool kOrientPF = true;
constexpr int kOrientWPE = 8;
template <typename OT>
using OrientFn = void (*)(const int64_t*, const int64_t*, const int32_t*, const int32_t*, int64_t, OT*,
                          unsigned long long*, unsigned long long*, unsigned long long*, int, int32_t, int64_t);
template <int AW, bool PV, typename OT, bool ONE>
OrientFn<OT> orient_fn_w(int groups) {
  switch (groups) {
    case 1: return k_count_orient<AW, 1, kOrientPF, kOrientWPE, PV, OT, ONE>;
    case 2: return k_count_orient<AW, 2, kOrientPF, kOrientWPE, PV, OT, ONE>;
    case 8: return k_count_orient<AW, 8, kOrientPF, kOrientWPE, PV, OT, ONE>;
    case 16: return k_count_orient<AW, 16, kOrientPF, kOrientWPE, PV, OT, ONE>;
    case 32: return k_count_orient<AW, 32, kOrientPF, kOrientWPE, PV, OT, ONE>;
    default: return k_count_orient<AW, 4, kOrientPF, kOrientWPE, PV, OT, ONE>;
  }
}
// partners per wave (groups), window, per-vertex output type (or the total
// only), one piece per wave -> kernel.  The windows other than the default
// are compiled for int32 counts only (the tuned paths); nullptr otherwise.
template <bool PV, typename OT, bool ONE = false>
OrientFn<OT> orient_fn(int groups, int window) {
  if (window == kOrientWindow) return orient_fn_w<kOrientWindow, PV, OT, ONE>(groups);
  if (!PV || !std::is_same<OT, int32_t>::value) return nullptr;
  if (window == kOrientWindowSmall) return orient_fn_w<kOrientWindowSmall, PV, OT, ONE>(groups);
  if (window == kOrientWindowLarge) return orient_fn_w<kOrientWindowLarge, PV, OT, ONE>(groups);
  return nullptr;
}
// Table slots per wave and window of the hash-lookup kernel: 24 TS bytes of
// LDS per workgroup (keys and 16-bit counters of four waves); a window of AW
// ids fills a table of 2 * roundup32(2 AW / 3) <= TS slots to at most 3/4.
constexpr int kHashTS = 832;
constexpr int kHashAW = 624;
constexpr int kHashWPE = 8;
// Rows with more ids in N+ than this go to the table kernel's long-row
// phase (at most kHashAW, the per-wave window).
constexpr int kHashLong = kHashAW;
template <typename OT>
using HashFn = void (*)(const int64_t*, const int64_t*, const int32_t*, const int32_t*, int64_t, OT*,
                        unsigned long long*, unsigned long long*, unsigned long long*, int, int32_t, int, int64_t,
                        const int32_t*,
                        const unsigned long long*, int64_t, unsigned long long*);
// WIN: rows longer than the window are walked in windows (without it the
// kernel counts them first, in its long-row phase)
template <bool PV, typename OT, bool ONE = false, bool WIN = true>
HashFn<OT> hash_fn(int groups) {
  switch (groups) {
    case 1: return k_count_hash<kHashTS, kHashAW, 1, kHashWPE, PV, OT, ONE, WIN>;
    case 2: return k_count_hash<kHashTS, kHashAW, 2, kHashWPE, PV, OT, ONE, WIN>;
    case 8: return k_count_hash<kHashTS, kHashAW, 8, kHashWPE, PV, OT, ONE, WIN>;
    case 16: return k_count_hash<kHashTS, kHashAW, 16, kHashWPE, PV, OT, ONE, WIN>;
    case 32: return k_count_hash<kHashTS, kHashAW, 32, kHashWPE, PV, OT, ONE, WIN>;
    default: return k_count_hash<kHashTS, kHashAW, 4, kHashWPE, PV, OT, ONE, WIN>;
  }
}
// Table slots of the long-row kernel: one table over the workgroup's LDS
// (6 TSW bytes), rows of up to 3 TSW / 4 = 2496 ids in one window.
constexpr int kLongTS = 3328;
template <typename OT>
using LongFn = void (*)(const int64_t*, const int64_t*, const int32_t*, const int32_t*, const unsigned long long*,
                        int64_t, OT*, unsigned long long*, const unsigned long long*, unsigned long long*, int);
// partners per wave at once: 1, 2 or 4 (groups above 4 take 4: long rows'
// partner lists are long)
template <bool PV, typename OT>
LongFn<OT> long_fn(int groups) {
  if (groups <= 1) return k_count_long<kLongTS, 1, PV, OT>;
  if (groups == 2) return k_count_long<kLongTS, 2, PV, OT>;
  return k_count_long<kLongTS, 4, PV, OT>;
}
int resident_long_wgs(int groups, int kind, int cus) {
  static std::mutex mu;
  static int occ[3][3] = {};
  const int gi = groups <= 1 ? 0 : groups == 2 ? 1 : 2;
  const int ki = kind == 1 || kind == 2 ? kind : 0;
  std::lock_guard<std::mutex> lk(mu);
  if (!occ[ki][gi]) {
    const void* f = ki == 1   ? reinterpret_cast<const void*>(long_fn<true, int32_t>(groups))
                    : ki == 2 ? reinterpret_cast<const void*>(long_fn<true, int64_t>(groups))
                              : reinterpret_cast<const void*>(long_fn<false, int32_t>(groups));
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 256, 0) != hipSuccess || n < 1) n = 1;
    occ[ki][gi] = n;
  }
  return occ[ki][gi] * cus;
}
// kind: 0 = the total only, 1 = int32 counts, 2 = int64 counts; lookup 0 =
// k_count_orient, else k_count_hash
const void* orient_kernel(int groups, int kind, int lookup, int window) {
  if (lookup) {
    if (kind == 1) return reinterpret_cast<const void*>(hash_fn<true, int32_t>(groups));
    if (kind == 2) return reinterpret_cast<const void*>(hash_fn<true, int64_t>(groups));
    return reinterpret_cast<const void*>(hash_fn<false, int32_t>(groups));
  }
  if (kind == 1) return reinterpret_cast<const void*>(orient_fn<true, int32_t>(groups, window));
  if (kind == 2) return reinterpret_cast<const void*>(orient_fn<true, int64_t>(groups, window));
  return reinterpret_cast<const void*>(orient_fn<false, int32_t>(groups, window));
}
bool orient_shape_ok(const OrientLaunch& o, int kind) {
  return (o.wg == 128 || o.wg == 256) && (o.lookup || orient_kernel(o.groups, kind, 0, o.window)) &&
         o.piece >= kOrientPieceMin && 2 * o.piece < 65536;
}
int resident_orient_wgs(const OrientLaunch& o, int kind, int cus) {
  static std::mutex mu;
  static std::map<std::array<int, 5>, int> occ;
  const int gi = o.groups == 1 || o.groups == 2 || o.groups == 8 || o.groups == 16 || o.groups == 32 ? o.groups : 4;
  const int ki = kind == 1 || kind == 2 ? kind : 0;
  const int li = o.lookup ? 1 : 0;
  // the table kernel has a fixed shape
  const std::array<int, 5> key{li, ki, gi, li ? 256 : o.wg, li ? 0 : o.window};
  std::lock_guard<std::mutex> lk(mu);
  auto it = occ.find(key);
  if (it == occ.end()) {
    const size_t lds = li ? 0 : orient_lds_bytes(o.wg / 64, o.window, ki != 0);
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, orient_kernel(gi, ki, li, o.window), key[3], lds) !=
            hipSuccess ||
        n < 1)
      n = 1;
    it = occ.emplace(key, n).first;
  }
  return it->second * cus;
}

template <bool PV, typename OT>
hipError_t launch_count_orient_t(const Prep& p, int64_t nv, int64_t nnz, OT* out, int64_t* total,
                                 const OrientLaunch& o, int kind, int cus, hipStream_t s, bool prepped) {
  if (!orient_shape_ok(o, kind) || nnz < 0) return hipErrorInvalidValue;
  unsigned long long* ctr = reinterpret_cast<unsigned long long*>(p.dscal + 4);
  unsigned long long* pq = p.otask + 8 * kOtaskStride;  // the piece counters
  hipError_t e = prepped ? hipSuccess : launch_fill_i64(reinterpret_cast<int64_t*>(pq), 8 * kOtaskStride, 0, s);
  if (e != hipSuccess) return e;
  if (nv <= 0) return hipSuccess;
  if (!p.orow || !p.oind || !p.prow) return hipErrorInvalidValue;
  const int nwg = resident_orient_wgs(o, kind, cus);
  const int wg = o.lookup ? 256 : o.wg;
  const int64_t piece = orient_piece(nnz, o.piece, (wg / 64) * (int64_t)nwg);
  if (!prepped)
    hipLaunchKernelGGL(k_piece_rows, dim3(grid_for(nv, 256, 4096)), dim3(256), 0, s, (const int64_t*)p.voff, nv,
                       piece, p.prow);
  // a wave for every piece: the kernel that takes one piece per wave
  const bool one = (wg / 64) * (int64_t)nwg * piece >= nnz;
  // rows longer than a wave's window are listed first; the table kernel
  // counts them in its first phase, the walked-array kernel leaves them to
  // k_count_long
  const int64_t aw = o.lookup ? kHashLong : o.window;
  const int64_t lmax = o.lng ? aw : INT64_MAX;
  unsigned long long* lq = reinterpret_cast<unsigned long long*>(p.qctr);  // [0] rows listed, [1] rows taken
  const int64_t lcap = nnz / aw + 1;
  if (o.lng) {
    if (!p.lrow) return hipErrorInvalidValue;
    if ((e = launch_fill_i64(reinterpret_cast<int64_t*>(p.qctr), 2, 0, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_long_rows, dim3(grid_for(nv, 256, 4096)), dim3(256), 0, s, (const int64_t*)p.orow, nv, aw,
                       p.lrow, lcap, lq);
  }
  if (o.lookup) {
    const HashFn<OT> fn = o.lng ? (one ? hash_fn<PV, OT, true, false>(o.groups) : hash_fn<PV, OT, false, false>(o.groups))
                                : (one ? hash_fn<PV, OT, true>(o.groups) : hash_fn<PV, OT>(o.groups));
    hipLaunchKernelGGL(fn, dim3(nwg), dim3(256), 0, s, (const int64_t*)p.voff, (const int64_t*)p.orow,
                       (const int32_t*)p.oind, (const int32_t*)p.prow, nv, out,
                       reinterpret_cast<unsigned long long*>(total), ctr, pq, o.side, (int32_t)piece,
                       o.lookup == 2 ? 1 : 0, lmax, (const int32_t*)p.lrow, (const unsigned long long*)lq, lcap,
                       lq + 1);
    return hipGetLastError();  // the long rows were counted in its first phase
  } else {
    const OrientFn<OT> fn = one ? orient_fn<PV, OT, true>(o.groups, o.window) : orient_fn<PV, OT>(o.groups, o.window);
    hipLaunchKernelGGL(fn, dim3(nwg), dim3(o.wg), orient_lds_bytes(o.wg / 64, o.window, PV), s,
                       (const int64_t*)p.voff, (const int64_t*)p.orow, (const int32_t*)p.oind,
                       (const int32_t*)p.prow, nv, out, reinterpret_cast<unsigned long long*>(total), ctr, pq,
                       o.side, (int32_t)piece, lmax);
  }
  if (o.lng) {
    const LongFn<OT> lf = long_fn<PV, OT>(o.groups);
    hipLaunchKernelGGL(lf, dim3(resident_long_wgs(o.groups, kind, cus)), dim3(256), 0, s, (const int64_t*)p.voff,
                       (const int64_t*)p.orow, (const int32_t*)p.oind, (const int32_t*)p.lrow,
                       (const unsigned long long*)lq, lcap, out, reinterpret_cast<unsigned long long*>(total),
                       (const unsigned long long*)ctr, lq + 1, o.side);
  }
  return hipGetLastError();
}
hipError_t launch_count_orient(const Prep& p, int64_t nv, int64_t nnz, int32_t* out, int64_t* total,
                               const OrientLaunch& o, int cus, hipStream_t s, bool prepped) {
  return out ? launch_count_orient_t<true>(p, nv, nnz, out, total, o, 1, cus, s, prepped)
             : launch_count_orient_t<false>(p, nv, nnz, out, total, o, 0, cus, s, prepped);
}
hipError_t launch_count_orient(const Prep& p, int64_t nv, int64_t nnz, int64_t* out, int64_t* total,
                               const OrientLaunch& o, int cus, hipStream_t s, bool prepped) {
  if (!out) return launch_count_orient_t<false>(p, nv, nnz, (int32_t*)nullptr, total, o, 0, cus, s, prepped);
  return launch_count_orient_t<true>(p, nv, nnz, out, total, o, 2, cus, s, prepped);
}

int64_t orient_piece_of(const OrientLaunch& o, int kind, int cus, int64_t nnz) {
  const int nwg = resident_orient_wgs(o, kind, cus);
  const int wg = o.lookup ? 256 : o.wg;
  return orient_piece(nnz, o.piece, (wg / 64) * (int64_t)nwg);
}

template <typename OffT, typename OT>
hipError_t launch_prep_single_t(const Prep& p, int64_t nv, int64_t nnz, const OffT* off, int64_t piece, OT* out,
                                int64_t* total, hipStream_t s) {
  if (nv <= 0) {  // nothing to orient or count: the total alone
    if (!total) return hipSuccess;
    return launch_fill_i64(total, 1, 0, s);
  }
  if (!p.huge || !p.odeg || !p.prow || piece < 1) return hipErrorInvalidValue;
  int64_t* voff = sizeof(OffT) == sizeof(int64_t) ? nullptr : p.voff;
  hipLaunchKernelGGL((k_prep_single<OffT, OT>), dim3(grid_for(nv, 256, 4096)), dim3(256), 0, s, off, nv, voff,
                     p.odeg, p.huge, nnz / kHugeRow + 1, reinterpret_cast<unsigned long long*>(p.dscal + 6), piece,
                     p.prow, out, total);
  return hipGetLastError();
}
hipError_t launch_prep_single(const Prep& p, int64_t nv, int64_t nnz, const int32_t* off, int64_t piece,
                              int32_t* out, int64_t* total, hipStream_t s) {
  return launch_prep_single_t(p, nv, nnz, off, piece, out, total, s);
}
hipError_t launch_prep_single(const Prep& p, int64_t nv, int64_t nnz, const int64_t* off, int64_t piece,
                              int64_t* out, int64_t* total, hipStream_t s) {
  return launch_prep_single_t(p, nv, nnz, off, piece, out, total, s);
}

hipError_t launch_mark_list(const int32_t* list, int64_t nlist, int64_t nv, int32_t* vmark, hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mark_list, dim3(grid_for(nlist, 256, 16384)), dim3(256), 0, s, list, nlist, nv, vmark);
  return hipGetLastError();
}
hipError_t launch_gather_list(const int32_t* list, int64_t nlist, int64_t nv, const int32_t* vcnt, int32_t* out,
                              int64_t* total, hipStream_t s) {
  if (nlist <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_list, dim3(grid_for(nlist, 256, 4096)), dim3(256), 0, s, list, nlist, nv, vcnt, out,
                     reinterpret_cast<unsigned long long*>(total));
  return hipGetLastError();
}

template <typename IT>
hipError_t launch_orient_t(const Prep& p, int64_t nv, int64_t nnz, const IT* ind, int cus, hipStream_t s,
                           bool prepped) {
  if (!p.orow || !p.oind) return hipErrorInvalidValue;
  // task counter, huge rows listed and taken
  hipError_t e = prepped ? hipSuccess : launch_fill_i64(p.dscal + 5, kScalars - 5, 0, s);
  if (e != hipSuccess || nv <= 0) return e;
  if (!p.huge || !p.odeg) return hipErrorInvalidValue;
  const int64_t cap = nnz / kHugeRow + 1;
  if (!prepped)
    hipLaunchKernelGGL(k_huge_rows, dim3(grid_for(nv, 256, 4096)), dim3(256), 0, s, (const int64_t*)p.voff, nv,
                       p.huge, cap, reinterpret_cast<unsigned long long*>(p.dscal + 6), p.odeg);
  // at most one wave per 64 vertices: ceil(nv / 256) workgroups of 4 waves, capped
  hipLaunchKernelGGL((k_orient_rows<IT, true>), dim3(grid_for(nv, 256, (int64_t)cus * 16)), dim3(256), 0, s,
                     (const int64_t*)p.voff, ind, (const int32_t*)p.vbase, nv, p.oind, p.orow, p.otask,
                     (const int32_t*)p.huge, cap,
                     reinterpret_cast<const unsigned long long*>(p.dscal + 6),
                     reinterpret_cast<unsigned long long*>(p.dscal + 7), (const int32_t*)p.odeg);
  return hipGetLastError();
}
hipError_t launch_orient(const Prep& p, int64_t nv, int64_t nnz, const int32_t* ind, int cus, hipStream_t s,
                         bool prepped) {
  return launch_orient_t(p, nv, nnz, ind, cus, s, prepped);
}
hipError_t launch_orient(const Prep& p, int64_t nv, int64_t nnz, const int64_t* ind, int cus, hipStream_t s,
                         bool prepped) {
  return launch_orient_t(p, nv, nnz, ind, cus, s, prepped);
}

int occupancy(void (*fn)(CountArgs), int bs) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(fn), bs, 0) != hipSuccess || n < 1)
    n = 1;
  return n;
}

// Resident workgroups of a variant's first (or second) kernel.
// Thread-safe: the per-variant figures are computed once under a lock.
int resident_wgs2(int vi, int cus, int which) {
  static std::mutex mu;
  static int occ[kNumVariants][2] = {};
  if (vi < 0 || vi >= kNumVariants || which < 0 || which > 1) return 0;
  std::lock_guard<std::mutex> lk(mu);
  if (!occ[vi][0]) {
    occ[vi][1] = kVariants[vi].fn2 ? occupancy(kVariants[vi].fn2, kVariants[vi].bs) : 0;
    occ[vi][0] = occupancy(kVariants[vi].fn, kVariants[vi].bs);
    if (!occ[vi][1]) occ[vi][1] = occ[vi][0];
  }
  return occ[vi][which] * cus;
}

int resident_wgs(int vi, int cus) { return resident_wgs2(vi, cus, 0); }

constexpr int kSearchBS = 256;
int resident_search_wgs(int cus) {
  static std::mutex mu;
  static int occ = 0;
  std::lock_guard<std::mutex> lk(mu);
  if (!occ) occ = occupancy(k_count_search<kSearchBS>, kSearchBS);
  return occ * cus;
}

hipError_t launch_count(const CountArgs& a, int algo, int vi, int nwg, int nwg2, int nwg3, hipStream_t s) {
  if (algo == 1) {
    CountArgs b = a;
    b.search_ratio = 0;
    hipLaunchKernelGGL(k_count_simple, dim3(nwg), dim3(256), 0, s, b);
    return hipGetLastError();
  }
  if (vi < 0 || vi >= kNumVariants) return hipErrorInvalidValue;
  const VariantDef& v = kVariants[vi];
  hipError_t e = launch_fill_i32(reinterpret_cast<int32_t*>(a.qctr), 24, 0, s);
  if (e != hipSuccess) return e;
  CountArgs b = a;
  b.wave_acap = v.awt;
  if (!v.fn2) b.search_ratio = 0;  // the ring kernel counts every edge itself
  hipLaunchKernelGGL(v.fn, dim3(nwg), dim3(v.bs), 0, s, b);
  if (v.fn2) {
    // same ranges (a.max_ranges), its own range queues
    b.qctr = a.qctr + 8;
    hipLaunchKernelGGL(v.fn2, dim3(nwg2 > 0 ? nwg2 : 1), dim3(v.bs), 0, s, b);
    if (b.search_ratio > 0) {
      b.qctr = a.qctr + 16;
      hipLaunchKernelGGL(k_count_search<kSearchBS>, dim3(nwg3 > 0 ? nwg3 : 1), dim3(kSearchBS), 0, s, b);
    }
  }
  return hipGetLastError();
}

hipError_t launch_lcc(const int32_t* off, int32_t nv, const int32_t* out, double* lcc, hipStream_t s) {
  if (nv <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lcc, dim3(grid_for(nv, 256, 16384)), dim3(256), 0, s, off, nv, out, lcc);
  return hipGetLastError();
}

hipError_t launch_validate(const int32_t* off, const int32_t* ind, int32_t nv, int64_t nnz, int32_t* flag, int cus,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_validate, dim3(grid_for((int64_t)nv * 64 + 64, 256, (int64_t)cus * 32)), dim3(256), 0, s,
                     off, ind, nv, nnz, flag);
  return hipGetLastError();
}

hipError_t launch_mismatch(const int32_t* a, const int32_t* b, int64_t n, int32_t* cnt, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mismatch, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, a, b, n, cnt);
  return hipGetLastError();
}

hipError_t launch_copy_i32(int32_t* dst, const int32_t* src, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t n4 = n / 4;
  const int32_t nt = (int32_t)(n - 4 * n4);
  hipLaunchKernelGGL(k_copy_i32, dim3(grid_for(n4 > 0 ? n4 : 1, 256, 16384)), dim3(256), 0, s,
                     reinterpret_cast<int4*>(dst), reinterpret_cast<const int4*>(src), n4, dst + 4 * n4,
                     src + 4 * n4, nt);
  return hipGetLastError();
}

hipError_t launch_graph_totals(const int64_t* tab, int32_t G, const int32_t* out, int64_t* totals, hipStream_t s,
                               bool keep_marked, int64_t nv_total) {
  if (G <= 0) return hipSuccess;
  if (nv_total < 0) {  // one workgroup per graph
    if (keep_marked) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_graph_totals, dim3(G < 65536 ? G : 65536), dim3(256), 0, s, tab, G, out, totals);
    return hipGetLastError();
  }
  hipError_t e = keep_marked ? hipSuccess : launch_fill_i64(totals, G, 0, s);
  if (e != hipSuccess || nv_total == 0) return e;
  const unsigned grid = grid_for((nv_total + 16 * 64 - 1) / (16 * 64) * 64, 256, 8192);
  unsigned long long* t = reinterpret_cast<unsigned long long*>(totals);
  if (keep_marked) hipLaunchKernelGGL(k_graph_sums<true>, dim3(grid), dim3(256), 0, s, tab, G, out, nv_total, t);
  else hipLaunchKernelGGL(k_graph_sums<false>, dim3(grid), dim3(256), 0, s, tab, G, out, nv_total, t);
  return hipGetLastError();
}

hipError_t launch_validate_batch(const int64_t* tab, int32_t G, int32_t* off_cat, int32_t* ind_cat, int64_t nnz_total,
                                 int64_t* totals, int cus, hipStream_t s) {
  if (G <= 0) return hipSuccess;
  hipError_t e = launch_fill_i64(totals, G, 0, s);
  if (e != hipSuccess || nnz_total <= 0) return e;
  hipLaunchKernelGGL(k_validate_batch, dim3(grid_for(nnz_total, 256, (int64_t)cus * 32)), dim3(256), 0, s, tab, G,
                     (const int32_t*)off_cat, (const int32_t*)ind_cat, nnz_total, totals);
  hipLaunchKernelGGL(k_drop_marked, dim3(G < 4096 ? G : 4096), dim3(256), 0, s, tab, G, (const int32_t*)off_cat,
                     ind_cat, (const int64_t*)totals);
  return hipGetLastError();
}

}  // namespace tcd
