// Per-vertex counts over the degree-ordered orientation (included by
// tc_kernels.hip after the streaming helpers).
//
// In the orientation (arc v -> w iff (d(v), v) < (d(w), w)) a triangle
// a < b < c is found exactly once: as the element c of N+(a) ∩ N+(b) on
// the arc a -> b.  Crediting 2 to each of a, b and c per triangle gives
// out[v] = 2 T(v), the reference's per-vertex value, and the total 6 T.
//
// One wave takes a whole oriented row u at a time (rows are handed out in
// small chunks from a counter): N+(u) is staged in the wave's LDS slice A
// next to a counter array X, and for every w in N+(u) the wave streams
// N+(w) against A exactly as the streamed count kernel does.  A match at A
// position p adds 2 to X[p] (each A element matches at most one element of
// one N+(w), so positions never collide within an edge); after the row,
// one atomic per nonzero X[p] credits vertex A[p].  A vertex that closes
// many triangles (a hub, which has the highest rank and so sits in many
// rows) thus gets about one atomic per row, not one per triangle.  out[w]
// gets 2 |N+(u) ∩ N+(w)| per arc and out[u] the row's sum.
//
// A row longer than AW is walked in windows of AW elements of N+(u); every
// N+(w) is then cut to the window's value range with a wave-wide search,
// so each match still falls in exactly one window.
template <int AW>
__global__ __launch_bounds__(256) void k_count_orient(const int64_t* __restrict__ ooff,
                                                      const int32_t* __restrict__ oind, int64_t nv,
                                                      int32_t* __restrict__ out, unsigned long long* total,
                                                      unsigned long long* ctr, int32_t probe_ratio) {
  constexpr int NW = 4;
  constexpr int CH = 16, SD = 2;
  constexpr int64_t kRows = 4;  // rows per grab
  __shared__ int32_t abuf_all[NW * (AW + 1)];
  __shared__ int32_t xcnt_all[NW * AW];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int32_t* A = abuf_all + wv * (AW + 1);
  int32_t* X = xcnt_all + wv * AW;
  const int32_t sh = (int32_t)((reinterpret_cast<uintptr_t>(oind) >> 2) & (CH - 1));
  const int32_t* ind4 = oind - sh;
  int64_t tri = 0;  // triangles found by this wave
  for (;;) {
    int64_t u0 = 0;
    if (lane == 0) u0 = (int64_t)atomicAdd(ctr, (unsigned long long)kRows);
    u0 = uniform64(__shfl(u0, 0, 64));
    if (u0 >= nv) break;
    const int64_t u1 = u0 + kRows < nv ? u0 + kRows : nv;
    for (int64_t u = u0; u < u1; ++u) {
      const int64_t ou = ooff[u], lu = ooff[u + 1] - ou;
      if (lu == 0) continue;
      int64_t cu = 0;
      for (int64_t ws = 0; ws < lu; ws += AW) {
        const int32_t n = (int32_t)(lu - ws < AW ? lu - ws : AW);
        for (int32_t y = lane; y < n; y += 64) {
          A[y] = oind[ou + ws + y];
          X[y] = 0;
        }
        if (lane == 0) A[n] = INT_MAX;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const bool whole = n == lu;
        const int32_t amin = oind[ou + ws], amax = oind[ou + ws + n - 1];
        // row offsets of w_j as 32-bit words in lanes 0-3, one w ahead
        auto off_load = [&](int64_t j) -> int32_t {
          if (j >= lu || lane >= 4) return 0;
          const int32_t w = oind[ou + j];
          return reinterpret_cast<const int32_t*>(ooff + w + (lane >> 1))[lane & 1];
        };
        int32_t offs = off_load(0);
        for (int64_t j = 0; j < lu; ++j) {
          const int32_t w = oind[ou + j];
          const int64_t ow = (int64_t)(uint32_t)__builtin_amdgcn_readlane(offs, 0) |
                             ((int64_t)__builtin_amdgcn_readlane(offs, 1) << 32);
          const int64_t we = (int64_t)(uint32_t)__builtin_amdgcn_readlane(offs, 2) |
                             ((int64_t)__builtin_amdgcn_readlane(offs, 3) << 32);
          offs = off_load(j + 1);
          const int64_t lw = we - ow;
          int64_t g0 = ow + sh, g1 = we + sh;  // positions in ind4
          if (lw > 0 && !whole) {
            g0 = wave_lower_bound(ind4, g0, g1, amin, lane);
            g1 = wave_lower_bound(ind4, g0, g1, amax == INT_MAX ? amax : amax + 1, lane);
          }
          int32_t c = 0;
          if (g0 < g1) {
            const int64_t cb = g0 / CH, nc = (g1 + CH - 1) / CH - cb;
            const int64_t c_lo = cb + (nc * lane) / 64, c_hi = cb + (nc * (lane + 1)) / 64;
            const bool by_a = probe_ratio > 0 && lw >= (int64_t)probe_ratio * n;
            c = by_a ? stream_chunks_a<CH, SD, true>(ind4, c_lo, c_hi, g0, g1, A, n, X)
                     : stream_chunks<CH, SD, true>(ind4, c_lo, c_hi, g0, g1, A, n, X);
          }
          c = wave_sum32(c);
          if (c) {
            if (lane == 0) atomicAdd(&out[w], 2 * c);
            cu += c;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int32_t y = lane; y < n; y += 64) {
          const int32_t x = X[y];
          if (x) atomicAdd(&out[A[y]], x);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();  // A and X are rewritten by the next window
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      if (cu) {
        if (lane == 0) atomicAdd(&out[u], (int32_t)(2 * cu));
        tri += cu;
      }
    }
  }
  if (lane == 0 && total && tri) atomicAdd(total, (unsigned long long)(6 * tri));
}
