// Streamed Merge Path count kernel (included by tc_kernels.hip after
// tc_mp_kernel.h, whose small-tile and ring code paths it reuses in spirit).
//
//   BS  threads per workgroup        CH  elements per streamed chunk (64 B)
//   P   LDS words = 2P               M   most edges per small tile
//   K   merge steps per lane in small and ring tiles
//   SD  streamed chunks in flight per lane
//
// Level 1 is the same as k_count_mp: the concatenated intersection work of
// all edges, [0, W) with |N(u)| + |N(w)| units per edge, is cut into equal
// ranges that a persistent grid pulls from per-XCD queues.
//
// Level 2 splits one edge's intersection by the longer list.  The shorter
// list A is staged whole in LDS (with an INT_MAX sentinel); the longer list
// B is read from memory in aligned 64-byte chunks straight into registers,
// each lane owning a contiguous run of chunks, and each lane walks A from
// the lower bound of its first element.  A match A[x] == B[y] is counted by
// the lane that owns B[y], so every match is counted once and no lane,
// wave or workgroup hands anything to another.  A range boundary inside an
// edge splits B proportionally: the part covering units [t0, t1) of an edge
// owns B[t0*|B|/L, t1*|B|/L), so the parts of all ranges tile B exactly.
//
// With PH bit 8 the wave phase takes every streamed edge: an A longer than
// the wave's LDS slice is walked in windows, each window meeting only the B
// elements of its value range (found by a wave-wide search), and no edge is
// left to phase 2.
//
// Per range, three phases:
//   1. waves   edges whose shorter list fits a per-wave LDS slice (AW) and
//              whose longer list is at most BWCAP: one wave per edge, edges
//              dealt round-robin to the waves, no workgroup barrier
//   2. group   the remaining edges (not packed into small tiles): shorter
//              list up to 2P-1 streams as above with the whole workgroup;
//              beyond that both lists go through the Merge Path LDS rings
//              (equal diagonals per lane, as in k_count_mp)
//   3. small   whole edges of at most small_max work inside the range,
//              packed up to M to a tile, Merge Path over the staged lists
//
// Edges that edge_search() classifies (B much longer than A) are left to
// k_count_search by phases 1 and 2.
//
// The phases are compiled into separate kernels (PH bit 1 = phase 1, bit 2
// = phase 2, bit 4 = phase 3) so the hot wave phase gets the register and
// LDS budget of its own code; both launches walk the same ranges and split
// the edges by the same rule (a.wave_acap = the phase-1 A capacity AWT).
template <int BS, int CH, int P, int M, int K, int SD, int PH, int AWT>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu((PH & 1) ? 4 : 1))) void k_count_bs(CountArgs a) {
  constexpr int NW = BS / 64;
  constexpr int AW = AWT;                // per-wave A capacity
  constexpr int BUFW = (PH & 6) ? 2 * P : NW * (AW + 1);
  static_assert(NW * (AW + 1) <= BUFW, "wave slices must fit the LDS buffer");
  constexpr int ACAP = 2 * P - 1;       // workgroup A capacity
  constexpr int BWCAP = 64 * CH * 64;   // longest B a single wave streams
  constexpr int LANECAP = BS * K;
  constexpr int RMASK = P - 1;
  constexpr int TMAX = P / 2 - 1;
  constexpr int PF = (TMAX + BS - 1) / BS;
  static_assert((P & (P - 1)) == 0, "P must be a power of two");
  static_assert(CH % 4 == 0, "chunks are read as 16-byte vectors");
  static_assert(LANECAP >= TMAX && LANECAP + 2 * M <= 2 * P && M <= BS, "tile sizes");

  __shared__ int32_t buf[BUFW];
  __shared__ int32_t m_ws[M + 1];
  __shared__ int32_t m_la[M], m_lb[M], m_cnt[M], m_u[M], m_w[M];
  __shared__ int64_t m_ag[M], m_bg[M];
  __shared__ int64_t red64[NW];
  __shared__ int32_t red32[NW];
  __shared__ int32_t s_first[BS];
  __shared__ int32_t s_ij[2][2];
  __shared__ int64_t s_e[2];
  __shared__ int32_t s_start[2];
  __shared__ int32_t s_rng;
  __shared__ int32_t s_ntake;

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t E = a.dscal[0], W = a.dscal[1];
  const int nr = level1_parts(W, a.max_ranges, a.min_units);
  // Chunks are CH*4-byte aligned in memory: element g of ind is element
  // g + sh of the aligned base ind4 (sh < CH words of the aligned block that
  // holds ind[0]).  An aligned chunk never crosses a page boundary, and every
  // chunk read holds at least one element of ind, so every read stays on a
  // page that ind occupies, whatever ind's own alignment and length.
  static_assert((CH & (CH - 1)) == 0, "chunk length must be a power of two");
  const int32_t sh = (int32_t)((reinterpret_cast<uintptr_t>(a.ind) >> 2) & (CH - 1));
  const int32_t* ind4 = a.ind - sh;
  int64_t my_sum = 0;
  int par = 1;
  for (;;) {
    if (tid == 0) s_rng = next_range(a.qctr, nr, blockIdx.x & 7);
    __syncthreads();
    const int lg = __builtin_amdgcn_readfirstlane(s_rng);
    __syncthreads();
    if (lg < 0) break;
    int64_t R0, R1;
    level1_range(W, nr, lg, R0, R1);
    if (R0 >= R1) continue;

    // First and last edge of the range; Merge Path start of a partial first
    // edge (only used if that edge goes through the rings).
    if (tid == 0) {
      int64_t lo = 0, hi = E + 1;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a.wpre[mid] <= R0) lo = mid + 1;
        else hi = mid;
      }
      const int64_t e0 = lo - 1;
      lo = e0;
      hi = E + 1;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a.wpre[mid] <= R1 - 1) lo = mid + 1;
        else hi = mid;
      }
      s_e[0] = e0;
      s_e[1] = lo - 1;
      const int64_t t = R0 - a.wpre[e0];
      int32_t i = 0;
      if (t > 0) {
        const int32_t u = a.eu[e0], w = a.ew[e0];
        const int64_t ab = a.voff[u], bb = a.voff[w];
        const int64_t la = a.voff[u + 1] - ab, lb = a.voff[w + 1] - bb;
        if ((PH & 2) != 0 && (la < lb ? la : lb) > ACAP) {
          int64_t l2 = t - lb > 0 ? t - lb : 0, h2 = t < la ? t : la;
          while (l2 < h2) {
            const int64_t mid = (l2 + h2) >> 1;
            if (a.ind[ab + mid] <= a.ind[bb + t - mid - 1]) l2 = mid + 1;
            else h2 = mid;
          }
          i = (int32_t)l2;
        }
      }
      s_start[0] = i;
      s_start[1] = (int32_t)(t - i);
    }
    __syncthreads();
    const int64_t e_first = uniform64(s_e[0]), e_last = uniform64(s_e[1]);
    const int32_t start_i = __builtin_amdgcn_readfirstlane(s_start[0]);
    const int32_t start_j = __builtin_amdgcn_readfirstlane(s_start[1]);
    __syncthreads();

    // ------------------------------------------------------------ phase 1
    if constexpr ((PH & 1) != 0) {
      const int wv = __builtin_amdgcn_readfirstlane(wid);
      int32_t* abuf = buf + wv * (AW + 1);
      int32_t cnt_w = 0;
      // Edge records are fetched two edges ahead of use: lanes 0-5 each load
      // one 32-bit word of (eu, ew, wpre[e], wpre[e+1]), lanes 0-7 one word
      // of the row offsets; readlane makes them uniform.  Vector loads, so
      // the merge's LDS waits do not wait on them.
      auto rec_load = [&](int64_t e) -> int32_t {
        if (e > e_last || lane >= 6) return 0;
        const int32_t* src = lane == 0 ? a.eu + e : lane == 1 ? a.ew + e
                                                 : reinterpret_cast<const int32_t*>(a.wpre + e) + (lane - 2);
        return *src;
      };
      // row offsets as 32-bit words: lane 2i loads the low and lane 2i+1 the
      // high word of offset i (voff[u], voff[u+1], voff[w], voff[w+1])
      auto off_load = [&](int32_t r, bool valid) -> int32_t {
        const int32_t u = __builtin_amdgcn_readlane(r, 0), w = __builtin_amdgcn_readlane(r, 1);
        if (!valid || lane >= 8) return 0;
        const int64_t* src = a.voff + (lane < 4 ? u : w) + ((lane >> 1) & 1);
        return reinterpret_cast<const int32_t*>(src)[lane & 1];
      };
      auto lo_hi = [](int32_t r, int l) -> int64_t {
        return (int64_t)(uint32_t)__builtin_amdgcn_readlane(r, l) | ((int64_t)__builtin_amdgcn_readlane(r, l + 1) << 32);
      };
      int64_t e = e_first + wv;
      int32_t rec = rec_load(e);
      int32_t rec_n = rec_load(e + NW);
      int32_t offs = off_load(rec, e <= e_last);
      for (; e <= e_last; e += NW) {
        const int32_t u = __builtin_amdgcn_readlane(rec, 0), w = __builtin_amdgcn_readlane(rec, 1);
        const int64_t we0 = lo_hi(rec, 2), L = lo_hi(rec, 4) - we0;
        const int64_t ab = lo_hi(offs, 0), ae = lo_hi(offs, 2), bb = lo_hi(offs, 4), be = lo_hi(offs, 6);
        // start the next two edges' fetches before this edge's work
        const int32_t rec_nn = rec_load(e + 2 * NW);
        const int32_t offs_n = off_load(rec_n, e + NW <= e_last);
        rec = rec_n;
        rec_n = rec_nn;
        offs = offs_n;
        const int64_t t0 = R0 > we0 ? R0 - we0 : 0;
        const int64_t t1 = R1 - we0 < L ? R1 - we0 : L;
        if (t0 == 0 && t1 == L && L <= a.small_max) continue;  // phase 3
        const int32_t la = (int32_t)(ae - ab), lb = (int32_t)(be - bb);
        const bool a_is_u = la <= lb;
        const int32_t lA = a_is_u ? la : lb, lB = a_is_u ? lb : la;
        if (edge_search(lA, lB, a.search_ratio)) continue;  // k_count_search
        if constexpr ((PH & 8) == 0) {
          if (lA > a.wave_acap || lB > BWCAP) continue;  // phase 2
        }
        const int64_t gA = a_is_u ? ab : bb, gB = a_is_u ? bb : ab;
        // stage A[x0, x0 + n) (n <= AW) into abuf[0, n) with abuf[n] = INT_MAX
        auto stage = [&](int64_t x0, int32_t n) {
          for (int32_t y0 = 0; y0 < n; y0 += 8 * 64) {
            int32_t v[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
              const int32_t y = y0 + lane + r * 64;
              v[r] = y < n ? a.ind[gA + x0 + y] : 0;
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) {
              const int32_t y = y0 + lane + r * 64;
              if (y < n) abuf[y] = v[r];
            }
          }
          if (lane == 0) abuf[n] = INT_MAX;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        };
        // B elements at ind4 positions [j0, j1) against the staged A; an
        // edge with B at least probe_ratio times longer than A is walked
        // from A's side (wave-uniform choice)
        const bool by_a = a.probe_ratio > 0 && (int64_t)lB >= (int64_t)a.probe_ratio * lA;
        auto stream = [&](int64_t j0, int64_t j1, int32_t n) -> int32_t {
          const int64_t cb = j0 / CH, nc = (j1 + CH - 1) / CH - cb;
          const int64_t c_lo = cb + (nc * lane) / 64, c_hi = cb + (nc * (lane + 1)) / 64;
          return by_a ? stream_chunks_a<CH, SD>(ind4, c_lo, c_hi, j0, j1, abuf, n)
                      : stream_chunks<CH, SD>(ind4, c_lo, c_hi, j0, j1, abuf, n);
        };
        const int64_t g0 = gB + sh + (t0 * lB) / L, g1 = gB + sh + (t1 * lB) / L;
        int32_t c = 0;
        if (lA <= AW) {
          stage(0, lA);
          if (g0 < g1) c = stream(g0, g1, lA);
        } else if (g0 < g1) {
          // A longer than the wave's slice (only with PH & 8): A in windows
          // of AW.  Only A[ia, ia_end) can match B's part; window A[ia, ia+n)
          // meets exactly the B elements from jb up to the first one above
          // its last element, and the next window starts after both.
          const int32_t* Ag = a.ind + gA;
          int64_t ia = wave_lower_bound(Ag, 0, lA, ind4[g0], lane);
          const int64_t ia_end = wave_lower_bound(Ag, ia, lA, ind4[g1 - 1] + 1, lane);
          int64_t jb = g0;
          while (ia < ia_end && jb < g1) {
            const int32_t n = (int32_t)(ia_end - ia < AW ? ia_end - ia : AW);
            stage(ia, n);
            const int64_t je = ia + n >= ia_end ? g1 : wave_lower_bound(ind4, jb, g1, Ag[ia + n - 1] + 1, lane);
            if (jb < je) c += stream(jb, je, n);
            jb = je;
            ia += n;
            __builtin_amdgcn_wave_barrier();  // abuf is rewritten by the next window
          }
        }
        c = wave_sum32(c);
        if (lane == 0 && c) {
          if (a.out) {
            atomicAdd(&a.out[u], c);
            if (!a.u_only) atomicAdd(&a.out[w], c);
          }
          cnt_w += c;
        }
        __builtin_amdgcn_wave_barrier();  // abuf is rewritten by the next edge
      }
      my_sum += cnt_w;
    }
    __syncthreads();

    // ------------------------------------------------------------ phase 2
    if constexpr ((PH & 2) != 0)
    for (int64_t e0 = e_first; e0 <= e_last; e0 += BS) {
      // which of the next BS edges belong to this phase
      bool mine = false;
      const int64_t ee = e0 + tid;
      if (ee <= e_last) {
        const int64_t we0 = a.wpre[ee], L = a.wpre[ee + 1] - we0;
        const int64_t t0 = R0 > we0 ? R0 - we0 : 0;
        const int64_t t1 = R1 - we0 < L ? R1 - we0 : L;
        if (!(t0 == 0 && t1 == L && L <= a.small_max)) {
          const int32_t u = a.eu[ee], w = a.ew[ee];
          const int64_t la = a.voff[u + 1] - a.voff[u], lb = a.voff[w + 1] - a.voff[w];
          const int64_t lA = la < lb ? la : lb, lB = la < lb ? lb : la;
          mine = !edge_search((int32_t)lA, (int32_t)lB, a.search_ratio) && (lA > a.wave_acap || lB > BWCAP);
        }
      }
      const uint64_t bal = __ballot(mine);
      if (lane == 0) red64[wid] = (int64_t)bal;
      __syncthreads();
      uint64_t masks[NW];
#pragma unroll
      for (int q = 0; q < NW; ++q) masks[q] = (uint64_t)red64[q];
      __syncthreads();
#pragma unroll 1
      for (int q = 0; q < NW; ++q) {
        uint64_t m = masks[q];
        while (m) {
          const int bit = __builtin_ctzll(m);
          m &= m - 1;
          const int64_t e = e0 + q * 64 + bit;
          const int64_t we0 = a.wpre[e], L = a.wpre[e + 1] - we0;
          const int64_t t0 = R0 > we0 ? R0 - we0 : 0;
          const int64_t t1 = R1 - we0 < L ? R1 - we0 : L;
          const int32_t u = a.eu[e], w = a.ew[e];
          const int64_t ab = a.voff[u], bb = a.voff[w];
          const int32_t la = (int32_t)(a.voff[u + 1] - ab), lb = (int32_t)(a.voff[w + 1] - bb);
          const bool a_is_u = la <= lb;
          const int32_t lA = a_is_u ? la : lb, lB = a_is_u ? lb : la;
          int32_t cnt = 0;
          if (lA <= ACAP) {
            // workgroup stream
            const int64_t gA = a_is_u ? ab : bb, gB = a_is_u ? bb : ab;
            for (int32_t x0 = 0; x0 < lA; x0 += 8 * BS) {
              int32_t v[8];
#pragma unroll
              for (int r = 0; r < 8; ++r) {
                const int32_t x = x0 + tid + r * BS;
                v[r] = x < lA ? a.ind[gA + x] : 0;
              }
#pragma unroll
              for (int r = 0; r < 8; ++r) {
                const int32_t x = x0 + tid + r * BS;
                if (x < lA) buf[x] = v[r];
              }
            }
            if (tid == 0) buf[lA] = INT_MAX;
            __syncthreads();
            const int64_t g0 = gB + sh + (t0 * lB) / L, g1 = gB + sh + (t1 * lB) / L;
            if (g0 < g1) {
              const int64_t cb = g0 / CH, nc = (g1 + CH - 1) / CH - cb;
              const int64_t c_lo = cb + (nc * tid) / BS, c_hi = cb + (nc * (tid + 1)) / BS;
              cnt = stream_chunks<CH, SD>(ind4, c_lo, c_hi, g0, g1, buf, lA);
            }
          } else {
            // both lists longer than the LDS: Merge Path through the rings
            int32_t ci = 0, cj = 0;
            if (t0 > 0) {
              ci = start_i;  // only the range's first edge can start inside
              cj = start_j;
            }
            int32_t a_hi, b_hi;
            {
              const int32_t a_to = ci + TMAX < la ? ci + TMAX : la;
              const int32_t b_to = cj + TMAX < lb ? cj + TMAX : lb;
              int32_t ra[PF], rb[PF];
#pragma unroll
              for (int r = 0; r < PF; ++r) {
                const int32_t x = ci + tid + r * BS, y = cj + tid + r * BS;
                ra[r] = x < a_to ? a.ind[ab + x] : 0;
                rb[r] = y < b_to ? a.ind[bb + y] : 0;
              }
#pragma unroll
              for (int r = 0; r < PF; ++r) {
                const int32_t x = ci + tid + r * BS, y = cj + tid + r * BS;
                if (x < a_to) buf[x & RMASK] = ra[r];
                if (y < b_to) buf[P + (y & RMASK)] = rb[r];
              }
              if (tid == 0) {
                if (a_to == la) buf[la & RMASK] = INT_MAX;
                if (b_to == lb) buf[P + (lb & RMASK)] = INT_MAX;
              }
              a_hi = a_to;
              b_hi = b_to;
            }
            __syncthreads();
            int64_t left = t1 - t0;
            while (left > 0) {
              const int32_t T = (int32_t)(left < TMAX ? left : TMAX);
              const bool more = left > T;
              int32_t ra[PF], rb[PF];
              int32_t pa_to = a_hi, pb_to = b_hi;
              if (more) {
                pa_to = ci + 2 * TMAX < la ? ci + 2 * TMAX : la;
                pb_to = cj + 2 * TMAX < lb ? cj + 2 * TMAX : lb;
#pragma unroll
                for (int r = 0; r < PF; ++r) {
                  const int32_t x = a_hi + tid + r * BS, y = b_hi + tid + r * BS;
                  ra[r] = x < pa_to ? a.ind[ab + x] : 0;
                  rb[r] = y < pb_to ? a.ind[bb + y] : 0;
                }
              }
              const int32_t d = tid * K;
              if (d < T) {
                int32_t l2 = d - (lb - cj);
                l2 = l2 > 0 ? l2 : 0;
                int32_t h2 = d < la - ci ? d : la - ci;
                while (l2 < h2) {
                  const int32_t mid = (l2 + h2) >> 1;
                  if (buf[(ci + mid) & RMASK] <= buf[P + ((cj + d - mid - 1) & RMASK)]) l2 = mid + 1;
                  else h2 = mid;
                }
                int32_t ia = ci + l2, jb = cj + d - l2;
                int32_t av = buf[ia & RMASK], bv = buf[P + (jb & RMASK)];
                const int32_t steps = K < T - d ? K : T - d;
                for (int s = 0; s < steps; ++s) {
                  const bool ta = av <= bv;
                  cnt += (av == bv);
                  ia += ta;
                  jb += !ta;
                  const int32_t nvv = buf[ta ? (ia & RMASK) : (P + (jb & RMASK))];
                  av = ta ? nvv : av;
                  bv = ta ? bv : nvv;
                }
                if (d + steps == T) {
                  s_ij[par][0] = ia;
                  s_ij[par][1] = jb;
                }
              }
              if (more) {
#pragma unroll
                for (int r = 0; r < PF; ++r) {
                  const int32_t x = a_hi + tid + r * BS, y = b_hi + tid + r * BS;
                  if (x < pa_to) buf[x & RMASK] = ra[r];
                  if (y < pb_to) buf[P + (y & RMASK)] = rb[r];
                }
                if (tid == 0) {
                  if (pa_to == la) buf[la & RMASK] = INT_MAX;
                  if (pb_to == lb) buf[P + (lb & RMASK)] = INT_MAX;
                }
                a_hi = pa_to;
                b_hi = pb_to;
              }
              __syncthreads();
              ci = s_ij[par][0];
              cj = s_ij[par][1];
              par ^= 1;
              left -= T;
            }
          }
          const int64_t c = block_sum64<BS>(cnt, red64);  // also orders buf reuse
          if (tid == 0 && c) {
            if (a.out) {
              atomicAdd(&a.out[u], (int32_t)c);
              if (!a.u_only) atomicAdd(&a.out[w], (int32_t)c);
            }
            my_sum += c;
          }
        }
      }
    }

    // ------------------------------------------------------------ phase 3
    if constexpr ((PH & 4) != 0)
    for (int64_t e_scan = e_first; e_scan <= e_last;) {
      const int64_t ee = e_scan + tid;
      bool q = false;
      int32_t wk = 0;
      if (ee <= e_last) {
        const int64_t we0 = a.wpre[ee], L = a.wpre[ee + 1] - we0;
        q = we0 >= R0 && we0 + L <= R1 && L <= a.small_max;
        wk = q ? (int32_t)L : 0;
      }
      // inclusive prefix over threads of (q, work)
      const uint64_t bq = __ballot(q);
      const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
      int32_t rank = __popcll(bq & lt) + (q ? 1 : 0);
      int64_t cw = wave_incl_scan64(wk);
      if (lane == 63) {
        red32[wid] = __popcll(bq);
        red64[wid] = cw;
      }
      __syncthreads();
      for (int w2 = 0; w2 < wid; ++w2) {
        rank += red32[w2];
        cw += red64[w2];
      }
      const bool take = q && rank <= M && cw <= LANECAP && cw + 2 * rank <= 2 * P;
      const uint64_t bt = __ballot(take);
      __syncthreads();
      if (lane == 0) red32[wid] = bt ? (int32_t)(wid * 64 + 63 - __builtin_clzll(bt)) : -1;
      __syncthreads();
      int32_t last = -1;
#pragma unroll
      for (int q2 = 0; q2 < NW; ++q2) last = red32[q2] > last ? red32[q2] : last;
      if (last < 0) {
        // no short whole edge among these BS edges
        __syncthreads();
        e_scan += BS;
        continue;
      }
      int32_t c = 0;
      if (take) {
        const int32_t slot = rank - 1;
        const int32_t u = a.eu[ee], w = a.ew[ee];
        const int64_t ag = a.voff[u], bg = a.voff[w];
        const int32_t ws = (int32_t)cw - wk;
        m_ws[slot] = ws;
        m_la[slot] = (int32_t)(a.voff[u + 1] - ag);
        m_lb[slot] = (int32_t)(a.voff[w + 1] - bg);
        m_cnt[slot] = 0;
        m_u[slot] = u;
        m_w[slot] = w;
        m_ag[slot] = ag;
        m_bg[slot] = bg;
        for (int32_t t = (ws + K - 1) / K; t * K < (int32_t)cw; ++t) s_first[t] = slot;
      }
      if (tid == last) {
        s_ntake = rank;  // number of edges taken
        m_ws[rank] = (int32_t)cw;
      }
      __syncthreads();
      c = s_ntake;
      const int32_t cum = m_ws[c];
      const int32_t d0 = tid * K;
      if (d0 < cum) {
        int32_t v[K], dst[K];
        int32_t k = s_first[tid];
        int32_t ws = m_ws[k], la = m_la[k], lab = la + m_lb[k];
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const int32_t x = d0 + r;
          dst[r] = -1;
          v[r] = 0;
          if (x < cum) {
            while (x - ws >= lab) {
              ++k;
              ws = m_ws[k];
              la = m_la[k];
              lab = la + m_lb[k];
            }
            const int32_t p = x - ws;
            const int64_t src = p < la ? m_ag[k] + p : m_bg[k] + (p - la);
            v[r] = a.ind[src];
            dst[r] = ws + 2 * k + p + (p >= la ? 1 : 0);
          }
        }
#pragma unroll
        for (int r = 0; r < K; ++r)
          if (dst[r] >= 0) buf[dst[r]] = v[r];
      }
      if (tid < c) {
        const int32_t base = m_ws[tid] + 2 * tid, la = m_la[tid];
        buf[base + la] = INT_MAX;
        buf[base + la + 1 + m_lb[tid]] = INT_MAX;
      }
      __syncthreads();
      if (d0 < cum) {
        int32_t k = s_first[tid];
        int32_t left = K < cum - d0 ? K : cum - d0;
        int32_t d = d0 - m_ws[k];
        int32_t la = m_la[k], lb = m_lb[k];
        int32_t abase = m_ws[k] + 2 * k, bbase = abase + la + 1;
        int32_t x0 = d - lb > 0 ? d - lb : 0, x1 = d < la ? d : la;
        while (x0 < x1) {
          const int32_t mid = (x0 + x1) >> 1;
          if (buf[abase + mid] <= buf[bbase + d - mid - 1]) x0 = mid + 1;
          else x1 = mid;
        }
        int32_t ia = x0, jb = d - x0;
        int32_t av = buf[abase + ia], bv = buf[bbase + jb];
        int32_t cnt = 0;
        for (;;) {
          const int32_t room = la + lb - d;
          const int32_t n = left < room ? left : room;
          for (int32_t s = 0; s < n; ++s) {
            const bool ta = av <= bv;
            cnt += (av == bv);
            ia += ta;
            jb += !ta;
            const int32_t nvv = buf[ta ? abase + ia : bbase + jb];
            av = ta ? nvv : av;
            bv = ta ? bv : nvv;
          }
          left -= n;
          d += n;
          if (d == la + lb) {
            if (cnt) atomicAdd(&m_cnt[k], cnt);
            cnt = 0;
            if (left == 0) break;
            ++k;
            la = m_la[k];
            lb = m_lb[k];
            abase = m_ws[k] + 2 * k;
            bbase = abase + la + 1;
            d = 0;
            ia = 0;
            jb = 0;
            av = buf[abase];
            bv = buf[bbase];
          } else {
            break;
          }
        }
        if (cnt) atomicAdd(&m_cnt[k], cnt);
      }
      __syncthreads();
      if (tid < c) {
        const int32_t cc = m_cnt[tid];
        if (cc) {
          if (a.out && !a.u_only) atomicAdd(&a.out[m_w[tid]], cc);
          my_sum += cc;
        }
        const int32_t u = m_u[tid];
        if (a.out && (tid == 0 || m_u[tid - 1] != u)) {
          int32_t s = 0;
          for (int32_t t2 = tid; t2 < c && m_u[t2] == u; ++t2) s += m_cnt[t2];
          if (s) atomicAdd(&a.out[u], s);
        }
      }
      __syncthreads();
      // continue after the last edge taken (later qualifying edges of this
      // window that did not fit are picked up by the next tile)
      e_scan = e_scan + last + 1;
    }
  }  // next level-1 range
  const int64_t tsum = block_sum64<BS>(my_sum, red64);
  if (tid == 0 && a.total && tsum) atomicAdd(a.total, (unsigned long long)(tsum * a.total_mult));
}
