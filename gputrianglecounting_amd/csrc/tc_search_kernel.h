// Lookup count kernel for skewed edges (included by tc_kernels.hip after
// the level-1 helpers and edge_search()).
//
// It walks the same level-1 ranges as k_count_bs, through its own queues,
// and takes exactly the edges edge_search() classifies: the longer list B
// is at least search_ratio times the shorter list A.  Streaming such a B
// spends almost all of its time on elements that match nothing; here every
// element x of A is instead looked up in the range's part of B by binary
// search in global memory.  A found x is a match owned by the B position it
// was found at, so the parts of an edge split by a range boundary (each
// owning B[t0*lB/L, t1*lB/L) as in k_count_bs) count every match once.
//
// The lookups of up to BS edges of a range are flattened over the
// workgroup's lanes, so the lanes stay busy however short each A is.
// Consecutive lanes take consecutive (sorted) x of the same edge, so the
// upper levels of their searches read the same cache lines.
template <int BS>
__global__ __launch_bounds__(BS) void k_count_search(CountArgs a) {
  constexpr int NW = BS / 64;
  __shared__ int64_t s_pre[BS + 1];  // exclusive prefix of lookups of the taken edges
  __shared__ int64_t s_ga[BS], s_g0[BS], s_len[BS];
  __shared__ int32_t s_u[BS], s_w[BS], s_cnt[BS];
  __shared__ int64_t red64[NW];
  __shared__ int32_t red32[NW];
  __shared__ int64_t s_e[2];
  __shared__ int32_t s_rng;

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t E = a.dscal[0], W = a.dscal[1];
  const int nr = level1_parts(W, a.max_ranges, a.min_units);
  int64_t my_sum = 0;
  for (;;) {
    if (tid == 0) s_rng = next_range(a.qctr, nr, blockIdx.x & 7);
    __syncthreads();
    const int lg = __builtin_amdgcn_readfirstlane(s_rng);
    __syncthreads();
    if (lg < 0) break;
    int64_t R0, R1;
    level1_range(W, nr, lg, R0, R1);
    if (R0 >= R1) continue;
    if (tid == 0) {
      // first edge e0 with wpre[e0] <= R0 < wpre[e0+1], last edge holding R1-1
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
    }
    __syncthreads();
    const int64_t e_first = uniform64(s_e[0]), e_last = uniform64(s_e[1]);

    for (int64_t e0 = e_first; e0 <= e_last; e0 += BS) {
      const int64_t ee = e0 + tid;
      int64_t nlook = 0, ga = 0, g0 = 0, glen = 0;
      int32_t u = 0, w = 0;
      if (ee <= e_last) {
        const int64_t we0 = a.wpre[ee], L = a.wpre[ee + 1] - we0;
        const int64_t t0 = R0 > we0 ? R0 - we0 : 0;
        const int64_t t1 = R1 - we0 < L ? R1 - we0 : L;
        if (!(t0 == 0 && t1 == L && L <= a.small_max)) {
          u = a.eu[ee];
          w = a.ew[ee];
          const int64_t ab = a.voff[u], bb = a.voff[w];
          const int32_t la = (int32_t)(a.voff[u + 1] - ab), lb = (int32_t)(a.voff[w + 1] - bb);
          const bool a_is_u = la <= lb;
          const int32_t lA = a_is_u ? la : lb, lB = a_is_u ? lb : la;
          if (edge_search(lA, lB, a.search_ratio)) {
            const int64_t gB = a_is_u ? bb : ab;
            const int64_t p0 = gB + (t0 * lB) / L, p1 = gB + (t1 * lB) / L;
            if (p0 < p1) {
              nlook = lA;
              ga = a_is_u ? ab : bb;
              g0 = p0;
              glen = p1 - p0;
            }
          }
        }
      }
      // compact the taken edges (order kept) and prefix their lookups
      const bool take = nlook > 0;
      const uint64_t bt = __ballot(take);
      const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
      const int64_t cw = wave_incl_scan64(nlook);
      if (lane == 63) {
        red32[wid] = __popcll(bt);
        red64[wid] = cw;
      }
      __syncthreads();
      int32_t rbase = 0, ntot = 0;
      int64_t wbase = 0, ltot = 0;
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        if (q < wid) {
          rbase += red32[q];
          wbase += red64[q];
        }
        ntot += red32[q];
        ltot += red64[q];
      }
      if (take) {
        const int32_t slot = rbase + __popcll(bt & lt);
        s_pre[slot] = wbase + cw - nlook;
        s_ga[slot] = ga;
        s_g0[slot] = g0;
        s_len[slot] = glen;
        s_u[slot] = u;
        s_w[slot] = w;
        s_cnt[slot] = 0;
      }
      if (tid == 0) s_pre[ntot] = ltot;
      __syncthreads();
      if (ntot == 0) continue;  // uniform: every thread summed the same red32

      int32_t k = 0;  // slot of this lane's current lookup; t only grows
      for (int64_t t = tid; t < ltot; t += BS) {
        // last slot k' >= k with s_pre[k'] <= t (s_pre strictly increases,
        // s_pre[0] = 0 <= t < s_pre[ntot], so k' < ntot)
        int32_t lo = k, hi = ntot;
        while (hi - lo > 1) {
          const int32_t mid = (lo + hi) >> 1;
          if (s_pre[mid] <= t) lo = mid;
          else hi = mid;
        }
        k = lo;
        const int32_t x = a.ind[s_ga[k] + (t - s_pre[k])];
        // lower bound of x in ind[g0, g0 + len): base + n stays g0 + len
        int64_t base = s_g0[k], n = s_len[k];
        const int64_t gend = base + n;
        while (n > 0) {
          const int64_t half = n >> 1;
          const bool less = a.ind[base + half] < x;
          base = less ? base + half + 1 : base;
          n = less ? n - half - 1 : half;
        }
        if (base < gend && a.ind[base] == x) atomicAdd(&s_cnt[k], 1);
      }
      __syncthreads();
      if (tid < ntot) {
        const int32_t cc = s_cnt[tid];
        if (cc) {
          if (a.out) {
            atomicAdd(&a.out[s_u[tid]], cc);
            if (!a.u_only) atomicAdd(&a.out[s_w[tid]], cc);
          }
          my_sum += cc;
        }
      }
      __syncthreads();  // the slot tables are rewritten by the next window
    }
  }  // next level-1 range
  const int64_t tsum = block_sum64<BS>(my_sum, red64);
  if (tid == 0 && a.total && tsum) atomicAdd(a.total, (unsigned long long)(tsum * a.total_mult));
}
