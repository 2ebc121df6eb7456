// Two-level Merge Path count kernel (included by tc_kernels.hip).
//
//   BS  threads per workgroup, K merge steps per lane per tile,
//   P   LDS ring length per adjacency list (power of two),
//   M   most edges packed into one small tile.
//
// Level 1: the concatenated Merge Paths of all edges, [0, W), are cut into
// equal ranges, one per workgroup (level1_range); a workgroup starts with a
// binary search for its first edge and that edge's diagonal.
// Level 2: the workgroup walks its range tile by tile, and lane l of a tile
// takes diagonals [l*K, l*K + K): one binary search on the staged lists for
// its start, then a serial merge from LDS.
//
// Ring tiles: one long edge (or the partial edge at a range boundary).  Both
// lists stream through LDS rings of P elements; a tile covers at most
// TMAX = P/2 - 1 diagonals, so while a tile is merged the ring has room for
// every element the next tile can need, and those are loaded into registers
// before the merge and stored after it (one barrier per tile).
// Small tiles: up to M whole edges of at most small_max work, each staged as
// [N(u) | INT_MAX | N(w) | INT_MAX]; lanes cross edge boundaries and add
// their partial counts to per-edge LDS counters.
//
// Match ownership: ties go to A (A[i] <= B[j] takes A), so when A[x] is
// consumed the head of B is the lower bound of A[x] in B, and the match
// A[x] == B[y] is counted exactly once, by the lane that consumes A[x].  No
// hand-off between lanes, tiles or workgroups is needed.
template <int BS, int K, int P, int M>
__global__ __launch_bounds__(BS) void k_count_mp(CountArgs a) {
  constexpr int NW = BS / 64;
  constexpr int LANECAP = BS * K;
  constexpr int RMASK = P - 1;
  constexpr int TMAX = P / 2 - 1;
  constexpr int PF = (TMAX + BS - 1) / BS;
  static_assert((P & (P - 1)) == 0, "ring length must be a power of two");
  static_assert(LANECAP >= TMAX, "lanes must cover a ring tile");
  static_assert(LANECAP + 2 * M <= 2 * P, "a small tile must fit the LDS buffer");
  static_assert(M <= BS, "one metadata slot per thread");

  __shared__ int32_t buf[2 * P];
  __shared__ int32_t m_ws[M + 1];
  __shared__ int32_t m_la[M], m_lb[M], m_cnt[M], m_u[M];
  __shared__ int64_t m_ag[M], m_bg[M];
  __shared__ int64_t red64[NW];
  __shared__ int32_t red32[NW];
  __shared__ int32_t s_ij[2][2];
  __shared__ int64_t s_e;
  __shared__ int32_t s_rng;
  __shared__ int32_t s_start[2];
  __shared__ int32_t s_first[BS];

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t E = a.dscal[0], W = a.dscal[1];
  const int nr = level1_parts(W, a.max_ranges, a.min_units);
  int64_t my_sum = 0;  // sum of edge counts flushed by this thread
  int par = 1;         // s_ij slot the next tile end is written to
  for (;;) {
  if (tid == 0) s_rng = next_range(a.qctr, nr, blockIdx.x & 7);
  __syncthreads();
  const int lg = s_rng;
  __syncthreads();
  if (lg < 0) break;
  int64_t R0, R1;
  level1_range(W, nr, lg, R0, R1);
  if (R0 >= R1) continue;

  if (tid == 0) {
    int64_t lo = 0, hi = E + 1;  // first index with wpre > R0
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (a.wpre[mid] <= R0) lo = mid + 1;
      else hi = mid;
    }
    const int64_t e = lo - 1;
    const int64_t t = R0 - a.wpre[e];
    int32_t i = 0;
    if (t > 0) {
      const int32_t u = a.eu[e], w = a.ew[e];
      const int64_t ab = a.voff[u], bb = a.voff[w];
      const int64_t la = a.voff[u + 1] - ab, lb = a.voff[w + 1] - bb;
      int64_t l2 = t - lb > 0 ? t - lb : 0, h2 = t < la ? t : la;
      while (l2 < h2) {
        int64_t mid = (l2 + h2) >> 1;
        if (a.ind[ab + mid] <= a.ind[bb + t - mid - 1]) l2 = mid + 1;
        else h2 = mid;
      }
      i = (int32_t)l2;
    }
    s_e = e;
    s_start[0] = i;
    s_start[1] = (int32_t)(t - i);
  }
  __syncthreads();
  int64_t e = s_e;
  int32_t ci = s_start[0], cj = s_start[1];
  int64_t pos = R0;
  __syncthreads();

  while (pos < R1) {
    const int64_t we0 = a.wpre[e], we1 = a.wpre[e + 1];
    const int64_t L = we1 - we0;
    const int64_t done = (int64_t)ci + cj;
    const int64_t rem = L - done;
    const bool ring = done > 0 || rem > a.small_max || (R1 - pos) < rem;
    if (ring) {
      // ---------------- one edge streamed through the LDS rings
      const int64_t units = rem < (R1 - pos) ? rem : (R1 - pos);
      const int32_t u = a.eu[e], w = a.ew[e];
      const int64_t ab = a.voff[u], bb = a.voff[w];
      const int32_t la = (int32_t)(a.voff[u + 1] - ab), lb = (int32_t)(a.voff[w + 1] - bb);
      int32_t cnt = 0;
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
      int64_t left = units;
      while (left > 0) {
        const int32_t T = (int32_t)(left < TMAX ? left : TMAX);
        const bool more = left > T;
        // Elements the next tile may need: its window lies in
        // [ci, ci + 2*TMAX); their ring slots held elements < ci (consumed).
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
            const int32_t av = buf[(ci + mid) & RMASK];
            const int32_t bv = buf[P + ((cj + d - mid - 1) & RMASK)];
            if (av <= bv) l2 = mid + 1;
            else h2 = mid;
          }
          int32_t ia = ci + l2, jb = cj + d - l2;
          int32_t av = buf[ia & RMASK], bv = buf[P + (jb & RMASK)];
          const int32_t steps = K < T - d ? K : T - d;
          if (steps == K) {
#pragma unroll
            for (int s = 0; s < K; ++s) {
              const bool ta = av <= bv;
              cnt += (av == bv);
              ia += ta;
              jb += !ta;
              const int32_t nvv = buf[ta ? (ia & RMASK) : (P + (jb & RMASK))];
              av = ta ? nvv : av;
              bv = ta ? bv : nvv;
            }
          } else {
            for (int s = 0; s < steps; ++s) {
              const bool ta = av <= bv;
              cnt += (av == bv);
              ia += ta;
              jb += !ta;
              const int32_t nvv = buf[ta ? (ia & RMASK) : (P + (jb & RMASK))];
              av = ta ? nvv : av;
              bv = ta ? bv : nvv;
            }
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
      const int64_t c = block_sum64<BS>(cnt, red64);
      if (tid == 0 && c) {
        if (a.out) {
          atomicAdd(&a.out[u], (int32_t)c);
          if (!a.u_only) atomicAdd(&a.out[w], (int32_t)c);
        }
        my_sum += c;
      }
      pos += units;
      if (units == rem) {
        ++e;
        ci = 0;
        cj = 0;
      }
    } else {
      // ---------------- several whole short edges
      bool ok = false;
      const int64_t ee = e + tid;
      int32_t cw = 0, wk = 0;
      if (tid < M && ee < E) {
        const int64_t w1 = a.wpre[ee + 1];
        const int64_t cw64 = w1 - we0;
        const int64_t wk64 = w1 - a.wpre[ee];
        ok = cw64 <= LANECAP && cw64 + 2 * (tid + 1) <= 2 * P && wk64 <= a.small_max && w1 <= R1;
        cw = (int32_t)(cw64 < INT_MAX ? cw64 : INT_MAX);
        wk = (int32_t)(wk64 < INT_MAX ? wk64 : INT_MAX);
      }
      const uint64_t bad = __ballot(!ok);
      if (lane == 0) red32[wid] = bad ? (wid * 64 + __builtin_ctzll(bad)) : BS;
      __syncthreads();
      int32_t c = BS;
#pragma unroll
      for (int q = 0; q < NW; ++q) c = red32[q] < c ? red32[q] : c;
      if (tid < c) {
        const int32_t u = a.eu[ee], w = a.ew[ee];
        const int64_t ag = a.voff[u], bg = a.voff[w];
        m_ws[tid] = cw - wk;
        m_la[tid] = (int32_t)(a.voff[u + 1] - ag);
        m_lb[tid] = (int32_t)(a.voff[w + 1] - bg);
        m_cnt[tid] = 0;
        m_u[tid] = u;
        m_ag[tid] = ag;
        m_bg[tid] = bg;
        if (tid == c - 1) m_ws[c] = cw;
        // lanes whose first diagonal t*K lies in this edge start here
        for (int32_t t = (cw - wk + K - 1) / K; t * K < cw; ++t) s_first[t] = tid;
      }
      __syncthreads();
      const int32_t cum = m_ws[c];
      const int32_t d0 = tid * K;
      if (d0 < cum) {
        // Stage this lane's diagonals [d0, d0 + K) of the concatenated
        // [N(u) | N(w)] lists: all loads issued before any store.
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
      {
        if (tid < c) {
          const int32_t base = m_ws[tid] + 2 * tid, la = m_la[tid];
          buf[base + la] = INT_MAX;
          buf[base + la + 1 + m_lb[tid]] = INT_MAX;
        }
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
          if (a.out && !a.u_only) atomicAdd(&a.out[a.ew[ee]], cc);
          my_sum += cc;
        }
        // u side: the first edge of each run of equal u adds the run's sum
        const int32_t u = m_u[tid];
        if (a.out && (tid == 0 || m_u[tid - 1] != u)) {
          int32_t s = 0;
          for (int32_t t2 = tid; t2 < c && m_u[t2] == u; ++t2) s += m_cnt[t2];
          if (s) atomicAdd(&a.out[u], s);
        }
      }
      __syncthreads();
      pos += cum;
      e += c;
      ci = 0;
      cj = 0;
    }
  }
  }  // next level-1 range
  const int64_t tsum = block_sum64<BS>(my_sum, red64);
  if (tid == 0 && a.total && tsum) atomicAdd(a.total, (unsigned long long)(tsum * a.total_mult));
}
