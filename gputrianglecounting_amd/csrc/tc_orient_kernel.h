// Per-vertex counts over the degree-ordered orientation (included by
// tc_kernels.hip after the streaming helpers).
//
// In the orientation (arc v -> w iff (d(v), v) < (d(w), w)) a triangle
// a < b < c is found exactly once: as the element c of N+(a) ∩ N+(b) on
// the arc a -> b.  Crediting 2 to each of a, b and c per triangle gives
// out[v] = 2 T(v), the reference's per-vertex value, and the total 6 T.
//
// The orientation lives in place of the input rows (k_orient_rows): row v
// keeps the positions voff[v] .. voff[v+1] of oind, N+(v) first (orow[2v]
// .. orow[2v+1], ascending), then v's in-neighbours.  Every row r has
// partners p: either its out-neighbours (arcs r -> p) or, with by_in, its
// in-neighbours (arcs p -> r).  Every arc is the pair (r, p) of exactly one
// row either way, and the arc's matches are N+(r) ∩ N+(p).  The second form
// streams N+ of the lower end of each arc, on average the shorter list
// under degree order, but a hub row then has as many partners as the hub
// has lower neighbours.  Work is therefore handed out in pieces of P
// consecutive arc positions (from eight counters pq, one per XCD, ticket n
// of counter x being piece 8 n + x, or with ONE piece i for wave i, when
// the grid has a wave for every piece); prow[b] is the row holding
// position b * P, and a piece walks the rows it overlaps, taking the
// partners that fall inside it, so a hub row spreads over many waves and a
// run of short rows shares one.
//
// N+(r) is staged in the wave's LDS slice A next to 16-bit counters X (two
// to a word), and for every partner p the wave streams N+(p) against A
// exactly as the streamed count kernel does.  A match at A position y adds
// 2 to X[y]; each A element matches at most one element of one N+(p), so
// X[y] <= 2 P within a piece.  After the row's part of the piece, one
// atomic per nonzero X[y] credits vertex A[y].  A vertex that closes many
// triangles (a hub, which has the highest rank and so sits in many rows)
// thus gets about one atomic per row, not one per triangle.  out[p] gets
// 2 |N+(r) ∩ N+(p)| per arc and out[r] the row's sum.
//
// A row longer than AW is walked in windows of AW elements of N+(r); every
// N+(p) is then cut to the window's value range with a wave-wide search,
// so each match still falls in exactly one window.
//
// Rows that fit one window take G partners at a time: the wave splits into
// G groups of 64/G lanes, group g streaming N+(p) of partner j + g, so
// short N+(p) (the usual case) keep every lane busy.  Groups may hit the
// same A position; X is updated with LDS atomics.
//
// Without PV (the total only) there is no X and no per-vertex credit.
//
// Occupancy decides this kernel's speed (it waits on memory and LDS, and
// the time falls in proportion to the waves per SIMD), so the lane's merge
// is kept small: 32-bit chunk and element offsets from a chunk-aligned
// base, and chunks of 8 elements (32 bytes), the current one in registers
// and (PF) the next one in flight.

// Lane part of one list of the orientation: chunks [c0, c1) of CH elements
// from bp (CH * 4-byte aligned); elements before offset e0 or from e1 on are
// masked out.  A: the staged window (n elements, A[n] = INT_MAX).  Every
// match at A position p adds 2 to X's 16-bit counter p; returns the
// matches.
template <int CH, bool PF, bool XC>
__device__ __forceinline__ int32_t orient_stream(const int32_t* __restrict__ bp, int32_t c0, int32_t c1, int32_t e0,
                                                 int32_t e1, const int32_t* A, int32_t n, int32_t* X) {
  constexpr int Q = CH / 4;
  if (c0 >= c1) return 0;
  const int4* b4 = reinterpret_cast<const int4*>(bp);
  int4 nx[Q];
  if (PF) {
#pragma unroll
    for (int q = 0; q < Q; ++q) nx[q] = b4[c0 * Q + q];
  }
  const int32_t bf = bp[c0 * CH > e0 ? c0 * CH : e0];
  int32_t lo = 0, hi = n;  // lower bound of bf in A
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (A[mid] < bf) lo = mid + 1;
    else hi = mid;
  }
  int32_t pos = lo;
  int32_t acur = A[pos];
  int32_t cnt = 0;
  for (int32_t c = c0; c < c1; ++c) {
    int32_t b[CH];
    if (!PF) {
#pragma unroll
      for (int q = 0; q < Q; ++q) nx[q] = b4[c * Q + q];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      b[4 * q] = nx[q].x;
      b[4 * q + 1] = nx[q].y;
      b[4 * q + 2] = nx[q].z;
      b[4 * q + 3] = nx[q].w;
    }
    if (PF && c + 1 < c1) {
#pragma unroll
      for (int q = 0; q < Q; ++q) nx[q] = b4[(c + 1) * Q + q];
    }
    const int32_t cb = c * CH;
    if (cb < e0 || cb + CH > e1) {
#pragma unroll
      for (int j = 0; j < CH; ++j)
        if (cb + j < e0 || cb + j >= e1) b[j] = -1;  // ids are >= 0: never matches, never advances
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int32_t bv = b[j];
      while (acur < bv) acur = A[++pos];
      if (acur == bv) {
        ++cnt;
        if (XC) x_credit(X, pos);
      }
    }
  }
  return cnt;
}

// Per-vertex credit into the caller's output type (int32, or int64 for an
// int64 CSR); v >= 0.
__device__ __forceinline__ void out_add(int32_t* out, int64_t i, int64_t v) { atomicAdd(out + i, (int32_t)v); }
__device__ __forceinline__ void out_add(int64_t* out, int64_t i, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(out + i), (unsigned long long)v);
}

// Dynamic LDS of k_count_orient for waves of aw-id windows: per wave aw + 1
// ids and, with per-vertex counts, (aw + 1) / 2 words of 16-bit counters,
// each wave's two arrays side by side.
inline size_t orient_lds_bytes(int waves, int aw, bool pv) {
  return 4 * (size_t)waves * ((size_t)aw + 1 + (pv ? ((size_t)aw + 1) / 2 : 0));
}

// AW: ids of N+(r) per wave window.  The waves per workgroup (blockDim.x /
// 64, 2 or 4) are a launch argument and the LDS is dynamic, sized by the
// launcher (orient_lds_bytes).
template <int AW, int G, bool PF, int WPE, bool PV, typename OT, bool ONE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_count_orient(
    const int64_t* __restrict__ voff, const int64_t* __restrict__ orow, const int32_t* __restrict__ oind,
    const int32_t* __restrict__ prow, int64_t nv, OT* __restrict__ out, unsigned long long* total,
    unsigned long long* ctr, unsigned long long* pq, int side, int32_t P, int64_t lmax) {
  constexpr int CH = 8;
  extern __shared__ int32_t orient_lds[];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // this wave's window, then its (AW + 1) / 2 words of 16-bit counters
  int32_t* A = orient_lds + wv * (PV ? AW + 1 + (AW + 1) / 2 : AW + 1);
  int32_t* X = A + AW + 1;
  const int32_t sh = (int32_t)((reinterpret_cast<uintptr_t>(oind) >> 2) & (CH - 1));
  const int32_t* ind4 = oind - sh;
  int64_t tri = 0;  // triangles found by this wave
  const int64_t M = voff[nv];  // arc positions
  // side 0: partners are out-neighbours, 1: in-neighbours, 2: in-neighbours
  // unless the orientation listed rows of more than 4096 arcs (their count
  // is ctr[2])
  const bool by_in = side == 2 ? uniform64((int64_t)ctr[2]) == 0 : side != 0;
  // ONE: the grid has a wave for every piece (a small graph), and wave i
  // takes piece i and nothing else: no counter atomics at all.  Otherwise
  // the pieces come from the per-XCD counters, which spread the atomics
  // over eight addresses (one address serialises them, about 9 ns each).
  const int64_t npieces = (M + P - 1) / P;
  const int xcd = blockIdx.x & 7;
  int qk = 0;
  for (;;) {
    int64_t b = 0;
    if (ONE) {
      b = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
    } else {
      b = next_task(pq, npieces, xcd, qk);
      if (b < 0) break;
    }
    const int64_t a0 = b * P;
    if (a0 >= M) break;
    const int64_t a1 = a0 + P < M ? a0 + P : M;
    // rows overlapping [a0, a1): from prow[b] until one ends at or past a1
    // (voff[nv] = M >= a1 stops the walk at row nv - 1 at the latest)
    int64_t u = prow[b];
    int64_t t_next = voff[u];
    for (;; ++u) {
      const int64_t t_lo = t_next;
      t_next = voff[u + 1];
      const int64_t ou = orow[2 * u], oe = orow[2 * u + 1], lu = oe - ou;  // N+(u)
      // partners: out-neighbours at [t_lo, oe), or in-neighbours at [oe, t_next)
      const int64_t p0 = by_in ? oe : t_lo, p1 = by_in ? t_next : oe;
      const int64_t tu = p0 > a0 ? p0 : a0;
      const int64_t lt = (p1 < a1 ? p1 : a1) - tu;  // partners of row u in this piece
      if (lu == 0 || lt <= 0 || lu > lmax) {  // rows over lmax are k_count_long's
        if (t_next >= a1) break;
        continue;
      }
      int64_t cu = 0;
      for (int64_t ws = 0; ws < lu; ws += AW) {
        const int32_t n = (int32_t)(lu - ws < AW ? lu - ws : AW);
        for (int32_t y = lane; y < n; y += 64) A[y] = oind[ou + ws + y];
        if (PV)
          for (int32_t y = lane; y < (n + 1) / 2; y += 64) X[y] = 0;
        if (lane == 0) A[n] = INT_MAX;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const bool whole = n == lu;
        const int32_t amin = oind[ou + ws], amax = oind[ou + ws + n - 1];
        if (G > 1 && whole) {
          constexpr int LG = 64 / G;
          const int gid = lane / LG, gl = lane % LG;
          int64_t cg = 0;  // matches of this group's arcs (summed in every lane of the group)
          for (int64_t j = 0; j < lt; j += G) {
            const int64_t jj = j + gid;
            int32_t c = 0;
            if (jj < lt) {
              const int32_t w = oind[tu + jj];
              const int64_t ow = orow[2 * w], we = orow[2 * w + 1];
              if (we > ow) {
                const int64_t cb = (ow + sh) / CH;  // first chunk of N+(w) in ind4
                const int32_t e0 = (int32_t)(ow + sh - cb * CH), e1 = (int32_t)(we + sh - cb * CH);
                const int32_t nc = (e1 + CH - 1) / CH;
                c = orient_stream<CH, PF, PV>(ind4 + cb * CH, (nc * gl) / LG, (nc * (gl + 1)) / LG, e0, e1, A, n, X);
              }
#pragma unroll
              for (int o = LG / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
              if (PV && gl == 0 && c) out_add(out, w, 2 * c);
            }
            cg += c;
          }
          // row sum: one lane per group holds its group's total
          int64_t rs = gl == 0 ? cg : 0;
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) rs += __shfl_xor(rs, o, 64);
          cu += uniform64(rs);
        } else {
          // (begin, end) of N+ of partner j as 32-bit words in lanes 0-3, one ahead
          auto off_load = [&](int64_t j) -> int32_t {
            if (j >= lt || lane >= 4) return 0;
            const int32_t w = oind[tu + j];
            return reinterpret_cast<const int32_t*>(orow + 2 * (int64_t)w)[lane];
          };
          int32_t offs = off_load(0);
          for (int64_t j = 0; j < lt; ++j) {
            const int32_t w = oind[tu + j];
            const int64_t ow = (int64_t)(uint32_t)__builtin_amdgcn_readlane(offs, 0) |
                               ((int64_t)__builtin_amdgcn_readlane(offs, 1) << 32);
            const int64_t we = (int64_t)(uint32_t)__builtin_amdgcn_readlane(offs, 2) |
                               ((int64_t)__builtin_amdgcn_readlane(offs, 3) << 32);
            offs = off_load(j + 1);
            int64_t g0 = ow + sh, g1 = we + sh;  // positions in ind4
            if (we > ow && !whole) {
              g0 = wave_lower_bound(ind4, g0, g1, amin, lane);
              g1 = wave_lower_bound(ind4, g0, g1, amax == INT_MAX ? amax : amax + 1, lane);
            }
            int32_t c = 0;
            if (g0 < g1) {
              const int64_t cb = g0 / CH;
              const int32_t e0 = (int32_t)(g0 - cb * CH), e1 = (int32_t)(g1 - cb * CH);
              const int32_t nc = (e1 + CH - 1) / CH;
              c = orient_stream<CH, PF, PV>(ind4 + cb * CH, (nc * lane) / 64, (nc * (lane + 1)) / 64, e0, e1, A, n,
                                            X);
            }
            c = wave_sum32(c);
            if (c) {
              if (PV && lane == 0) out_add(out, w, 2 * c);
              cu += c;
            }
          }
        }
        if (PV) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          for (int32_t y = lane; y < n; y += 64) {
            const int32_t x = x_read(X, y);
            if (x) out_add(out, A[y], x);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();  // A and X are rewritten by the next window
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      if (cu) {
        if (PV && lane == 0) out_add(out, u, 2 * cu);
        tri += cu;
      }
      if (t_next >= a1) break;
    }
    if (ONE) break;  // its one piece done
  }
  if (lane == 0 && total && tri) atomicAdd(total, (unsigned long long)(6 * tri));
}
