// Per-vertex counts over the degree-ordered orientation with the row's
// list held in an LDS hash table (included by tc_kernels.hip after
// tc_orient_kernel.h, whose work split, orientation and credit rule it
// shares: pieces of P arc positions taken from a counter, or piece i for
// wave i; row parts; partners by out- or in-neighbours; 2 per triangle to
// each of its three vertices).
//
// What differs is how a partner's list N+(p) meets the row's N+(r):
//  - A window of at most AW elements of N+(r) is inserted into a two-choice
//    cuckoo table of buckets in the wave's LDS: nb buckets of two slots
//    (keys K[2b], K[2b+1]: a vertex id, or -1 when empty), an id x living in
//    bucket h1(x) or h2(x) (multiplicative hashes scaled to nb), beside a
//    16-bit match counter per slot.  The load is at most 3/4.  The lanes
//    insert in parallel with LDS compare-and-swaps into free slots of
//    either bucket and, when both are full, an exchange that displaces an
//    id to its other bucket; an id still displaced after kCuckooSteps moves
//    fails the build, which is redone with the next of kCuckooSeeds hash
//    pairs.  If every pair fails, the window is laid out as a sorted array
//    instead and looked up by binary search (also chosen for every window
//    with `sorted`).
//  - N+(p) is read by the lanes of a group as consecutive 16-byte vectors:
//    lane l of a group of LG lanes holds elements 4l .. 4l+3 of every
//    4·LG-element block, so one load instruction covers whole cache lines
//    and no line is fetched again by a later instruction of the wave.
//  - Every element is looked up with exactly two independent 8-byte LDS
//    reads (its two candidate buckets), the same for every lane: no lane
//    walks the row, searches for where its part of N+(p) starts, or follows
//    a probe chain that the other lanes of the wave wait for.
// A match in slot t adds 2 to the counter of t; after the row part, one
// atomic per nonzero counter credits the vertex K[t].  A slot's counter
// gets at most one match per partner, so it stays below 2 P.
//
// A row longer than AW is walked in windows; every N+(p) is then cut to the
// window's value range with a wave-wide search, so each match falls in
// exactly one window.

constexpr int kCuckooSteps = 128;
// The slot of a full bucket an insertion displaces at step it: a hash bit
// of the id and the step.  Alternating slots by step alone sent many chains
// round the same few buckets: replaying the builds of every window of
// R-MAT s22 on the host, 14.6% of the windows (weighted by their partners)
// needed a second hash pair and 1.8% fell back to the sorted layout; with
// the hashed slot and 128 steps, 0.3% and 0.2% (more steps alone changed
// nothing: those chains were cycles).
__device__ __forceinline__ int cuckoo_slot(int32_t id, int it) {
  return (int)(((((uint32_t)(id + it)) & 0xffffffu) * 0x9E3779u >> 13) & 1u);
}
constexpr int kCuckooSeeds = 4;
// Odd multipliers of the hash pairs, one pair per build attempt.
__device__ __forceinline__ uint32_t cuckoo_mult(int i) {
  switch (i) {
    case 0: return 0x9E3779B1u;
    case 1: return 0x85EBCA77u;
    case 2: return 0xC2B2AE3Du;
    case 3: return 0x27D4EB2Fu;
    case 4: return 0x165667B1u;
    case 5: return 0xD3A2646Du;
    case 6: return 0xFD7046C5u;
    default: return 0xB55A4F09u;
  }
}

// Bucket of an id among nb buckets (nb < 2^16) under multiplier c: the top
// 24 bits x of the id's low 24 bits times c (mod 2^32), scaled to nb as
// x * nb / 2^24.  Both products are full-rate 24-bit multiplies (a 32-bit
// multiply issues at a quarter of the rate, and the lookups are bound by
// instruction issue).  Ids of 2^24 and more share the buckets of their low
// 24 bits: a window where that overfills a bucket pair fails its builds and
// is laid out sorted instead, so counts never depend on it.  The scaling is
// the high word of a 24 x 24-bit product (x times nb * 2^8).
//
// Both products are written as the instructions themselves (mul_u24,
// mulhi_u24): the compiler picks the 24-bit multiplies only when it can see,
// in the same basic block, that both operands fit 24 bits, and the
// multipliers come into the probe loops from outside them, so __umul24 and
// the 64-bit product there became masks and quarter-rate 32-bit multiplies.
__device__ __forceinline__ uint32_t mul_u24(uint32_t a, uint32_t c) {
  uint32_t r;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "s"(c), "v"(a));  // low 32 bits of a[23:0] * c[23:0]
  return r;
}
__device__ __forceinline__ uint32_t mulhi_u24(uint32_t a, uint32_t c) {
  uint32_t r;
  asm("v_mul_hi_u32_u24 %0, %1, %2" : "=v"(r) : "s"(c), "v"(a));  // bits 32..47 of a[23:0] * c[23:0]
  return r;
}
__device__ __forceinline__ uint32_t hash_bucket(int32_t key, uint32_t c, uint32_t nb) {
  return mulhi_u24(mul_u24((uint32_t)key, c) >> 8, (nb << 8) & 0xffffffu);
}

// Bucket of an id among 2^(32 - shift) buckets: the top bits of its low
// 24 bits times c (mod 2^32), one multiply and one shift (small tables).
__device__ __forceinline__ uint32_t hash_bucket_p2(int32_t key, uint32_t c, uint32_t shift) {
  return mul_u24((uint32_t)key, c) >> shift;
}

// Slots of the table of a window of n ids on at most cap slots: nb buckets
// of two, nb >= (2n + 2) / 3 (load at most 3/4): a power of two from 32 up
// when 2 nb fits cap (buckets from one multiply and a shift), else the next
// multiple of 32.
__device__ __forceinline__ uint32_t table_slots(int32_t n, uint32_t cap) {
  const uint32_t need = (uint32_t)((2 * n + 2) / 3);
  uint32_t p2 = 32;
  while (p2 < need) p2 *= 2;
  return 2 * p2 <= cap ? 2 * p2 : 2 * ((need + 31) / 32 * 32);
}

// The table of one window: K (2 nb slots), the hash pair (c1, c2), or
// sorted (n ids ascending in K[0 .. n), searched in `top` halving steps,
// top = the largest power of two <= n).  shift > 0: nb = 2^(32 - shift) and
// buckets come from hash_bucket_p2.
struct HashTab {
  const int32_t* K;
  uint32_t nb, c1, c2, shift;
  int32_t n, top;
  bool sorted;
};

// Looks the elements of e[0..N) up in the table (an element outside the
// list holds kNoId, which equals no key and no empty slot): a hit in slot
// 2b + h adds 2 to the slot's 16-bit counter, half h of counter word b
// (XC); returns the hits.
constexpr int32_t kNoId = -2;
template <int N, bool XC, bool P2>
__device__ __forceinline__ int32_t hash_probe(const int32_t (&e)[N], const HashTab& t, int32_t* X) {
  int32_t hits = 0;
  if (!t.sorted) {
    const int2* K2 = reinterpret_cast<const int2*>(t.K);
    uint32_t b1[N], b2[N];
    int2 k1[N], k2[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
      b1[q] = P2 ? hash_bucket_p2(e[q], t.c1, t.shift) : hash_bucket(e[q], t.c1, t.nb);
      b2[q] = P2 ? hash_bucket_p2(e[q], t.c2, t.shift) : hash_bucket(e[q], t.c2, t.nb);
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
      k1[q] = K2[b1[q]];
      k2[q] = K2[b2[q]];
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
      const bool in1 = k1[q].x == e[q] || k1[q].y == e[q];
      const bool hit = in1 || k2[q].x == e[q] || k2[q].y == e[q];
      if (hit) {  // a key sits in one slot only
        ++hits;
        if (XC) {
          const bool upper = k1[q].y == e[q] || k2[q].y == e[q];
          atomicAdd(&X[in1 ? b1[q] : b2[q]], upper ? 0x20000 : 2);
        }
      }
    }
  } else {
    // one element at a time: this path only serves windows whose cuckoo
    // builds failed (or `sorted`), so it is kept small in registers
#pragma unroll
    for (int q = 0; q < N; ++q) {
      int32_t pos = 0;
      for (int32_t st = t.top; st > 0; st >>= 1)
        if (pos + st <= t.n && t.K[pos + st - 1] < e[q]) pos += st;
      if (pos < t.n && t.K[pos] == e[q]) {
        ++hits;
        if (XC) x_credit(X, pos);
      }
    }
  }
  return hits;
}

// The elements at positions [g0, g1) of bp (16-byte aligned) looked up in
// the table by lane gl of a group of LG lanes: vectors gl, gl + LG, ...,
// the next one in flight while the current one is probed.  Only a list's
// first and last vectors hold elements of other lists, which are set to
// kNoId.  The next vector is loaded whatever the trip count (the last one
// again past the end), so no branch keeps the old value alive: a
// conditional load cost four register copies per vector.
template <int LG, bool XC, bool P2>
__device__ __forceinline__ int32_t hash_stream_t(const int4* __restrict__ b4, int32_t nvec, int32_t lo, int32_t hi,
                                                 int gl, const HashTab& t, int32_t* X) {
  int32_t cnt = 0;
  if (gl >= nvec) return 0;
  int4 nx = b4[gl];
  for (int32_t j = gl; j < nvec; j += LG) {
    const int4 a = nx;
    nx = b4[j + LG < nvec ? j + LG : nvec - 1];  // past the end: the last vector again (cached, unused)
    int32_t e[4] = {a.x, a.y, a.z, a.w};
    if ((uint32_t)(j - 1) >= (uint32_t)(nvec - 2)) {  // j == 0 or j == nvec - 1
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if ((uint32_t)(4 * j + q - lo) >= (uint32_t)(hi - lo)) e[q] = kNoId;
    }
    cnt += hash_probe<4, XC, P2>(e, t, X);
  }
  return cnt;
}

template <int LG, bool XC>
__device__ __forceinline__ int32_t hash_stream(const int32_t* __restrict__ bp, int64_t g0, int64_t g1, int gl,
                                               const HashTab& t, int32_t* X) {
  if (g0 >= g1) return 0;
  const int4* b4 = reinterpret_cast<const int4*>(bp) + (g0 >> 2);
  const int32_t nvec = (int32_t)(((g1 + 3) >> 2) - (g0 >> 2));
  const int32_t lo = (int32_t)(g0 & 3), hi = (int32_t)(g1 - (g0 & ~int64_t(3)));  // valid element offsets
  if (t.shift) return hash_stream_t<LG, XC, true>(b4, nvec, lo, hi, gl, t, X);
  return hash_stream_t<LG, XC, false>(b4, nvec, lo, hi, gl, t, X);
}

// Builds the table of the window src[0 .. n) (ascending ids) in K, tsz
// slots: cuckoo insertion under successive hash pairs, or the sorted layout
// when `sorted` is asked for or every pair fails.  Whole wave; K is ready
// for lookups on return.
__device__ __forceinline__ HashTab hash_build(const int32_t* __restrict__ src, int32_t n, int32_t* K, uint32_t tsz,
                                              bool sorted, int lane) {
  HashTab t;
  t.K = K;
  t.nb = tsz / 2;
  t.shift = (t.nb & (t.nb - 1)) == 0 ? 32u - (uint32_t)__builtin_ctz(t.nb) : 0u;
  t.n = n;
  t.top = 0;
  t.sorted = true;
  t.c1 = t.c2 = 0;
  auto bucket = [&](int32_t key, uint32_t c) {
    return t.shift ? hash_bucket_p2(key, c, t.shift) : hash_bucket(key, c, t.nb);
  };
  for (int r = 0; r < kCuckooSeeds && !sorted; ++r) {
    const uint32_t c1 = cuckoo_mult(2 * r), c2 = cuckoo_mult(2 * r + 1);
    for (uint32_t y = lane; y < tsz; y += 64) K[y] = -1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bool fail = false;
    for (int32_t y = lane; y < n; y += 64) {
      int32_t cur = src[y];
      uint32_t b = bucket(cur, c1);  // the bucket cur is tried in first
      for (int it = 0;; ++it) {
        if (atomicCAS(&K[2 * b], -1, cur) == -1 || atomicCAS(&K[2 * b + 1], -1, cur) == -1) break;
        const uint32_t h1 = bucket(cur, c1);
        const uint32_t o = b == h1 ? bucket(cur, c2) : h1;  // cur's other bucket
        if (atomicCAS(&K[2 * o], -1, cur) == -1 || atomicCAS(&K[2 * o + 1], -1, cur) == -1) break;
        if (it == kCuckooSteps) {
          fail = true;
          break;
        }
        // both buckets full: displace a resident of the other bucket, which
        // then tries its own other bucket
        const int32_t old = atomicExch(&K[2 * o + cuckoo_slot(cur, it)], cur);
        if (old == -1) break;
        cur = old;
        b = o;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!__ballot(fail)) {
      t.sorted = false;
      t.c1 = c1;
      t.c2 = c2;
      return t;
    }
  }
  for (uint32_t y = lane; y < tsz; y += 64) K[y] = (int32_t)y < n ? src[y] : INT_MAX;
  int32_t top = 1;
  while (2 * top <= n) top *= 2;
  t.top = n > 0 ? top : 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return t;
}

// The table of src[0 .. n) built by the whole workgroup (256 threads) in K,
// tsz slots: as hash_build, with the workgroup's barriers.
__device__ __forceinline__ HashTab hash_build_wg(const int32_t* __restrict__ src, int32_t n, int32_t* K, uint32_t tsz) {
  HashTab t;
  t.K = K;
  t.nb = tsz / 2;
  t.shift = (t.nb & (t.nb - 1)) == 0 ? 32u - (uint32_t)__builtin_ctz(t.nb) : 0u;
  t.n = n;
  t.top = 0;
  t.sorted = true;
  t.c1 = t.c2 = 0;
  auto bucket = [&](int32_t key, uint32_t c) {
    return t.shift ? hash_bucket_p2(key, c, t.shift) : hash_bucket(key, c, t.nb);
  };
  for (int r = 0; r < kCuckooSeeds; ++r) {
    const uint32_t c1 = cuckoo_mult(2 * r), c2 = cuckoo_mult(2 * r + 1);
    for (uint32_t y = threadIdx.x; y < tsz; y += blockDim.x) K[y] = -1;
    __syncthreads();
    int fail = 0;
    for (int32_t y = threadIdx.x; y < n; y += blockDim.x) {
      int32_t cur = src[y];
      uint32_t b = bucket(cur, c1);
      for (int it = 0;; ++it) {
        if (atomicCAS(&K[2 * b], -1, cur) == -1 || atomicCAS(&K[2 * b + 1], -1, cur) == -1) break;
        const uint32_t h1 = bucket(cur, c1);
        const uint32_t o = b == h1 ? bucket(cur, c2) : h1;
        if (atomicCAS(&K[2 * o], -1, cur) == -1 || atomicCAS(&K[2 * o + 1], -1, cur) == -1) break;
        if (it == kCuckooSteps) {
          fail = 1;
          break;
        }
        const int32_t old = atomicExch(&K[2 * o + cuckoo_slot(cur, it)], cur);
        if (old == -1) break;
        cur = old;
        b = o;
      }
    }
    if (!__syncthreads_or(fail)) {
      t.sorted = false;
      t.c1 = c1;
      t.c2 = c2;
      return t;
    }
  }
  for (uint32_t y = threadIdx.x; y < tsz; y += blockDim.x) K[y] = (int32_t)y < n ? src[y] : INT_MAX;
  int32_t top = 1;
  while (2 * top <= n) top *= 2;
  t.top = n > 0 ? top : 0;
  __syncthreads();
  return t;
}

// The long rows (listed by k_long_rows: N+ longer than a wave's window), one
// row at a time per workgroup: the four waves build one cuckoo table of
// N+(u) over the workgroup's LDS (TSW slots: up to 3 TSW / 4 ids, so no row
// of the graphs measured needs more than one window), then take the row's
// partners in groups of 64 / G lanes (group g of wave v takes partners
// 4 G j + G v + g), streaming each N+(p) in wave-coalesced vectors against
// the shared table, then flush the counters together.  A longer row is
// walked in windows of 3 TSW / 4 ids with every N+(p) cut to the window's
// range, as in the per-wave kernels.
constexpr int64_t kLongRun = 16384;
// Keys of the wave's 64 lanes sorted in descending order (bitonic network:
// lane l ends with the l-th largest key).
__device__ __forceinline__ uint32_t wave_sort_desc(uint32_t key, int lane) {
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint32_t o = (uint32_t)__shfl_xor((int)key, j, 64);
      const bool keep_max = ((lane & j) == 0) == ((lane & k) == 0);
      key = keep_max ? (key > o ? key : o) : (key < o ? key : o);
    }
  }
  return key;
}

// The long-row phase: the workgroup takes listed long rows from lctr until
// none is left, counting each with one table of TSW slots in K (and 16-bit
// counters in X) shared by its four waves; returns this wave's matches.
template <int TSW, int G, bool PV, typename OT>
__device__ __forceinline__ int64_t long_rows_phase(int32_t* K, int32_t* X, int64_t* bcast,
                                                   const int64_t* __restrict__ voff, const int64_t* __restrict__ orow,
                                                   const int32_t* __restrict__ oind, const int32_t* __restrict__ lrow,
                                                   int64_t nl, OT* __restrict__ out, unsigned long long* lctr,
                                                   bool by_in) {
  constexpr int NW = 4;
  constexpr int AWW = 3 * TSW / 4 / 32 * 32;  // ids per table: 2 roundup32(2 AWW / 3) <= TSW
  static_assert(TSW % 64 == 0 && 2 * (((2 * AWW + 2) / 3 + 31) / 32 * 32) <= TSW, "table size");
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int32_t sh = (int32_t)((reinterpret_cast<uintptr_t>(oind) >> 2) & 3);
  const int32_t* ind4 = oind - sh;
  int64_t tri = 0;
  for (;;) {
    if (threadIdx.x == 0) *bcast = (int64_t)atomicAdd(lctr, 1ull);
    __syncthreads();
    const int64_t i = uniform64(*bcast);
    __syncthreads();
    if (i >= nl) break;
    const int64_t u = lrow[i];
    const int64_t ou = orow[2 * u], oe = orow[2 * u + 1], lu = oe - ou;
    const int64_t tu = by_in ? oe : voff[u];
    const int64_t lt = (by_in ? voff[u + 1] : oe) - tu;
    if (lt <= 0) continue;
    int64_t cu = 0;  // this wave's matches of the row
    for (int64_t ws = 0; ws < lu; ws += AWW) {
      const int32_t n = (int32_t)(lu - ws < AWW ? lu - ws : AWW);
      const uint32_t tsz = table_slots(n, TSW);
      const HashTab tab = hash_build_wg(oind + ou + ws, n, K, tsz);
      const bool whole = n == lu;
      const int32_t amin = oind[ou + ws], amax = oind[ou + ws + n - 1];
      // partners in runs of kLongRun between flushes: a slot gains at most
      // one match per partner, so its 16-bit counter stays below 2 kLongRun
      for (int64_t r0 = 0; r0 < lt; r0 += kLongRun) {
        const int64_t r1 = r0 + kLongRun < lt ? r0 + kLongRun : lt;
        if (PV) {
          for (uint32_t y = threadIdx.x; y < tsz / 2; y += blockDim.x) X[y] = 0;
          __syncthreads();
        }
        if (whole) {
          constexpr int LG = 64 / G;
          const int gid = lane / LG, gl = lane % LG;
          int64_t cg = 0;
          for (int64_t j = r0 + (int64_t)wv * G; j < r1; j += NW * G) {
            const int64_t jj = j + gid;
            int32_t c = 0;
            if (jj < r1) {
              const int32_t w = oind[tu + jj];
              const int64_t ow = orow[2 * w], we = orow[2 * w + 1];
              c = hash_stream<LG, PV>(ind4, ow + sh, we + sh, gl, tab, X);
              c = group_sum32<LG>(c);  // G = 1: every lane is here (j is the same in all)
              if (PV && gl == 0 && c) out_add(out, w, 2 * c);
            }
            cg += c;
          }
          if (G == 1) {
            cu += uniform64(cg);  // every lane holds the same sums
          } else {
            int64_t rs = gl == 0 ? cg : 0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) rs += __shfl_xor(rs, o, 64);
            cu += uniform64(rs);
          }
        } else {
          for (int64_t j = r0 + wv; j < r1; j += NW) {
            const int32_t w = oind[tu + j];
            const int64_t ow = orow[2 * w], we = orow[2 * w + 1];
            int64_t g0 = ow + sh, g1 = we + sh;
            if (we > ow) {
              g0 = wave_lower_bound(ind4, g0, g1, amin, lane);
              g1 = wave_lower_bound(ind4, g0, g1, amax == INT_MAX ? amax : amax + 1, lane);
            }
            int32_t c = hash_stream<64, PV>(ind4, g0, g1, lane, tab, X);
            c = wave_sum32_dpp(c);
            if (c) {
              if (PV && lane == 0) out_add(out, w, 2 * c);
              cu += c;
            }
          }
        }
        __syncthreads();  // every match of the run counted
        if (PV) {
          for (uint32_t y = threadIdx.x; y < tsz; y += blockDim.x) {
            const int32_t x = x_read(X, (int32_t)y);
            if (x) out_add(out, K[y], x);
          }
          __syncthreads();  // X is cleared for the next run
        }
      }
      __syncthreads();  // K is rewritten by the next window or row
    }
    if (cu) {
      if (PV && lane == 0) out_add(out, u, 2 * cu);
      tri += cu;
    }
  }
  return tri;
}

// sorted: every window as a sorted array (binary-search lookups) instead of
// a cuckoo table.  Rows with more than lmax ids in N+ are skipped (their
// partners are left to the long-row phase).  Without WIN, lmax <= AW: every
// row taken fits one window, and the window walk is compiled out; the
// workgroup then first counts the listed long rows lrow[0 .. *nlong) (from
// lctr, its four waves sharing one table over the whole LDS), and only
// then takes pieces, so the largest tasks start first and run beside the
// pieces of other workgroups.
template <int TS, int AW, int G, int WPE, bool PV, typename OT, bool ONE, bool WIN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_count_hash(
    const int64_t* __restrict__ voff, const int64_t* __restrict__ orow, const int32_t* __restrict__ oind,
    const int32_t* __restrict__ prow, int64_t nv, OT* __restrict__ out, unsigned long long* total,
    unsigned long long* ctr, unsigned long long* pq, int side, int32_t P, int sorted, int64_t lmax,
    const int32_t* __restrict__ lrow,
    const unsigned long long* nlong, int64_t lcap, unsigned long long* lctr) {
  static_assert(TS % 64 == 0 && 2 * (((2 * AW + 2) / 3 + 31) / 32 * 32) <= TS, "table of whole wave strides");
  constexpr int NW = 4;
  constexpr int XW = TS / 2;  // words of 16-bit counters per wave
  __shared__ __align__(16) int32_t keys_all[NW * TS];
  __shared__ int32_t xcnt_all[PV ? NW * XW : 1];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int32_t* K = keys_all + wv * TS;
  int32_t* X = xcnt_all + (PV ? wv * XW : 0);
  const int32_t sh = (int32_t)((reinterpret_cast<uintptr_t>(oind) >> 2) & 3);
  const int32_t* ind4 = oind - sh;  // 16-byte aligned; position q of oind is position q + sh of ind4
  int64_t tri = 0;
  const int64_t M = voff[nv];
  const bool by_in = side == 2 ? uniform64((int64_t)ctr[2]) == 0 : side != 0;
  if (!WIN) {
    __shared__ int64_t bcast;
    int64_t nl = (int64_t)*nlong;
    if (nl > lcap) nl = lcap;
    constexpr int GL = G < 2 ? 1 : 2;
    tri += long_rows_phase<NW * TS, GL, PV, OT>(keys_all, xcnt_all, &bcast, voff, orow, oind, lrow, nl, out, lctr,
                                                by_in);
  }
  const int64_t npieces = (M + P - 1) / P;
  const int xcd = blockIdx.x & 7;
  int qk = 0;
  for (;;) {
    int64_t b = 0;
    if (ONE) {
      b = (int64_t)blockIdx.x * NW + wv;
    } else {
      b = next_task(pq, npieces, xcd, qk);
      if (b < 0) break;
    }
    const int64_t a0 = b * P;
    if (a0 >= M) break;
    const int64_t a1 = a0 + P < M ? a0 + P : M;
    int64_t u = prow[b];
    int64_t t_next = voff[u];
    for (;; ++u) {
      const int64_t t_lo = t_next;
      t_next = voff[u + 1];
      const int64_t ou = orow[2 * u], oe = orow[2 * u + 1], lu = oe - ou;
      const int64_t p0 = by_in ? oe : t_lo, p1 = by_in ? t_next : oe;
      const int64_t tu = p0 > a0 ? p0 : a0;
      const int64_t lt = (p1 < a1 ? p1 : a1) - tu;
      if (lu == 0 || lt <= 0 || lu > lmax) {
        if (t_next >= a1) break;
        continue;
      }
      int64_t cu = 0;
      for (int64_t ws = 0; ws < lu; ws += AW) {
        const int32_t n = (int32_t)(lu - ws < AW ? lu - ws : AW);
        // nb >= 2n/3 buckets of two slots (load <= 3/4), a power of two when it fits
        const uint32_t tsz = table_slots(n, TS);
        if (PV)
          for (uint32_t y = lane; y < tsz / 2; y += 64) X[y] = 0;
        const HashTab tab = hash_build(oind + ou + ws, n, K, tsz, sorted != 0, lane);
        const bool whole = !WIN || n == lu;
        const int32_t amin = oind[ou + ws], amax = oind[ou + ws + n - 1];
        if ((G > 1 || !WIN) && whole) {
          constexpr int LG = 64 / G;
          const int gid = lane / LG, gl = lane % LG;
          int64_t cg = 0;
          if (G > 1 && G <= 8) {
            // The row's partners 64 at a time, one per lane (its id and N+
            // bounds), ordered by list length, longest first: the G lists
            // streamed together then take about as many steps, where in
            // the CSR's order the longest of each G set the pace for all.
            // (With 16 or 32 partners per wave the lists are short, and
            // ordering them cost more than it saved: the 4096-graph batch
            // at 16 per wave went from 32.0 to 41.5 ms.)
            for (int64_t j0 = 0; j0 < lt; j0 += 64) {
              const int32_t nb = (int32_t)(lt - j0 < 64 ? lt - j0 : 64);
              if (nb <= G) {  // one set of G: the order changes nothing
                int32_t c = 0;
                if (gid < nb) {
                  const int32_t w = oind[tu + j0 + gid];
                  const int64_t ow = orow[2 * w], we = orow[2 * w + 1];
                  c = hash_stream<LG, PV>(ind4, ow + sh, we + sh, gl, tab, X);
                  c = group_sum32<LG>(c);
                  if (PV && gl == 0 && c) out_add(out, w, 2 * c);
                }
                cg += c;
                continue;
              }
              uint32_t key = 0u;  // lanes past nb sort last
              if (lane < nb) {
                const int32_t pw = oind[tu + j0 + lane];
                const int64_t len = orow[2 * pw + 1] - orow[2 * pw];
                const int64_t kl = len < 0 ? 1 : len < (int64_t(1) << 25) - 2 ? len + 1 : (int64_t(1) << 25) - 1;
                key = ((uint32_t)kl << 6) | (uint32_t)lane;  // distinct, above 0
              }
              // only the order is kept; each set reloads its partners' ids
              // and bounds (cached), which holds fewer registers than
              // carrying them through the loop
              const int order = (int)(wave_sort_desc(key, lane) & 63u);
              for (int32_t b = 0; b < nb; b += G) {
                const int32_t si = b + gid;
                const int src = __shfl(order, si & 63, 64);
                int32_t c = 0;
                if (si < nb) {
                  const int32_t w = oind[tu + j0 + src];
                  const int64_t ow = orow[2 * w], we = orow[2 * w + 1];
                  c = hash_stream<LG, PV>(ind4, ow + sh, we + sh, gl, tab, X);
                  c = group_sum32<LG>(c);
                  if (PV && gl == 0 && c) out_add(out, w, 2 * c);
                }
                cg += c;
              }
            }
          } else {
            for (int64_t j = 0; j < lt; j += G) {
              const int64_t jj = j + gid;
              int32_t c = 0;
              if (jj < lt) {
                const int32_t w = oind[tu + jj];
                const int64_t ow = orow[2 * w], we = orow[2 * w + 1];
                c = hash_stream<LG, PV>(ind4, ow + sh, we + sh, gl, tab, X);
                c = group_sum32<LG>(c);  // G = 1: every lane is here (j is the same in all)
                if (PV && gl == 0 && c) out_add(out, w, 2 * c);
              }
              cg += c;
            }
          }
          if (G == 1) {
            cu += uniform64(cg);  // every lane holds the same sums
          } else {
            int64_t rs = gl == 0 ? cg : 0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) rs += __shfl_xor(rs, o, 64);
            cu += uniform64(rs);
          }
        } else if (WIN) {
          for (int64_t j = 0; j < lt; ++j) {
            const int32_t w = oind[tu + j];
            const int64_t ow = orow[2 * w], we = orow[2 * w + 1];
            int64_t g0 = ow + sh, g1 = we + sh;
            if (we > ow && !whole) {
              g0 = wave_lower_bound(ind4, g0, g1, amin, lane);
              g1 = wave_lower_bound(ind4, g0, g1, amax == INT_MAX ? amax : amax + 1, lane);
            }
            int32_t c = hash_stream<64, PV>(ind4, g0, g1, lane, tab, X);
            c = wave_sum32_dpp(c);
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
          for (uint32_t y = lane; y < tsz; y += 64) {
            const int32_t x = x_read(X, (int32_t)y);
            if (x) out_add(out, K[y], x);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();  // K and X are rewritten by the next window
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      if (cu) {
        if (PV && lane == 0) out_add(out, u, 2 * cu);
        tri += cu;
      }
      if (t_next >= a1) break;
    }
    if (ONE) break;
  }
  if (lane == 0 && total && tri) atomicAdd(total, (unsigned long long)(6 * tri));
}

// Rows whose N+ is longer than a wave's window (lu > aw) into lrow[0 ..
// *cnt), one atomic per wave (at most cap are listed: cap >= nnz / aw + 1
// holds them all).  Their partners are left to the long-row phase by the
// per-wave count kernels.
__global__ void k_long_rows(const int64_t* __restrict__ orow, int64_t nv, int64_t aw, int32_t* __restrict__ lrow,
                            int64_t cap, unsigned long long* cnt) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < nv; base += step) {
    const int64_t u = base + threadIdx.x;
    const bool l = u < nv && orow[2 * u + 1] - orow[2 * u] > aw;
    const uint64_t m = __ballot(l);
    if (!m) continue;
    unsigned long long at = 0;
    if (lane == 0) at = atomicAdd(cnt, (unsigned long long)__popcll(m));
    const int64_t i = (int64_t)__shfl(at, 0, 64) + __popcll(m & lt);
    if (l && i < cap) lrow[i] = (int32_t)u;
  }
}

// The listed long rows (k_long_rows) on their own, after a per-wave count
// kernel that left them (k_count_orient's rows over its window).  Same
// credit rule and sides as k_count_hash; ctr[2] is the listed hub-row count
// that side 2 reads.
template <int TSW, int G, bool PV, typename OT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_count_long(
    const int64_t* __restrict__ voff, const int64_t* __restrict__ orow, const int32_t* __restrict__ oind,
    const int32_t* __restrict__ lrow, const unsigned long long* nlong, int64_t cap, OT* __restrict__ out,
    unsigned long long* total, const unsigned long long* ctr, unsigned long long* lctr, int side) {
  __shared__ __align__(16) int32_t K[TSW];
  __shared__ int32_t X[PV ? TSW / 2 : 1];
  __shared__ int64_t bcast;
  int64_t nl = (int64_t)*nlong;
  if (nl > cap) nl = cap;
  const bool by_in = side == 2 ? uniform64((int64_t)ctr[2]) == 0 : side != 0;
  const int64_t tri = long_rows_phase<TSW, G, PV, OT>(K, X, &bcast, voff, orow, oind, lrow, nl, out, lctr, by_in);
  if ((threadIdx.x & 63) == 0 && total && tri) atomicAdd(total, (unsigned long long)(6 * tri));
}
