/*
 * tc.h — C-ABI of the MI355X (gfx950) per-vertex triangle counter.
 *
 * Drop-in for the reference GPU path:
 *   kernelCall<T>(blocks, bs, nv, d_off, d_ind, out, tpi, groups, shifter)
 *       src/clusteringCount.cu:415-467, declared include/clusteringCount.cuh:99-104
 *   count_all_trianglesGPU<T, bs>  (the __global__ it launches)
 *       src/clusteringCount.cu:314-359
 *   thrust::reduce(out) total in the caller
 *       src/gpurun.cu:40-41
 *
 * Contract kept from the reference: CSR in (device int32 row offsets
 * off[nv+1] and sorted neighbour ids ind[nnz]), per-vertex counts out
 * (out[v] = sum over neighbours w of |N(v) ∩ N(w)| = 2·T(v), the same value
 * the reference kernel and CPU path produce), every out[v] written, launch
 * asynchronous on a stream.  Changes: the total is int64 (the reference
 * sums in T and overflows at scale), errors are reported as return codes,
 * and the input must be a simple undirected graph — rows strictly
 * increasing, no self-loops, every arc (u,w) matched by (w,u) — which
 * tc_validate_i32 checks on the device.  Per-vertex int32 is enough for every
 * configuration measured (largest 2·T(v) at R-MAT scale 24 is ~2.0e8).
 *
 * Streams are passed as void* (a hipStream_t; NULL = the default stream).
 * No torch or HIP types appear in the signatures.
 */
#ifndef TC_H
#define TC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tc_ctx tc_ctx;

/* Count configuration.  Zero fields take the context's default.  The
 * default counts on the degree-oriented scheme (`oriented`, `row_groups`);
 * the other fields size the edge-list scheme's kernels (oriented = -1, or
 * algo != 0).
 *   wg_size        threads per workgroup (256 or 512)
 *   items          merge-path steps per lane per tile (8, 16 or 32)
 *   ring           LDS ring length per adjacency list, elements (power of 2);
 *                  a tile covers at most ring-2 steps of one long edge
 *   small_max      edges with at most this much work (|N(u)|+|N(w)|) are
 *                  packed several to a tile; longer ones stream through the
 *                  LDS rings
 *   ranges_per_wg  level-1 split: the edge-work range [0, W) is cut into
 *                  ranges_per_wg * (resident workgroups) equal ranges that a
 *                  persistent grid pulls from per-XCD queues
 *   algo           0 = streamed kernel with every long edge on one wave
 *                  (the shorter list walked in LDS windows) plus packed
 *                  small-edge tiles (default); 3 = streamed kernel whose
 *                  longest edges go to workgroup streams and LDS rings;
 *                  1 = one lane per edge (simple, for cross-checks);
 *                  2 = Merge Path through LDS rings for every long edge
 *   chunk          algo 0/3: elements per streamed chunk (16)
 *   depth          algo 0/3: streamed chunks in flight per lane (2 or 4)
 *   search_ratio   algo 0/3, edge-list scheme: an edge whose longer list B
 *                  has at least 512 and at least search_ratio times as many
 *                  elements as the shorter list A is counted by
 *                  binary-searching every element of A in B instead of
 *                  streaming B (default 64; negative: never)
 *   probe_ratio    algo 0/3, edge-list scheme: a streamed edge whose B is at
 *                  least probe_ratio times longer than A is merged from A's
 *                  side, each A value tested against a 16-element chunk of B
 *                  in registers (default 4; negative: never)
 *   oriented       algo 0 (tc_count_i32, tc_batch_count_i32): 1 =
 *                  intersect the degree-ordered orientation, built in place
 *                  of the input rows, each triangle found once and its three
 *                  vertices credited; 2 = the same, but a row takes its
 *                  in-neighbours as partners (streams the lower end's row of
 *                  each arc); 3 (default) = automatic: 2 for batches, for
 *                  graphs below 2^24 arcs and for the table lookup below
 *                  2^30 arcs; otherwise 2 unless the graph has rows of
 *                  more than 4096 arcs (then 1, decided on the device);
 *                  tc_oriented_plan reports the choice; -1 = every
 *                  undirected edge's full lists
 *                  (the reference's scheme, the edge-list kernels).  Same
 *                  counts either way; other algos always use the full lists.
 *                  tc_last_edge_work reports (0, 0) for the oriented scheme,
 *                  which builds no edge list.
 *   row_groups     oriented scheme: partners of a row a wave streams at once,
 *                  each by 64/row_groups lanes (1, 2, 4, 8, 16 or 32;
 *                  default -1 = chosen from the size and mean degree: 1
 *                  for single graphs from 2^30 arcs up; single graphs with
 *                  the table lookup 8 with in-neighbour partners, else 1
 *                  from 2^28 arcs and 2 below; with the walked array 2 from
 *                  2^28 arcs; else 32 below mean degree 10, 16 below 20, 8
 *                  above)
 *   lookup         oriented scheme: how a partner's list meets the row's
 *                  window: 1 = the window as a sorted LDS array that each
 *                  lane walks (k_count_orient); 2 = an LDS cuckoo table
 *                  that every element of wave-coalesced 16-byte reads
 *                  probes twice (k_count_hash); 3 = the same reads, each
 *                  element binary-searched in the sorted window;
 *                  0 (default) = automatic: 2 for single graphs from
 *                  3 * 2^25 arcs up, and from 2^24 arcs with a mean degree
 *                  of at least 16; else 1
 *   long_rows      oriented scheme: 2 = rows whose N+ is longer than one
 *                  wave's LDS window (`window` ids for lookup 1, 624 for 2
 *                  and 3) are counted by whole workgroups sharing one table of
 *                  up to 2496 ids (k_count_long) instead of in windows by
 *                  single waves; 1 = windows; 0 (default) = automatic: 2
 *                  for single graphs from 2^26 arcs up, else 1
 *   window         oriented scheme, lookup 1: ids of a row's N+ per wave LDS
 *                  window: 851 (default), 425 or 1701; the int64 path and
 *                  the total-only count always use 851.  Rows longer than
 *                  it are walked in windows, or go to k_count_long with
 *                  long_rows
 *   orient_wg      oriented scheme, lookup 1: threads per workgroup of the
 *                  count kernel (128 or 256, default 256)
 *   piece          oriented scheme: arc positions per work piece, the unit
 *                  the count kernel's waves take from a counter (256..16383;
 *                  default -1 = 256 for single graphs (512 on the table
 *                  lookup with in-neighbour partners from 2^26 arcs),
 *                  1024 for batches; a
 *                  graph with fewer pieces than waves takes smaller ones) */
typedef struct tc_cfg {
  int32_t wg_size;
  int32_t items;
  int32_t ring;
  int32_t small_max;
  int32_t ranges_per_wg;
  int32_t algo;
  int32_t chunk;
  int32_t depth;
  int32_t search_ratio;
  int32_t probe_ratio;
  int32_t oriented;
  int32_t row_groups;
  int32_t lookup;
  int32_t long_rows;
  int32_t window;
  int32_t orient_wg;
  int32_t piece;
} tc_cfg;

enum {
  TC_OK = 0,
  TC_EINVAL = -1,    /* bad argument or unsupported launch configuration */
  TC_ERANGE = -2,    /* sizes beyond the int32 path (nnz >= 2^31) */
  TC_EHIP = -3,      /* a HIP runtime call failed (tc_last_error has the text) */
  TC_ENOMEM = -4,    /* device scratch could not be allocated */
  TC_EGRAPH = -5,    /* tc_validate_i32, tc_server_submit/_wait: input is not a simple undirected CSR */
  TC_ECANCELED = -6  /* tc_server_wait: the request was canceled (tc_server_cancel) */
};

/* Context: owns device scratch (edge list, work prefix, scan partials), the
 * tuned configuration table and timing events.  One context per host
 * thread.  device < 0 uses the current device. */
int tc_ctx_create(int device, tc_ctx** out);
int tc_ctx_destroy(tc_ctx* ctx);
const char* tc_last_error(const tc_ctx* ctx);
/* Scratch bytes tc_count_i32 needs for a graph of this size (so a caller
 * can reserve up front; tc_ctx_reserve allocates it). */
size_t tc_scratch_bytes(int32_t nv, int64_t nnz);
int tc_ctx_reserve(tc_ctx* ctx, size_t bytes);
/* Default configuration used when cfg == NULL. */
int tc_ctx_set_default_cfg(tc_ctx* ctx, const tc_cfg* cfg);
int tc_ctx_get_default_cfg(const tc_ctx* ctx, tc_cfg* cfg);
/* When enabled, each count records HIP events around its phases on the
 * launch stream; tc_last_timing returns (after the stream is synchronised)
 * ms[0] = prep (the orientation, or the edge list), ms[1] = the count
 * kernels, ms[2] = finalisation, ms[3] = whole call. */
int tc_ctx_set_timing(tc_ctx* ctx, int enable);
int tc_last_timing(tc_ctx* ctx, float ms[4]);

/* Per-vertex triangle counts of a simple undirected graph.
 *   d_off[nv+1], d_ind[nnz]  device CSR (int32), rows sorted ascending
 *   d_out[nv]                device int32, fully overwritten: 2·T(v)
 *   d_total                  device int64 or NULL: sum of d_out (= 6·T)
 * Replaces kernelCall<int32_t> + thrust::reduce (gpurun.cu:36-41). */
int tc_count_i32(tc_ctx* ctx, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                 int32_t* d_out, int64_t* d_total, const tc_cfg* cfg, void* stream);

/* Per-vertex count that allocates nothing: tc_count_i32 on the oriented
 * scheme (the default), with its scratch in a fixed device array of
 * TC_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs on
 * until the process ends (its address is resolved once per device by
 * tc_ctx_create).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree or allocating library routine.  Same arguments, output and
 * asynchrony as tc_count_i32; TC_ERANGE when the graph's scratch
 * (about 28 (nv + 1) + 4 nnz + 0.3 MB bytes) exceeds the array, TC_EINVAL
 * when cfg asks for a scheme other than the oriented one.  Every call on
 * one GPU shares that GPU's array and leaves it holding its own working
 * data (each call clears what must start at zero): calls that overlap on
 * one GPU, from different streams or host threads, race and must be
 * ordered by the caller. */
#define TC_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int tc_count_noalloc_i32(tc_ctx* ctx, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                         int32_t* d_out, int64_t* d_total, const tc_cfg* cfg, void* stream);

/* Shape-compatible stand-in for kernelCall<int32_t>
 * (src/clusteringCount.cu:415-467): same argument list and meaning.
 * numberThreads must be one of the reference block sizes
 * {32,64,96,128,192,256,512,1024} and threads_per_block (threads per
 * intersection) one of {1,2,4,8,16,32} with number_blocks == numberThreads /
 * threads_per_block and shifter == log2(threads_per_block), as
 * param.cu:225-247 requires; otherwise TC_EINVAL (the reference silently
 * launches nothing).  The legacy launch shape is validated, then the
 * default count (the degree-oriented scheme) runs on the default stream of
 * the current device with a library-owned context.  outPutTriangles is
 * fully written. */
int tc_kernelCall_i32(unsigned numberBlocks, unsigned numberThreads, int32_t nv, const int32_t* d_off,
                      const int32_t* d_ind, int32_t* outPutTriangles, int32_t threads_per_block,
                      int32_t number_blocks, int32_t shifter);

/* The int64 CSR (the reference's kernelCall<int64_t>, instantiated at
 * src/clusteringCount.cu:518-524 and run by main, src/main.cu:36):
 *   d_off[nv+1], d_ind[nnz]  device int64 CSR, rows sorted ascending; nnz
 *                            may exceed 2^31, vertex ids must be below 2^31
 *   d_out[nv]                device int64, fully overwritten: 2·T(v)
 *   d_total                  device int64 or NULL: sum of d_out (= 6·T)
 * Runs the oriented scheme (TC_EINVAL if cfg asks for another); the
 * offsets are read in place, the ids once into the orientation. */
int tc_count_i64(tc_ctx* ctx, int64_t nv, int64_t nnz, const int64_t* d_off, const int64_t* d_ind,
                 int64_t* d_out, int64_t* d_total, const tc_cfg* cfg, void* stream);

/* Shape-compatible stand-in for kernelCall<int64_t>: the same argument
 * rules as tc_kernelCall_i32, on tc_count_i64. */
int tc_kernelCall_i64(unsigned numberBlocks, unsigned numberThreads, int64_t nv, const int64_t* d_off,
                      const int64_t* d_ind, int64_t* outPutTriangles, int64_t threads_per_block,
                      int64_t number_blocks, int64_t shifter);

/* Vertex-list count: the reference's vertex-list overload of kernelCall<T>
 * (src/clusteringCount.cu:361-413, kernel :266-312; CPU form
 * src/clusteringCount.cpp:125-155).  d_list[nlist] holds vertex ids of the
 * graph (any order, repeats allowed); d_out[i] receives, for v = d_list[i],
 * the sum over arcs v -> w with w != v, d(v) >= 2 and d(w) >= 2 of
 * |N(v) ∩ N(w)| (2·T(v) on a simple graph), and d_total (nullable) the sum
 * of d_out.  An entry outside [0, nv) gets d_out[i] = -1 and adds nothing.
 * The reference overload takes neither the graph's vertex count nor its arc
 * count; this entry point needs both (nv, nnz) for its scratch.  Partition
 * lists of a .part.K file (tch_read_partition in tc_host.h) are the
 * intended input. */
int tc_count_vertices_i32(tc_ctx* ctx, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                          int32_t nlist, const int32_t* d_list, int32_t* d_out, int64_t* d_total, const tc_cfg* cfg,
                          void* stream);

/* Vertex-list count with shadow vertices: the reference's intersection
 * with a lim argument (intersectCount, src/clusteringCount.cu:150-177, and
 * count_triangles, :214-246; declared at include/clusteringCount.cuh:56,73),
 * where a common neighbour x counts 1 + [x > lim] instead of 1.  d_out[i]
 * receives, for v = d_list[i], the sum over the arcs v -> w of
 * tc_count_vertices_i32 of the sum over x in N(v) ∩ N(w) of 1 + [x > lim];
 * lim = -1 doubles every value, lim >= nv - 1 gives tc_count_vertices_i32's.
 * The reference counts a match met at a thread's split of the merge path
 * once (fixThreadWorkEdges, :108-121), so its value depends on its launch
 * shape there; this entry point applies the rule to every match.  The
 * reference's overload has no caller in its sources.  nnz must stay below
 * 2^30 (TC_ERANGE): every listed arc may become two edges. */
int tc_count_vertices_lim_i32(tc_ctx* ctx, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                              int32_t nlist, const int32_t* d_list, int32_t lim, int32_t* d_out, int64_t* d_total,
                              const tc_cfg* cfg, void* stream);

/* FP64 local clustering coefficient lcc[v] = out[v] / (d(v)·(d(v)-1)),
 * 0 when d(v) < 2; one IEEE division of exactly converted integers. */
int tc_lcc_f64(tc_ctx* ctx, int32_t nv, const int32_t* d_off, const int32_t* d_out, double* d_lcc,
               void* stream);

/* Device-side check of the simple-graph contract.  *h_status (host) gets 0
 * when valid, else TC_EGRAPH; the call synchronises the stream. */
int tc_validate_i32(tc_ctx* ctx, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                    int* h_status, void* stream);

/* Batched count: G graphs packed as one concatenated CSR (layout of
 * tch_batch in tc_host.h):
 *   d_tab[2*(G+1)]  int64 {v_base, arc_base} per graph, entry G = totals
 *   d_off_cat       int32, graph g's local offsets at d_off_cat[v_base+g ..]
 *   d_ind_cat       int32 local ids, graph g's at d_ind_cat[arc_base ..]
 *   d_out_cat       int32[total vertices], graph g's counts at [v_base ..]
 *   d_totals        int64[G] or NULL: per-graph sum of counts
 * h_nv_total / h_nnz_total are tab[2G] / tab[2G+1] (host copies, used to
 * size scratch and launches without a device read).  One launch sequence
 * for all graphs; results equal counting each graph alone. */
int tc_batch_count_i32(tc_ctx* ctx, int32_t G, const int64_t* d_tab, int64_t h_nv_total, int64_t h_nnz_total,
                       const int32_t* d_off_cat, const int32_t* d_ind_cat, int32_t* d_out_cat, int64_t* d_totals,
                       const tc_cfg* cfg, void* stream);
size_t tc_batch_scratch_bytes(int32_t G, int64_t nv_total, int64_t nnz_total);

/* The same G graphs counted one tc_count_i32 call per graph, in order, on
 * one stream (the plain path, one launch sequence per graph): the baseline
 * the batched launch is compared against.  h_tab is the host copy of
 * d_tab's layout; d_totals (nullable) gets each graph's total. */
int tc_batch_count_pergraph_i32(tc_ctx* ctx, int32_t G, const int64_t* h_tab, const int32_t* d_off_cat,
                                const int32_t* d_ind_cat, int32_t* d_out_cat, int64_t* d_totals, const tc_cfg* cfg,
                                void* stream);

/* ---------------------------------------------------------------- graphs
 * One count call recorded as a HIP graph, replayable on any stream with no
 * host-side work and no allocation: scratch is reserved before recording,
 * every address is fixed, and buffers are cleared by fill kernels (no
 * memset nodes).  Input contents may change between replays; sizes and
 * addresses may not.  No reference counterpart (the reference launches
 * from the host each time, clusteringCount.cu:425-463). */
typedef struct tc_graph tc_graph;
int tc_count_capture_i32(tc_ctx* ctx, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                         int32_t* d_out, int64_t* d_total, const tc_cfg* cfg, tc_graph** out);
int tc_batch_capture_i32(tc_ctx* ctx, int32_t G, const int64_t* d_tab, int64_t h_nv_total, int64_t h_nnz_total,
                         const int32_t* d_off_cat, const int32_t* d_ind_cat, int32_t* d_out_cat, int64_t* d_totals,
                         const tc_cfg* cfg, tc_graph** out);
int tc_graph_launch(tc_graph* g, void* stream);
int tc_graph_destroy(tc_graph* g);

/* ---------------------------------------------------------------- autotuner
 * The counterpart of the reference's sweep of 8 block sizes x 6 threads
 * per intersection (src/gpurun.cu:151-172).  Candidates, in five stages,
 * each around the fastest identical candidate so far: (1) the oriented
 * scheme, oriented in {1, 2} x row_groups in {1, 2, 4, 8, 16, 32} (walked
 * array); (2) window in {425, 851, 1701} x orient_wg in {128, 256};
 * (3) piece in {256, 512, 1024, 2048, 4096} (the partition granularity);
 * (4) lookup in {1, 2} x long_rows in {1, 2}; (5) the edge scheme
 * (oriented = -1) at every compiled streamed kernel variant (wg_size,
 * items, ring; algo 0 and 3) times small_max in {64, 256, 1024} (at most
 * wg_size*items) times ranges_per_wg in {8, 32, 128}.  Each candidate
 * runs once, its per-vertex output is
 * compared with the first candidate's (identical must hold: the counts are
 * exact integers), then it is timed over reps launches with HIP events.
 * The fastest identical candidate is returned in *best and, when
 * set_default is nonzero, becomes the context default.  Up to n_results
 * entries are written to results (nullable).  Returns TC_EGRAPH if any
 * candidate's output differed. */
typedef struct tc_tune_result {
  tc_cfg cfg;
  float ms;
  int32_t identical;
} tc_tune_result;
int tc_autotune_i32(tc_ctx* ctx, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                    int32_t* d_out, int reps, int set_default, tc_tune_result* results, int n_results,
                    int* n_written, tc_cfg* best);
int tc_autotune_batch_i32(tc_ctx* ctx, int32_t G, const int64_t* d_tab, int64_t h_nv_total, int64_t h_nnz_total,
                          const int32_t* d_off_cat, const int32_t* d_ind_cat, int32_t* d_out_cat, int reps,
                          int set_default, tc_tune_result* results, int n_results, int* n_written, tc_cfg* best);

/* Global-count mode (README.md:6-10): only the int64 total 6·T, via a
 * degree-ordered orientation so each triangle is intersected once. */
int tc_count_total_i32(tc_ctx* ctx, int32_t nv, int64_t nnz, const int32_t* d_off, const int32_t* d_ind,
                       int64_t* d_total, void* stream);
/* The same for an int64 CSR (as tc_count_i64: any nnz, ids below 2^31). */
int tc_count_total_i64(tc_ctx* ctx, int64_t nv, int64_t nnz, const int64_t* d_off, const int64_t* d_ind,
                       int64_t* d_total, void* stream);

/* The oriented count's launch plan for a count of nv vertices and nnz arcs
 * (a batch's totals when batch != 0) under cfg (NULL: the context's
 * default), with every automatic choice resolved: plan[0] = lookup (0 the
 * walked array, k_count_orient; 1 the LDS table, k_count_hash; 2 the table
 * kernel with sorted windows), plan[1] = partner side (0 out-neighbours, 1
 * in-neighbours, 2 in-neighbours unless the input has rows of more than
 * 4096 arcs, decided on the device), plan[2] = partners per wave,
 * plan[3] = 1 when rows longer than a wave's table are counted by
 * workgroups.  The edge-list scheme gives {-1, -1, 0, 0}.  Launches
 * nothing.  No reference counterpart (the reference has one kernel shape
 * per launch). */
int tc_oriented_plan(tc_ctx* ctx, int64_t nv, int64_t nnz, int32_t batch, const tc_cfg* cfg, int32_t plan[4]);

/* Edge count E and total merge work W = sum over intersected edges of
 * |N(u)| + |N(w)| of the last edge-list-scheme count on this context
 * (synchronises the context's device); (0, 0) after an oriented count,
 * which builds no edge list.  out[0] = E, out[1] = W. */
int tc_last_edge_work(tc_ctx* ctx, int64_t out[2]);

/* Plain device copy d_dst[i] = d_src[i] of n int32 values (one read and one
 * write per element, 16-byte vector accesses).  No reference counterpart:
 * it calibrates hardware byte counters (rocprofv3 FETCH_SIZE / WRITE_SIZE)
 * against a transfer of known size. */
int tc_copy_i32(tc_ctx* ctx, int32_t* d_dst, const int32_t* d_src, int64_t n, void* stream);

/* ---------------------------------------------------------------- serving
 * A request loop that counts submitted graphs in batches as they arrive.  A
 * worker thread forms batches first in first out: the next batch is the
 * longest prefix of the queue (at most max_batch_graphs requests) whose
 * block fits the largest free block of a device arena of hbm_budget bytes.
 * The block holds the batch in tc_batch_count_i32's packed layout (table,
 * offsets, ids), its counts and totals, and the scratch of its count, and
 * stays allocated until the batch completes: hbm_budget bounds all the
 * device memory of the requests in flight (the server allocates nothing on
 * the device after tc_server_create).  Each request is copied once, at
 * submit, into a pinned host slab; a batch's slabs are uploaded by DMA
 * straight into its block's layout, checked, counted and copied back on one
 * of `slots` launch slots (each with its own stream), so batches overlap on
 * the device while later requests wait for memory.  tc_server_destroy must
 * not race with other calls on the same server.  Counts are those of
 * tc_count_i32 on each graph alone.  No reference counterpart (the
 * reference counts one graph per process, gpurun.cu:2-56). */
typedef struct tc_server tc_server;
typedef struct tc_server_cfg {
  int32_t device;
  int32_t slots;            /* batches in flight, 1..8 */
  int32_t max_batch_graphs; /* most graphs per launch */
  int32_t reserved;
  int64_t hbm_budget;       /* bytes of the arena holding the batches in flight */
} tc_server_cfg;
typedef struct tc_server_stats {
  int64_t submitted, completed, rejected, batches;
  int64_t bytes_in_use, peak_bytes, waiting, in_flight;
  int64_t canceled;
} tc_server_stats;
int tc_server_create(const tc_server_cfg* cfg, tc_server** out);
/* Finishes all submitted work, then frees everything. */
int tc_server_destroy(tc_server* s);
/* Checks the host CSR's offsets (h_off[0] = 0, never decreasing, h_off[nv]
 * = nnz), copies it into a slab of the server's pinned host pool (hbm_budget
 * bytes; while the requests queued or in flight fill it, the call waits for
 * the worker to retire some) and queues it; returns a request id >= 0, or a negative
 * TC_E* code (TC_EGRAPH for offsets that fail the check, TC_ENOMEM if the
 * request can never fit the budget).  The rest of the simple-graph contract
 * (ids in [0, nv), rows strictly increasing, no self-loops, every arc with
 * its reverse) is checked on the device before the request's batch is
 * counted: a request that breaks it is set aside (the count reads none of
 * its rows and no other request's results change) and tc_server_wait
 * returns TC_EGRAPH for it. */
int64_t tc_server_submit(tc_server* s, int32_t nv, int64_t nnz, const int32_t* h_off, const int32_t* h_ind);
/* Waits up to timeout_ms (< 0: no limit) for request id; on completion
 * copies its nv counts to h_out (nullable), the total to h_total
 * (nullable) and the submit-to-completion time to latency_ms (nullable),
 * and forgets the request.  Returns TC_OK, 1 on timeout, TC_ECANCELED for
 * a canceled request, or a TC_E* code. */
int tc_server_wait(tc_server* s, int64_t id, int32_t* h_out, int64_t* h_total, int32_t timeout_ms,
                   double* latency_ms);
/* Cancels request id if it is still waiting for admission: it leaves the
 * queue and tc_server_wait returns TC_ECANCELED for it.  Returns TC_OK, 1
 * when the request is already admitted or done (it then completes as
 * usual), or TC_EINVAL for an unknown id. */
int tc_server_cancel(tc_server* s, int64_t id);
int tc_server_get_stats(tc_server* s, tc_server_stats* st);

/* Number of compute units of the context's device (for sizing). */
int tc_device_cus(const tc_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
