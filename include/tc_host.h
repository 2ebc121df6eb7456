/*
 * tc_host.h — C-ABI of the host-side graph driver (CPU only, no HIP).
 *
 * This is the part of the host driver that feeds the GPU counter: seeded
 * synthetic graph generators, the simple-graph CSR builder, and the text
 * formats the reference reads and writes.  It replaces:
 *   - readGraph<T>(path, off, ind, nv, nnz)      src/graphRead.cpp:2-26
 *   - readGraph (METIS adjacency -> CSR)         src/csrdatacreate.cpp:3-49
 *   - writeGraph (CSR -> .info text)             src/csrdatacreate.cpp:86-114
 * and adds what the reference lacks: generators (ER, R-MAT, 2D geometric),
 * self-loop removal, de-duplication, symmetrisation, a simple-graph check and
 * a packer that concatenates many graphs into one batch.
 *
 * Every function is thread-safe except that tch_last_error() is per thread.
 * Returned handles own their memory; free them with tch_graph_free /
 * tch_batch_free.
 */
#ifndef TC_HOST_H
#define TC_HOST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A CSR graph: off[n+1] (int32 arc offsets), ind[nnz] (int32 vertex ids). */
typedef struct tch_graph tch_graph;

/* Error text of the last failing call on this thread ("" if none). */
const char* tch_last_error(void);

int32_t        tch_graph_n(const tch_graph* g);
int64_t        tch_graph_nnz(const tch_graph* g);
const int32_t* tch_graph_off(const tch_graph* g);
const int32_t* tch_graph_ind(const tch_graph* g);
void           tch_graph_free(tch_graph* g);

/* Simple undirected graph from an edge list: self-loops dropped, both arc
 * directions added, rows sorted, duplicates removed.  Endpoints must lie in
 * [0, n).  NULL on error. */
tch_graph* tch_from_edges(int32_t n, int64_t m, const int32_t* src, const int32_t* dst);

/* CSR copied verbatim from caller arrays (no normalisation). */
tch_graph* tch_from_csr(int32_t n, int64_t nnz, const int32_t* off, const int32_t* ind);

/* Normalise any CSR into a simple symmetric one (same rules as
 * tch_from_edges applied to its arcs). */
tch_graph* tch_normalize(const tch_graph* g);

/* Generators.  All are counter-based (SplitMix64 stream indexed by draw
 * number) so the result does not depend on thread count or machine.
 *  ER:   m uniform (u,v) draws on n vertices.
 *  RMAT: m draws on 2^scale vertices with quadrant probabilities a,b,c
 *        (d = 1-a-b-c) quantised to 32-bit thresholds; endpoints >= n_keep
 *        are dropped (n_keep <= 0 keeps all) and the vertex count is
 *        n_keep (or 2^scale).
 *  RGG:  n points uniform in [0,1)^2, edge iff squared distance < r^2 with
 *        r^2 = mean_deg / (pi * n), evaluated in IEEE double without FMA. */
tch_graph* tch_gen_er(int32_t n, int64_t m, uint64_t seed);
tch_graph* tch_gen_rmat(int32_t scale, int64_t m, double a, double b, double c,
                        uint64_t seed, int32_t n_keep);
tch_graph* tch_gen_rgg(int32_t n, double mean_deg, uint64_t seed);

/* Check the simple-graph contract the GPU counter relies on.  Returns 0 if
 * off is monotone from 0 to nnz, every row is strictly increasing, no
 * self-loops, every id is in range and every arc has its reverse;
 * otherwise a negative TCH_E* code (tch_last_error names the first
 * offending vertex). */
int tch_check_simple(int32_t n, int64_t nnz, const int32_t* off, const int32_t* ind);

/* Reference text formats.
 * .info: "nv+1 nnz" header, nv+1 offsets, nnz 0-based ids (graphRead.cpp:2-26).
 * tch_read_info appends ".info" to path_no_ext like the reference does.
 * tch_write_info writes the exact layout of csrdatacreate.cpp:86-114.
 * METIS: optional '%' comment lines, header "n m", then n lines of 1-based
 * neighbour ids; rows are sorted, nothing else is changed (csrdatacreate.cpp:3-49). */
tch_graph* tch_read_info(const char* path_no_ext);
int        tch_write_info(const tch_graph* g, const char* path);
tch_graph* tch_read_metis(const char* path);

/* Partition file (graphRead.cpp:28-45): path_no_ext + ".part.K" holds a
 * count n and then n integers, whitespace separated.  *vals receives them
 * (malloc'd; free with tch_free) and *n their number.  With a METIS
 * partition vector (one partition id per vertex, the count first) the
 * members of partition p are tch_partition_members(n, vals, p, ...): the
 * vertices v with part[v] == p in ascending order (malloc'd). */
int  tch_read_partition(const char* path_no_ext, int32_t K, int32_t* n, int32_t** vals);
int  tch_partition_members(int32_t n, const int32_t* part, int32_t p, int32_t* count, int32_t** ids);
void tch_free(void* p);

/* Binary CSR cache: a 32-byte header ("TCCSR001", n, nnz, reserved) then
 * off and ind as little-endian int32. */
int        tch_write_bin(const tch_graph* g, const char* path);
tch_graph* tch_read_bin(const char* path);

/* ---------------------------------------------------------------- int64 CSR
 * The reference binary counts with T = int64 (src/main.cu:36 runs
 * run<int64_t>; readGraph<int64_t>, src/graphRead.cpp:2-26, reads .info
 * into int64 arrays).  tch_graph64 holds off[n+1] and ind[nnz] as int64 with
 * no cap on nnz; n and every id stay below 2^31 (tc_count_i64's limit).
 *   tch_read_info64   the .info text (appends ".info"), any nnz
 *   tch_read_bin64    the int64 cache ("TCCSR064", n, nnz, reserved, then
 *                     int64 off and ind) or the int32 cache, widened
 *   tch_write_bin64   writes the int64 cache
 *   tch_widen         an int32 graph as int64
 *   tch_check_simple64  tch_check_simple for the int64 layout
 *   tch_peek_sizes    n and nnz from the header of <path>.info, or of path
 *                     itself when it ends in ".bin", without reading the rest */
typedef struct tch_graph64 tch_graph64;
int64_t        tch_graph64_n(const tch_graph64* g);
int64_t        tch_graph64_nnz(const tch_graph64* g);
const int64_t* tch_graph64_off(const tch_graph64* g);
const int64_t* tch_graph64_ind(const tch_graph64* g);
void           tch_graph64_free(tch_graph64* g);
tch_graph64*   tch_read_info64(const char* path_no_ext);
tch_graph64*   tch_read_bin64(const char* path);
int            tch_write_bin64(const tch_graph64* g, const char* path);
tch_graph64*   tch_widen(const tch_graph* g);
int            tch_check_simple64(int64_t n, int64_t nnz, const int64_t* off, const int64_t* ind);
int            tch_peek_sizes(const char* path, int64_t* nv, int64_t* nnz);

/* ---------------------------------------------------------------- batches
 * G graphs packed into one concatenated CSR:
 *   tab[2*(G+1)]  int64: tab[2g] = v_base (sum of earlier n), tab[2g+1] =
 *                 arc_base (sum of earlier nnz); entry G holds the totals.
 *   off_cat       int32, sum(n_g + 1) entries: graph g's local offsets
 *                 (0 .. nnz_g) start at off_cat[v_base + g].
 *   ind_cat       int32, sum(nnz_g): graph g's local vertex ids start at
 *                 ind_cat[arc_base].
 * kind[g]: 0 = ER, 1 = R-MAT, 2 = 2D geometric. */
typedef struct tch_batch tch_batch;

/* The mixed batch: n_g = floor(exp(U(ln n_lo, ln n_hi))), kind uniform over
 * {ER mean degree 8 (m = 4n), R-MAT a=.57 b=.19 c=.19 with edge factor
 * U{ef_lo..ef_hi} and scale ceil(log2 n) dropping ids >= n, RGG mean
 * degree 6}; graph g uses seed SplitMix64(seed ^ g). kinds_mask selects
 * which kinds may be drawn (bit k = kind k). */
tch_batch* tch_gen_batch(int32_t G, int32_t n_lo, int32_t n_hi, int32_t ef_lo,
                         int32_t ef_hi, uint32_t kinds_mask, uint64_t seed);
/* Pack existing graphs (copied). */
tch_batch* tch_pack(int32_t G, const tch_graph* const* graphs);
int32_t        tch_batch_count(const tch_batch* b);
const int64_t* tch_batch_tab(const tch_batch* b);
const int32_t* tch_batch_off(const tch_batch* b);
const int32_t* tch_batch_ind(const tch_batch* b);
const int32_t* tch_batch_kind(const tch_batch* b);
/* Graph g of the batch as a standalone CSR (copied). */
tch_graph*     tch_batch_graph(const tch_batch* b, int32_t g);
void           tch_batch_free(tch_batch* b);

/* ------------------------------------------------------- serving admission
 * The serving loop's admission policy run without a device.  Requests wait
 * first in first out; the next batch is the longest prefix of the queue (at
 * most max_batch requests) whose block fits the largest free block of a
 * first-fit arena of `budget` bytes; the head is never overtaken.
 * script[i] > 0 submits a request whose share of a batch block is that many
 * bytes, and each block adds tcs::kBatchFixed (the serving loop's blocks
 * add tcs::kBlockFixed, which also covers the scratch's fixed part);
 * script[i] < 0 completes request -script[i]-1, and a batch's block is
 * released when all its requests have completed.  After every step batches
 * are admitted while the head fits.  admitted_at[r] gets the step at which
 * request r was admitted (-1 if never) and batch_of[r] its batch number;
 * returns the number of requests admitted, or a negative TCH_E* code
 * (TCH_ENOMEM for a request that can never fit). */
int tch_admission_sim(int64_t budget, int32_t max_batch, int32_t nsteps, const int64_t* script,
                      int32_t* admitted_at, int32_t* batch_of, int64_t* peak_used);
/* Arena bytes of a request of nv vertices and nnz arcs batched alone
 * (its input, counts and total, and the scratch of its count). */
int64_t tch_request_bytes(int64_t nv, int64_t nnz);
/* Exact bytes of the block of a batch of G graphs with these totals: the
 * batched counter's packed input and output, then the count's scratch
 * (tcs::scratch_bound bytes). */
int64_t tch_batch_block_bytes(int32_t G, int64_t nv_total, int64_t nnz_total);
/* Packs G graphs into a host block as the serving loop uploads it: the
 * int64 table at byte 0, offsets and ids at the regions tcs::batch_layout
 * gives (each 256-byte aligned).  Returns the bytes written (the upload
 * size) or a negative TCH_E* code. */
int64_t tch_pack_block(int32_t G, const int32_t* nv, const int64_t* nnz, const int32_t* const* off,
                       const int32_t* const* ind, void* block, int64_t block_bytes);
/* The serving loop's copy pool (tcs::CopyPool of `threads` threads) run by
 * `callers` threads at once, each copying `rounds` times n buffers
 * (src[i] -> dst[c][i], bytes[i] bytes; dst[c] is caller c's own set).
 * Returns 0, or a negative TCH_E* code. */
int tch_copy_pool_run(int32_t threads, int32_t callers, int32_t rounds, int32_t n, void* const* dst,
                      const void* const* src, const int64_t* bytes);

enum {
  TCH_OK = 0,
  TCH_EINVAL = -1,     /* bad argument */
  TCH_EIO = -2,        /* file missing, unreadable or truncated */
  TCH_ERANGE = -3,     /* id out of range or count beyond int32 */
  TCH_EUNSORTED = -4,  /* row not strictly increasing (includes duplicates) */
  TCH_ESELFLOOP = -5,
  TCH_EASYM = -6,      /* an arc without its reverse */
  TCH_ENOMEM = -7
};

#ifdef __cplusplus
}
#endif
#endif
