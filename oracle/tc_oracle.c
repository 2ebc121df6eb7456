/*
 * tc_oracle.c — CPU restatement of the reference's triangle-count path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker; the product path (gputrianglecounting_amd) never calls it.
 *
 * What it restates (reference files, paths relative to the reference root):
 *   oracle_count_intersections   src/clusteringCount.cpp:2-48
 *       count_intersectionsReg: merge of two sorted rows with a self-id skip
 *       (a[ka] == ai, b[kb] == bi are stepped over once per iteration) and a
 *       linear skip-ahead; early exit when the value ranges are disjoint.
 *   oracle_count_arcs            src/clusteringCount.cpp:50-70
 *       count_all_trianglesCPU: one intersection per arc in CSR order into
 *       triNE[arc] (int64) and the int64 total.
 *   oracle_count_vertex_list     src/clusteringCount.cpp:125-155
 *       the vertex-list overload (only the listed source rows).
 *   oracle_per_vertex            per-vertex value = sum of triNE over the row
 *       [off[v], off[v+1]); the reference GPU kernel stores exactly this sum
 *       per source vertex (src/clusteringCount.cu:331-357), which is 2*T(v)
 *       on a simple graph.
 *   oracle_per_vertex_omp        same numbers, rows split over OpenMP threads.
 *   oracle_vertex_list_omp       the vertex-list overload's per-position row
 *       sums with the listed rows' arcs split over OpenMP threads (the
 *       sampled check of graphs too large for a full CPU count).
 *   oracle_vertex_list_lim       the vertex-list GPU kernel's per-position
 *       sums (src/clusteringCount.cu:266-312: arcs with src != dest and both
 *       lengths >= 2) with the shadow-vertex intersection of
 *       src/clusteringCount.cu:150-177, each match x adding 1 + (x > lim),
 *       applied to every match of a plain merge.
 *   oracle_count_arc_sample      the per-arc intersection of clusteringCount.cpp:57-68
 *       on a caller-chosen arc list (bounded CPU baseline).
 *   oracle_lcc                   local clustering coefficient
 *       out[v] / (d(v) * (d(v) - 1)) in IEEE double, 0 for d < 2 (no
 *       reference counterpart; derived value).
 *
 * How each function is pinned (tests/test_oracle.py): against the compiled
 * reference CPU code (oracle/_ref) on random graphs including self-loops,
 * duplicates and directed rows, against committed golden vectors produced
 * from it, against closed forms (K_n, wheels, friendship graphs, cycles,
 * bipartite graphs), against diag(A^3) by dense and sparse matrix products
 * and against networkx.triangles.  oracle_vertex_list_lim (the reference's
 * lim overload has no caller and no CPU form to compile): lim = -1 doubles
 * and lim >= nv - 1 equals oracle_count_vertex_list's sums, closed forms
 * on K_n with lim = k, and a brute-force sum over vertex triples.
 */
#include <stdint.h>
#include <stdlib.h>

/* clusteringCount.cpp:2-48 */
int64_t oracle_count_intersections(int32_t ai, int32_t alen, const int32_t* a, int32_t bi, int32_t blen,
                                   const int32_t* b) {
  int32_t ka = 0, kb = 0;
  int64_t out = 0;
  if (!alen || !blen || a[alen - 1] < b[0] || b[blen - 1] < a[0]) return 0;
  for (;;) {
    if (ka >= alen || kb >= blen) break;
    int32_t va = a[ka];
    int32_t vb = b[kb];
    if (va == ai) {
      ++ka;
      if (ka >= alen) break;
      va = a[ka];
    }
    if (vb == bi) {
      ++kb;
      if (kb >= blen) break;
      vb = b[kb];
    }
    if (va == vb) {
      ++ka;
      ++kb;
      ++out;
    } else if (va < vb) {
      ++ka;
      while (ka < alen && a[ka] < vb) ++ka;
    } else {
      ++kb;
      while (kb < blen && va > b[kb]) ++kb;
    }
  }
  return out;
}

/* clusteringCount.cpp:50-70 */
void oracle_count_arcs(int32_t nv, const int32_t* off, const int32_t* ind, int64_t* triNE, int64_t* total) {
  int64_t edge = 0, sum = 0;
  for (int32_t src = 0; src < nv; ++src) {
    int32_t srcLen = off[src + 1] - off[src];
    for (int32_t it = off[src]; it < off[src + 1]; ++it) {
      int32_t dest = ind[it];
      int32_t destLen = off[dest + 1] - off[dest];
      triNE[edge] = oracle_count_intersections(src, srcLen, ind + off[src], dest, destLen, ind + off[dest]);
      sum += triNE[edge++];
    }
  }
  *total = sum;
}

/* clusteringCount.cpp:125-155: rows of vertexList[0..count), triNE packed in
 * visiting order. */
void oracle_count_vertex_list(int32_t count, const int32_t* vertexList, const int32_t* off, const int32_t* ind,
                              int64_t* triNE, int64_t* total) {
  int64_t edge = 0, sum = 0;
  for (int32_t s = 0; s < count; ++s) {
    int32_t src = vertexList[s];
    int32_t srcLen = off[src + 1] - off[src];
    for (int32_t it = off[src]; it < off[src + 1]; ++it) {
      int32_t dest = ind[it];
      int32_t destLen = off[dest + 1] - off[dest];
      triNE[edge] = oracle_count_intersections(src, srcLen, ind + off[src], dest, destLen, ind + off[dest]);
      sum += triNE[edge++];
    }
  }
  *total = sum;
}

/* src/clusteringCount.cu:266-312 with the intersection of :150-177: out[s]
 * = sum over the arcs src -> dest of vertexList[s] (src != dest, both rows
 * of length >= 2) of sum over x in N(src) ∩ N(dest) of 1 + (x > lim). */
void oracle_vertex_list_lim(int32_t count, const int32_t* vertexList, const int32_t* off, const int32_t* ind,
                            int32_t lim, int64_t* out, int64_t* total) {
  int64_t sum = 0;
  for (int32_t s = 0; s < count; ++s) {
    const int32_t src = vertexList[s];
    const int32_t srcLen = off[src + 1] - off[src];
    const int32_t* a = ind + off[src];
    int64_t row = 0;
    for (int32_t it = off[src]; it < off[src + 1]; ++it) {
      const int32_t dest = ind[it];
      const int32_t destLen = off[dest + 1] - off[dest];
      if (src == dest || srcLen < 2 || destLen < 2) continue;
      const int32_t* b = ind + off[dest];
      int32_t ka = 0, kb = 0;
      while (ka < srcLen && kb < destLen) {
        const int32_t comp = a[ka] < b[kb] ? -1 : a[ka] > b[kb] ? 1 : 0;
        if (comp == 0) row += 1 + (a[ka] > lim);
        ka += comp <= 0;
        kb += comp >= 0;
      }
    }
    out[s] = row;
    sum += row;
  }
  *total = sum;
}

/* Per-arc intersections for a caller-chosen arc list (arc index arc[s] whose
 * source row is src[s]), in the order given; the bounded CPU baseline. */
int64_t oracle_count_arc_sample(const int32_t* off, const int32_t* ind, const int32_t* src, const int64_t* arc,
                                int64_t count) {
  int64_t sum = 0;
  for (int64_t s = 0; s < count; ++s) {
    int32_t u = src[s], w = ind[arc[s]];
    sum += oracle_count_intersections(u, off[u + 1] - off[u], ind + off[u], w, off[w + 1] - off[w], ind + off[w]);
  }
  return sum;
}

/* Per-vertex row sums of the per-arc counts.  Returns 0, or -1 when a row
 * sum does not fit int32 (out[] then holds the low 32 bits). */
static int per_vertex_rows(int32_t lo, int32_t hi, const int32_t* off, const int32_t* ind, int32_t* out,
                           int64_t* partial) {
  int rc = 0;
  int64_t sum = 0;
  for (int32_t src = lo; src < hi; ++src) {
    int32_t srcLen = off[src + 1] - off[src];
    int64_t row = 0;
    for (int32_t it = off[src]; it < off[src + 1]; ++it) {
      int32_t dest = ind[it];
      int32_t destLen = off[dest + 1] - off[dest];
      row += oracle_count_intersections(src, srcLen, ind + off[src], dest, destLen, ind + off[dest]);
    }
    if (row > INT32_MAX) rc = -1;
    out[src] = (int32_t)(uint32_t)(uint64_t)row;
    sum += row;
  }
  *partial = sum;
  return rc;
}

int oracle_per_vertex(int32_t nv, const int32_t* off, const int32_t* ind, int32_t* out, int64_t* total) {
  return per_vertex_rows(0, nv, off, ind, out, total);
}

int oracle_per_vertex_omp(int32_t nv, const int32_t* off, const int32_t* ind, int32_t* out, int64_t* total,
                          int32_t threads) {
  int64_t sum = 0;
  int rc = 0;
  const int32_t chunk = 256;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : sum) reduction(| : rc) num_threads(threads)
  for (int32_t lo = 0; lo < nv; lo += chunk) {
    int32_t hi = lo + chunk < nv ? lo + chunk : nv;
    int64_t part = 0;
    rc |= per_vertex_rows(lo, hi, off, ind, out, &part) ? 1 : 0;
    sum += part;
  }
  *total = sum;
  return rc ? -1 : 0;
}

void oracle_lcc(int32_t nv, const int32_t* off, const int32_t* out, double* lcc) {
  for (int32_t v = 0; v < nv; ++v) {
    int64_t d = off[v + 1] - off[v];
    lcc[v] = d < 2 ? 0.0 : (double)out[v] / (double)(d * (d - 1));
  }
}

/* clusteringCount.cpp:125-155 with the work of the listed rows spread over
 * OpenMP threads: out[s] = the sum over the arcs of row vertexList[s] of
 * count_intersectionsReg (the row sum of the reference's triNE for that
 * source, 2*T(v) on a simple graph), *total = the sum of out.  pre[s] is
 * the caller's exclusive prefix of the listed rows' lengths (pre[count] =
 * arcs of all listed rows); the arcs are dealt to the threads in chunks,
 * so one hub row is shared by all of them.  Pinned against
 * oracle_count_vertex_list and the compiled reference on small graphs. */
void oracle_vertex_list_omp(int32_t count, const int32_t* vertexList, const int64_t* pre, const int32_t* off,
                            const int32_t* ind, int64_t* out, int64_t* total, int32_t threads) {
  const int64_t arcs = count > 0 ? pre[count] : 0;
  int64_t sum = 0;
  for (int32_t s = 0; s < count; ++s) out[s] = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : sum) num_threads(threads)
  for (int64_t i = 0; i < arcs; ++i) {
    int32_t lo = 0, hi = count; /* last s with pre[s] <= i */
    while (hi - lo > 1) {
      int32_t mid = lo + (hi - lo) / 2;
      if (pre[mid] <= i) lo = mid;
      else hi = mid;
    }
    const int32_t src = vertexList[lo];
    const int32_t srcLen = off[src + 1] - off[src];
    const int32_t dest = ind[off[src] + (i - pre[lo])];
    const int32_t destLen = off[dest + 1] - off[dest];
    const int64_t c = oracle_count_intersections(src, srcLen, ind + off[src], dest, destLen, ind + off[dest]);
    if (c) {
#pragma omp atomic
      out[lo] += c;
      sum += c;
    }
  }
  *total = sum;
}
