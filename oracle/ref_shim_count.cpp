// C-ABI wrappers around the compiled reference CPU code (src/clusteringCount.cpp
// and src/graphRead.cpp, built unmodified from the reference tree by
// oracle/Makefile).  TEST INFRASTRUCTURE ONLY: loaded by tests/ and by the
// cpu_baseline leg of bench.py, never by the product path.
#include <cstdint>
#include <string>

// Declarations matching the reference templates (clusteringCount.hpp:11-33,
// graphRead.hpp:4-6); the definitions and explicit instantiations live in the
// reference objects this file is linked with.
template <typename T>
int64_t count_intersectionsReg(const T ai, const T alen, const T* a, const T bi, const T blen, const T* b);
template <typename T>
void count_all_trianglesCPU(const T nv, const T* off, const T* ind, int64_t* triNE, int64_t* allTriangles);
template <typename T>
void count_all_trianglesCPU(const T nv, const T* vertexList, const T* off, const T* ind, int64_t* triNE,
                            int64_t* allTriangles);
template <typename T>
void readGraph(std::string filePath, T*& offset, T*& index, T& numVertices, T& numEdges);

extern "C" {
int64_t tcref_count_intersections(int32_t ai, int32_t alen, const int32_t* a, int32_t bi, int32_t blen,
                                  const int32_t* b) {
  return count_intersectionsReg<int32_t>(ai, alen, a, bi, blen, b);
}
void tcref_count_all(int32_t nv, const int32_t* off, const int32_t* ind, int64_t* triNE, int64_t* total) {
  count_all_trianglesCPU<int32_t>(nv, off, ind, triNE, total);
}
void tcref_count_vertex_list(int32_t count, const int32_t* list, const int32_t* off, const int32_t* ind,
                             int64_t* triNE, int64_t* total) {
  count_all_trianglesCPU<int32_t>(count, list, off, ind, triNE, total);
}
// Bounded CPU baseline: the reference intersection (count_intersectionsReg,
// clusteringCount.cpp:2-48) on a caller-chosen list of arcs (k, src[k]) in
// the order given, exactly as count_all_trianglesCPU calls it per arc
// (clusteringCount.cpp:57-68).  Returns the sum of the per-arc counts.
int64_t tcref_count_arc_sample(const int32_t* off, const int32_t* ind, const int32_t* src, const int64_t* arc,
                               int64_t count) {
  int64_t sum = 0;
  for (int64_t s = 0; s < count; ++s) {
    const int32_t u = src[s];
    const int32_t w = ind[arc[s]];
    sum += count_intersectionsReg<int32_t>(u, off[u + 1] - off[u], ind + off[u], w, off[w + 1] - off[w], ind + off[w]);
  }
  return sum;
}
// readGraph<int32_t>: returns nv and nnz; copies into caller buffers of the
// given capacity (call once with zero capacity to size them).
int tcref_read_info(const char* path_no_ext, int32_t* nv, int32_t* nnz, int32_t* off, int64_t off_cap,
                    int32_t* ind, int64_t ind_cap) {
  int32_t *o = nullptr, *i = nullptr, v = 0, e = 0;
  readGraph<int32_t>(std::string(path_no_ext), o, i, v, e);
  *nv = v;
  *nnz = e;
  int rc = 0;
  if (off && off_cap >= (int64_t)v + 1 && ind && ind_cap >= e) {
    for (int64_t k = 0; k <= v; ++k) off[k] = o[k];
    for (int64_t k = 0; k < e; ++k) ind[k] = i[k];
  } else {
    rc = 1;
  }
  delete[] o;
  delete[] i;
  return rc;
}
}
