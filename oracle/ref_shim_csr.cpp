// C-ABI wrappers around the reference METIS reader and .info writer
// (src/csrdatacreate.cpp:3-49 and 86-114).  That file defines templates
// without explicit instantiations, so this unit compiles it in place (found
// through the include path oracle/Makefile passes) and instantiates the two
// functions used.  TEST INFRASTRUCTURE ONLY.
#include <cstdint>
#include <string>
#include <vector>
#include "csrdatacreate.cpp"

extern "C" {
// Reads a METIS file; copies CSR into caller buffers when they are large
// enough (call with zero capacity first to size them).
int tcref_read_metis(const char* path, int32_t* nv, int64_t* nnz, int32_t* off, int64_t off_cap, int32_t* ind,
                     int64_t ind_cap) {
  std::vector<int32_t> o, i;
  int32_t v = 0, e = 0;
  try {
    readGraph<int32_t>(std::string(path), o, i, v, e);
  } catch (...) {
    return -1;
  }
  *nv = v;
  *nnz = (int64_t)i.size();
  if (!off || off_cap < (int64_t)o.size() || !ind || ind_cap < (int64_t)i.size()) return 1;
  for (size_t k = 0; k < o.size(); ++k) off[k] = o[k];
  for (size_t k = 0; k < i.size(); ++k) ind[k] = i[k];
  return 0;
}
void tcref_write_info(const char* path, int32_t nv, const int32_t* off, const int32_t* ind) {
  std::vector<int32_t> o(off, off + (size_t)nv + 1), i(ind, ind + off[nv]);
  writeGraph<int32_t>(std::string(path), o, i);
}
}
