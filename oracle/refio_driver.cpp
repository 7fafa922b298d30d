// C entry points over the reference's own graph loaders and writers
// (graphio.cpp, compiled from its source by oracle/Makefile), so the tests
// can compare the gas_load_* / gas_write_* entry points with them.  Test
// infrastructure only.
// The reference's loaders exit(1) on malformed input: call them with
// well-formed files only.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "graphio.h"

static int *dup_vec(const std::vector<int> &v) {
  int *p = static_cast<int *>(malloc(sizeof(int) * (v.size() ? v.size() : 1)));
  if (!v.empty()) memcpy(p, v.data(), sizeof(int) * v.size());
  return p;
}

extern "C" {

// loadGraph (graphio.cpp:441-473); *vals is NULL when no values were read.
int refio_load(const char *path, int want_values, int *n_vertices, int *n_edges,
               int **srcs, int **dsts, int **vals) {
  int nV = 0;
  std::vector<int> s, d, w;
  loadGraph(path, nV, s, d, want_values ? &w : nullptr);
  *n_vertices = nV;
  *n_edges = (int)d.size();
  *srcs = dup_vec(s);
  *dsts = dup_vec(d);
  *vals = (want_values && w.size() == d.size() && !d.empty()) ? dup_vec(w) : nullptr;
  return 0;
}

// writeGraph_binaryCSR (graphio.cpp:365-414) from CSR arrays.
int refio_write_gr(const char *path, int n_vertices, int n_edges,
                   const int *offsets, const int *dsts, const int *vals) {
  writeGraph_binaryCSR(path, n_vertices, n_edges, offsets, dsts, vals);
  return 0;
}

// writeGraph_mtx (graphio.cpp:416-439).
int refio_write_mtx(const char *path, int n_vertices, int n_edges,
                    const int *srcs, const int *dsts, const int *vals) {
  writeGraph_mtx(path, n_vertices, n_edges, srcs, dsts, vals);
  return 0;
}

void refio_free(void *p) { free(p); }

}  // extern "C"
