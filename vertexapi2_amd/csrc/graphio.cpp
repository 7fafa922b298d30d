// Graph files with the reference's reading rules (graphio.cpp):
//   .mtx  MatrixMarket coordinate             (graphio.cpp:190-262)
//   .edge GraphLab SNAP edge list             (graphio.cpp:178-188)
//   .gr   Lonestar binary CSR, version 1      (graphio.cpp:266-414)
// plus the MatrixMarket writer (graphio.cpp:416-439) and the dispatch by
// file extension (graphio.cpp:441-473).  Every file is read through zlib,
// which passes uncompressed files through unchanged, so "x.mtx.gz" and
// "x.mtx" load alike.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <climits>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "gas.h"
#include "gas_internal.hpp"

namespace gas {
namespace {

struct GzFile {
  gzFile f = nullptr;
  explicit GzFile(const char *path) : f(gzopen(path, "rb")) {}
  ~GzFile() {
    if (f) gzclose(f);
  }
  // Next line without its terminator; false at end of file or on error
  // (err set for a read error).
  bool line(std::string &out, bool &err) {
    out.clear();
    err = false;
    char buf[4096];
    for (;;) {
      if (!gzgets(f, buf, sizeof(buf))) {
        int code = 0;
        gzerror(f, &code);
        if (code != Z_OK && code != Z_STREAM_END) err = true;
        return !out.empty() && !err;
      }
      out += buf;
      if (!out.empty() && out.back() == '\n') {
        out.pop_back();
        return true;
      }
    }
  }
  bool read(void *p, size_t n) {
    char *c = static_cast<char *>(p);
    while (n) {
      const unsigned chunk = (unsigned)std::min<size_t>(n, 1u << 30);
      const int got = gzread(f, c, chunk);
      if (got <= 0) return false;
      c += got;
      n -= (size_t)got;
    }
    return true;
  }
};

enum Symmetry { SYM_NONE, SYM_SYMMETRIC, SYM_SKEW, SYM_HERMITIAN };

// Rules of the line-based loaders (loadGraph_common, graphio.cpp:69-175).
struct TextRules {
  const char *who;
  char comment;
  bool skip_size_line;  // MatrixMarket "rows cols nnz"
  bool one_based;
  Symmetry sym;
  bool values;
};

bool blank(const std::string &s) {
  for (char c : s)
    if (!isspace((unsigned char)c)) return false;
  return true;
}

// Edge lines of an open text file: comments and blank lines skipped; a self
// loop is dropped, and (as in graphio.cpp:133-141) counts toward the
// largest id only when no values are read; a symmetric file stores both
// directions (a skew-symmetric one negates the mirrored value).
int read_edge_lines(GzFile &f, const TextRules &r, int64_t line_no,
                    gas_edges_t *out) {
  std::unique_ptr<gas_edges> e(new gas_edges);
  int64_t max_vertex = 0;
  bool size_line = r.skip_size_line;
  std::string line;
  bool err = false;
  while (f.line(line, err)) {
    ++line_no;
    if (line.empty() || line[0] == r.comment || blank(line)) continue;
    if (size_line) {
      size_line = false;
      continue;
    }
    long s, d;
    int nb = 0;
    if (sscanf(line.c_str(), "%ld%ld%n", &s, &d, &nb) != 2) {
      return set_error(GAS_ERR_IO, "%s: error parsing src/dst at line %ld",
                       r.who, (long)line_no);
    }
    if (r.one_based) {
      --s;
      --d;
    }
    if (s < 0 || d < 0 || s > INT32_MAX - 1 || d > INT32_MAX - 1) {
      return set_error(GAS_ERR_IO, "%s: bad vertex id at line %ld", r.who,
                       (long)line_no);
    }
    if (s == d) {
      if (r.values) continue;
    } else {
      // edge ids are int32 (gpugas.h:75): a list that would outgrow them
      // (a symmetric file stores two per line) is refused, not wrapped
      if ((int64_t)e->src.size() + (r.sym != SYM_NONE ? 2 : 1) > INT32_MAX)
        return set_error(GAS_ERR_IO, "%s: more than 2^31 - 1 edges at line %ld", r.who,
                         (long)line_no);
      e->src.push_back((int32_t)s);
      e->dst.push_back((int32_t)d);
      if (r.sym != SYM_NONE) {
        e->src.push_back((int32_t)d);
        e->dst.push_back((int32_t)s);
      }
      if (r.values) {
        float val = 0.f;
        sscanf(line.c_str() + nb, "%f", &val);
        e->w.push_back((int32_t)val);
        if (r.sym == SYM_SYMMETRIC) e->w.push_back((int32_t)val);
        if (r.sym == SYM_SKEW) e->w.push_back((int32_t)(-val));
      }
    }
    max_vertex = std::max(max_vertex, (int64_t)std::max(s, d));
  }
  if (err) {
    return set_error(GAS_ERR_IO, "%s: read error after line %ld", r.who,
                     (long)line_no);
  }
  e->nV = (int32_t)(max_vertex + 1);
  *out = e.release();
  return GAS_OK;
}

// Stable counting sort of an edge list by source: CSR offsets (V+1) and,
// per slot, the index of its edge in the list.
void csr_order(const gas_edges &e, std::vector<int64_t> &off,
               std::vector<int32_t> &order) {
  off.assign((size_t)e.nV + 1, 0);
  for (int32_t s : e.src) off[(size_t)s + 1]++;
  for (int32_t v = 0; v < e.nV; ++v) off[(size_t)v + 1] += off[(size_t)v];
  std::vector<int64_t> pos(off.begin(), off.end() - 1);
  order.resize(e.src.size());
  for (size_t i = 0; i < e.src.size(); ++i)
    order[(size_t)pos[(size_t)e.src[i]]++] = (int32_t)i;
}

bool has_suffix(const char *path, const char *ext) {
  // graphio.cpp:447-460: the extension after the last '.', or after the
  // one before it when the file ends in ".gz", compared by prefix
  const char *dot = strrchr(path, '.');
  if (!dot) return false;
  if (strcmp(dot, ".gz") == 0) {
    const char *p = dot - 1;
    while (p >= path && *p != '.') --p;
    if (p < path) return false;
    dot = p;
  }
  return strncmp(dot, ext, strlen(ext)) == 0;
}

// Host allocation failures and other exceptions become error codes at the
// C boundary.
template <class F>
int io_guard(const char *who, F f) {
  try {
    return f();
  } catch (const std::bad_alloc &) {
    return set_error(GAS_ERR_NOMEM, "%s: out of host memory", who);
  } catch (const std::exception &x) {
    return set_error(GAS_ERR_IO, "%s: %s", who, x.what());
  }
}

}  // namespace
}  // namespace gas

using namespace gas;

static int load_mtx_impl(const char *path, int want_values, gas_edges_t *out) {
  if (!path || !out) return set_error(GAS_ERR_ARG, "gas_load_mtx: null");
  GzFile f(path);
  if (!f.f) return set_error(GAS_ERR_IO, "gas_load_mtx: cannot open %s", path);
  std::string line;
  bool err = false;
  if (!f.line(line, err))
    return set_error(GAS_ERR_IO, "gas_load_mtx: error reading header");
  // banner, then "matrix coordinate <field> <symmetry>" (graphio.cpp:207-258)
  char tok[5][64] = {{0}};
  const int nt = sscanf(line.c_str(), "%63s %63s %63s %63s %63s", tok[0],
                        tok[1], tok[2], tok[3], tok[4]);
  if (nt < 2 || strcmp(tok[1], "matrix") != 0)
    return set_error(GAS_ERR_IO, "gas_load_mtx: unrecognized token %s", tok[1]);
  if (nt < 3 || strcmp(tok[2], "coordinate") != 0)
    return set_error(GAS_ERR_IO,
                     "gas_load_mtx: only coordinate format is supported");
  TextRules r{"gas_load_mtx", '%', true, true, SYM_NONE, want_values != 0};
  if (strcmp(tok[3], "pattern") == 0) r.values = false;
  if (strcmp(tok[3], "complex") == 0)
    return set_error(GAS_ERR_IO, "gas_load_mtx: complex values unsupported");
  if (strcmp(tok[4], "general") == 0) r.sym = SYM_NONE;
  else if (strcmp(tok[4], "symmetric") == 0) r.sym = SYM_SYMMETRIC;
  else if (strcmp(tok[4], "skew-symmetric") == 0) r.sym = SYM_SKEW;
  else if (strcmp(tok[4], "hermitian") == 0) r.sym = SYM_HERMITIAN;
  else
    return set_error(GAS_ERR_IO, "gas_load_mtx: unrecognized symmetry '%s'",
                     tok[4]);
  if (r.sym == SYM_HERMITIAN && r.values)
    return set_error(GAS_ERR_IO,
                     "gas_load_mtx: hermitian values are not supported");
  return read_edge_lines(f, r, 1, out);
}

static int load_edge_impl(const char *path, gas_edges_t *out) {
  if (!path || !out) return set_error(GAS_ERR_ARG, "gas_load_edge: null");
  GzFile f(path);
  if (!f.f) return set_error(GAS_ERR_IO, "gas_load_edge: cannot open %s", path);
  // '#' comments, 0-based ids, no values, self loops dropped
  // (graphio.cpp:178-188)
  const TextRules r{"gas_load_edge", '#', false, false, SYM_NONE, false};
  return read_edge_lines(f, r, 0, out);
}

static int load_gr_impl(const char *path, int want_values, gas_edges_t *out) {
  if (!path || !out) return set_error(GAS_ERR_ARG, "gas_load_gr: null");
  GzFile f(path);
  if (!f.f) return set_error(GAS_ERR_IO, "gas_load_gr: cannot open %s", path);
  uint64_t hdr[4];  // version, edge value bytes, vertices, edges
  if (!f.read(hdr, sizeof(hdr)))
    return set_error(GAS_ERR_IO, "gas_load_gr: error reading header of %s", path);
  const uint64_t wbytes = hdr[1], nV = hdr[2], nE = hdr[3];
  if (wbytes == 0 && want_values > 0)
    return set_error(GAS_ERR_IO, "gas_load_gr: file does not have edge values");
  if (wbytes != 0 && wbytes != 4)
    return set_error(GAS_ERR_IO, "gas_load_gr: file edge data is %llu bytes wide",
                     (unsigned long long)wbytes);
  if (nV > (uint64_t)INT32_MAX - 1 || nE > (uint64_t)INT32_MAX)
    return set_error(GAS_ERR_IO, "gas_load_gr: %llu vertices / %llu edges exceed int32",
                     (unsigned long long)nV, (unsigned long long)nE);
  std::vector<uint64_t> ends((size_t)nV);
  std::unique_ptr<gas_edges> e(new gas_edges);
  e->nV = (int32_t)nV;
  e->src.resize((size_t)nE);
  e->dst.resize((size_t)nE);
  uint32_t pad = 0;
  bool ok = (nV == 0 || f.read(ends.data(), 8 * (size_t)nV)) &&
            (nE == 0 || f.read(e->dst.data(), 4 * (size_t)nE)) &&
            (nE % 2 == 0 || f.read(&pad, 4));
  if (ok && want_values != 0 && wbytes == 4) {
    e->w.resize((size_t)nE);
    ok = nE == 0 || f.read(e->w.data(), 4 * (size_t)nE);
  }
  if (!ok) {
    return set_error(GAS_ERR_IO, "gas_load_gr: error reading lonestar binary CSR file %s",
                     path);
  }
  // offsets are each vertex's end; the list is expanded to one source per
  // edge (graphio.cpp:339-346)
  uint64_t j = 0;
  for (uint64_t v = 0; v < nV; ++v) {
    if (ends[v] < j || ends[v] > nE) {
      return set_error(GAS_ERR_IO, "gas_load_gr: bad offset of vertex %llu",
                       (unsigned long long)v);
    }
    for (; j < ends[v]; ++j) e->src[j] = (int32_t)v;
  }
  if (j != nE) {
    return set_error(GAS_ERR_IO, "gas_load_gr: offsets end at %llu of %llu edges",
                     (unsigned long long)j, (unsigned long long)nE);
  }
  for (int32_t d : e->dst)
    if (d < 0 || (uint64_t)d >= nV) {
      return set_error(GAS_ERR_IO, "gas_load_gr: destination %d out of range", d);
    }
  *out = e.release();
  return GAS_OK;
}

static int write_gr_impl(const char *path, gas_edges_t e) {
  if (!path || !e) return set_error(GAS_ERR_ARG, "gas_write_gr: null");
  std::vector<int64_t> off;
  std::vector<int32_t> order;
  csr_order(*e, off, order);
  const bool values = !e->w.empty();
  const size_t nE = e->src.size();
  FILE *f = fopen(path, "wb");
  if (!f) return set_error(GAS_ERR_IO, "gas_write_gr: unable to write to file %s", path);
  // version 1, edge value bytes, vertices, edges, each vertex's end offset
  // (graphio.cpp:365-414)
  const uint64_t hdr[4] = {1, values ? 4u : 0u, (uint64_t)e->nV, (uint64_t)nE};
  bool ok = fwrite(hdr, 8, 4, f) == 4;
  for (int32_t v = 0; ok && v < e->nV; ++v) {
    const uint64_t end = (uint64_t)off[(size_t)v + 1];
    ok = fwrite(&end, 8, 1, f) == 1;
  }
  std::vector<int32_t> col(nE);
  for (size_t k = 0; k < nE; ++k) col[k] = e->dst[(size_t)order[k]];
  ok = ok && (nE == 0 || fwrite(col.data(), 4, nE, f) == nE);
  const uint32_t pad = 0;
  ok = ok && (nE % 2 == 0 || fwrite(&pad, 4, 1, f) == 1);
  if (values) {
    for (size_t k = 0; k < nE; ++k) col[k] = e->w[(size_t)order[k]];
    ok = ok && (nE == 0 || fwrite(col.data(), 4, nE, f) == nE);
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? GAS_OK : set_error(GAS_ERR_IO, "gas_write_gr: short write to %s", path);
}

static int write_mtx_impl(const char *path, gas_edges_t e) {
  if (!path || !e) return set_error(GAS_ERR_ARG, "gas_write_mtx: null");
  FILE *f = fopen(path, "w");
  if (!f) return set_error(GAS_ERR_IO, "gas_write_mtx: unable to write to file %s", path);
  // graphio.cpp:416-439: a banner with one '%' (what the reference's
  // printf of "%%MatrixMarket" writes), "V V E", then 1-based
  // "src dst [value]" lines in list order
  const bool values = !e->w.empty();
  const size_t nE = e->src.size();
  bool ok = fprintf(f, "%%MatrixMarket matrix coordinate %s general\n",
                    values ? "Integer" : "pattern") > 0 &&
            fprintf(f, "%d %d %zu\n", e->nV, e->nV, nE) > 0;
  for (size_t i = 0; ok && i < nE; ++i)
    ok = (values ? fprintf(f, "%d %d %d\n", e->src[i] + 1, e->dst[i] + 1, e->w[i])
                 : fprintf(f, "%d %d\n", e->src[i] + 1, e->dst[i] + 1)) > 0;
  ok = (fclose(f) == 0) && ok;
  return ok ? GAS_OK : set_error(GAS_ERR_IO, "gas_write_mtx: short write to %s", path);
}

int gas_load_graph(const char *path, int want_values, gas_edges_t *out) {
  if (!path || !out) return set_error(GAS_ERR_ARG, "gas_load_graph: null");
  if (has_suffix(path, ".edge")) return gas_load_edge(path, out);
  if (has_suffix(path, ".mtx")) return gas_load_mtx(path, want_values, out);
  if (has_suffix(path, ".gr")) return gas_load_gr(path, want_values, out);
  return set_error(GAS_ERR_IO, "gas_load_graph: unrecognized filetype extension of %s",
                   path);
}

int gas_load_mtx(const char *path, int want_values, gas_edges_t *out) {
  return io_guard("gas_load_mtx", [&] { return load_mtx_impl(path, want_values, out); });
}

int gas_load_edge(const char *path, gas_edges_t *out) {
  return io_guard("gas_load_edge", [&] { return load_edge_impl(path, out); });
}

int gas_load_gr(const char *path, int want_values, gas_edges_t *out) {
  return io_guard("gas_load_gr", [&] { return load_gr_impl(path, want_values, out); });
}

int gas_write_gr(const char *path, gas_edges_t e) {
  return io_guard("gas_write_gr", [&] { return write_gr_impl(path, e); });
}

int gas_write_mtx(const char *path, gas_edges_t e) {
  return io_guard("gas_write_mtx", [&] { return write_mtx_impl(path, e); });
}
