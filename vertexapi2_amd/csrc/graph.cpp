// Host graph toolkit: seeded synthetic generators, CSR/CSC construction
// and upload of a graph into HBM (graph files: graphio.cpp).
//
// CSR/CSC layout follows util.h:79-121 (offsets of V+1 entries, neighbour
// ids sorted by key vertex, plus the original edge index of every slot);
// the sort here is a stable counting sort instead of the reference's
// std::sort on (key, index) pairs (util.h:48-72).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_set>

#include "gas.h"
#include "gas_internal.hpp"

namespace gas {

// ---------------------------------------------------------------------------
// Random numbers: SplitMix64 for seeding, xoshiro256** for streams
// (Blackman & Vigna, "Scrambled linear pseudorandom number generators").
static inline uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Xoshiro256 {
  uint64_t s[4];
  explicit Xoshiro256(uint64_t seed) {
    uint64_t x = seed;
    for (auto &v : s) v = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t x, int k) {
    return (x << k) | (x >> (64 - k));
  }
  inline uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform in [0, 1) with 53 bits
  inline double uniform() { return (next() >> 11) * 0x1.0p-53; }
  // uniform integer in [0, n) (Lemire's multiply-shift; bias < 2^-32 for
  // the n used here)
  inline uint64_t below(uint64_t n) {
    return (uint64_t)(((unsigned __int128)next() * n) >> 64);
  }
};

// Stateless hash of a vertex pair, used so both directions of a link get
// the same weight regardless of generation order.
static inline uint64_t pair_hash(uint64_t seed, uint32_t u, uint32_t v) {
  uint64_t x = seed ^ (((uint64_t)u << 32) | v);
  splitmix64(x);
  return splitmix64(x);
}

static int hw_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 4;
  return (int)std::min(n, 16u);
}

template <class F>
static void parallel_for(int64_t n, F f) {
  int nt = hw_threads();
  if (n < 1 << 16 || nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { f(b, e); });
  }
  for (auto &x : th) x.join();
}

void build_compressed(int32_t nV, int32_t nE, const int32_t *keys,
                      const int32_t *vals, std::vector<int32_t> &off,
                      std::vector<int32_t> &out_vals,
                      std::vector<int32_t> &eid) {
  off.assign((size_t)nV + 1, 0);
  for (int32_t k = 0; k < nE; ++k) off[(size_t)keys[k] + 1]++;
  for (int32_t v = 0; v < nV; ++v) off[v + 1] += off[v];
  std::vector<int32_t> cur(off.begin(), off.end() - 1);
  out_vals.resize(nE);
  eid.resize(nE);
  for (int32_t k = 0; k < nE; ++k) {
    int32_t p = cur[keys[k]]++;
    out_vals[p] = vals[k];
    eid[p] = k;
  }
}

// Sorts every adjacency list of a CSR by neighbour id and drops repeats;
// returns the new edge list in (src, dst) order.
static void dedupe_csr(int32_t nV, std::vector<int32_t> &off,
                       std::vector<int32_t> &adj) {
  std::vector<int32_t> cnt(nV, 0);
  parallel_for(nV, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) {
      int32_t *p = adj.data() + off[v];
      int32_t *q = adj.data() + off[v + 1];
      std::sort(p, q);
      cnt[v] = (int32_t)(std::unique(p, q) - p);
    }
  });
  std::vector<int32_t> noff((size_t)nV + 1, 0);
  for (int32_t v = 0; v < nV; ++v) noff[v + 1] = noff[v] + cnt[v];
  std::vector<int32_t> nadj(noff[nV]);
  parallel_for(nV, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v)
      std::copy(adj.begin() + off[v], adj.begin() + off[v] + cnt[v],
                nadj.begin() + noff[v]);
  });
  off.swap(noff);
  adj.swap(nadj);
}

// Host allocation failures and other exceptions become error codes at the
// C boundary.
template <class F>
int host_guard(const char *who, F f) {
  try {
    return f();
  } catch (const std::bad_alloc &) {
    return set_error(GAS_ERR_NOMEM, "%s: out of host memory", who);
  } catch (const std::exception &x) {
    return set_error(GAS_ERR_STATE, "%s: %s", who, x.what());
  }
}

}  // namespace gas

namespace gas {

// Vertices renumbered by descending in-degree (ties by id): hub
// destinations and, on skewed graphs, the hottest sources share cache lines,
// and each in-list stays sorted by (new) source id.  The edge list is laid
// out by new source, so the stable CSC sort keeps every in-list ascending.
gas_graph *ordered_graph(gas_graph *g) {
  // threads of different engine sets may make the first whole-run call on
  // a graph at once: one builds the copy, the others wait for it
  std::lock_guard<std::mutex> lock(g->order_mu);
  if (!g->vertex_order || (int64_t)g->nE < kOrderEdges || g->order_unskewed) return g;
  if (g->ordered) return g->ordered;
  const int32_t V = g->nV;
  // only skewed graphs: on near-regular ones (road networks, meshes) the
  // numbering as given is the local one
  int32_t max_in = 0;
  for (int32_t v = 0; v < V; ++v)
    max_in = std::max(max_in, g->h_csc_off[(size_t)v + 1] - g->h_csc_off[(size_t)v]);
  if ((int64_t)max_in < kOrderSkew * std::max<int64_t>(1, (int64_t)g->nE / std::max(V, 1))) {
    g->order_unskewed = true;  // the per-query entry points ask every call
    return g;
  }
  std::vector<int32_t> order((size_t)V);
  for (int32_t v = 0; v < V; ++v) order[(size_t)v] = v;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return g->h_csc_off[(size_t)a + 1] - g->h_csc_off[(size_t)a] >
           g->h_csc_off[(size_t)b + 1] - g->h_csc_off[(size_t)b];
  });
  std::vector<int32_t> new_of((size_t)V);
  for (int32_t k = 0; k < V; ++k) new_of[(size_t)order[(size_t)k]] = k;
  // edge weights (kept on the device only) in CSR order
  std::vector<int32_t> w;
  if (g->csr_w && g->nE) {
    w.resize((size_t)g->nE);
    if (gasrt_memcpy_d2h(w.data(), g->csr_w, sizeof(int32_t) * (size_t)g->nE, nullptr))
      throw std::runtime_error(std::string("vertex order: ") + gas_last_error());
  }
  gas_edges e;
  e.nV = V;
  e.src.reserve((size_t)g->nE);
  e.dst.reserve((size_t)g->nE);
  if (!w.empty()) e.w.reserve((size_t)g->nE);
  for (int32_t k = 0; k < V; ++k) {
    const int32_t u = order[(size_t)k];
    for (int32_t j = g->h_csr_off[(size_t)u]; j < g->h_csr_off[(size_t)u + 1]; ++j) {
      e.src.push_back(k);
      e.dst.push_back(new_of[(size_t)g->h_csr_dst[(size_t)j]]);
      if (!w.empty()) e.w.push_back(w[(size_t)j]);
    }
  }
  gas_graph_t twin = nullptr;
  int rc = gas_graph_create(&e, &twin);
  int32_t *d_new = nullptr, *d_old = nullptr;
  const size_t vb = sizeof(int32_t) * std::max<size_t>(V, 1);
  if (!rc) rc = gasrt_malloc((void **)&d_new, vb);
  if (!rc) rc = gasrt_malloc((void **)&d_old, vb);
  if (!rc) rc = gasrt_memcpy_h2d(d_new, new_of.data(), sizeof(int32_t) * (size_t)V, nullptr);
  if (!rc) rc = gasrt_memcpy_h2d(d_old, order.data(), sizeof(int32_t) * (size_t)V, nullptr);
  if (rc) {
    if (twin) gas_graph_free(twin);
    gasrt_free(d_new);
    gasrt_free(d_old);
    throw std::runtime_error(std::string("vertex order: ") + gas_last_error());
  }
  g->ordered = twin;
  g->new_of.swap(new_of);
  g->old_of.swap(order);
  g->d_new_of = d_new;
  g->d_old_of = d_old;
  return twin;
}

}  // namespace gas

using gas::set_error;

extern "C" {

int gas_edges_from_arrays(int32_t n_vertices, int32_t n_edges,
                          const int32_t *srcs, const int32_t *dsts,
                          const int32_t *weights, gas_edges_t *out) {
  return gas::host_guard("gas_edges_from_arrays", [&]() -> int {
    if (!out || n_vertices < 0 || n_edges < 0 ||
        (n_edges && (!srcs || !dsts)))
      return set_error(GAS_ERR_ARG, "gas_edges_from_arrays: bad arguments");
    for (int32_t k = 0; k < n_edges; ++k) {
      if (srcs[k] < 0 || srcs[k] >= n_vertices || dsts[k] < 0 ||
          dsts[k] >= n_vertices)
        return set_error(GAS_ERR_ARG,
                         "gas_edges_from_arrays: edge %d (%d -> %d) outside "
                         "[0, %d)",
                         k, srcs[k], dsts[k], n_vertices);
    }
    auto *e = new gas_edges;
    e->nV = n_vertices;
    e->src.assign(srcs, srcs + n_edges);
    e->dst.assign(dsts, dsts + n_edges);
    if (weights) e->w.assign(weights, weights + n_edges);
    *out = e;
    return GAS_OK;
  });
}

int gas_edges_view_get(gas_edges_t e, gas_edges_view *view) {
  if (!e || !view) return set_error(GAS_ERR_ARG, "gas_edges_view_get: null");
  view->n_vertices = e->nV;
  view->n_edges = (int32_t)e->src.size();
  view->srcs = e->src.data();
  view->dsts = e->dst.data();
  view->weights = e->w.empty() ? nullptr : e->w.data();
  return GAS_OK;
}

int gas_edges_free(gas_edges_t e) {
  delete e;
  return GAS_OK;
}

int gas_gen_grid(int32_t rows, int32_t cols, gas_edges_t *out) {
  return gas::host_guard("gas_gen_grid", [&]() -> int {
    if (!out || rows <= 0 || cols <= 0 || (int64_t)rows * cols > INT32_MAX)
      return set_error(GAS_ERR_ARG, "gas_gen_grid: bad size %d x %d", rows,
                       cols);
    auto *e = new gas_edges;
    e->nV = rows * cols;
    int64_t links = (int64_t)rows * (cols - 1) + (int64_t)(rows - 1) * cols;
    e->src.reserve(2 * links);
    e->dst.reserve(2 * links);
    for (int32_t r = 0; r < rows; ++r)
      for (int32_t c = 0; c < cols; ++c) {
        int32_t v = r * cols + c;
        if (c + 1 < cols) {
          e->src.push_back(v); e->dst.push_back(v + 1);
          e->src.push_back(v + 1); e->dst.push_back(v);
        }
        if (r + 1 < rows) {
          e->src.push_back(v); e->dst.push_back(v + cols);
          e->src.push_back(v + cols); e->dst.push_back(v);
        }
      }
    e->w.assign(e->src.size(), 1);
    *out = e;
    return GAS_OK;
  });
}

int gas_gen_road(int32_t rows, int32_t cols, double drop_frac,
                 double long_frac, int32_t wmin, int32_t wmax, uint64_t seed,
                 gas_edges_t *out) {
  return gas::host_guard("gas_gen_road", [&]() -> int {
    if (!out || rows <= 1 || cols <= 1 || (int64_t)rows * cols > INT32_MAX ||
        drop_frac < 0 || drop_frac > 1 || long_frac < 0 || wmin < 1 ||
        wmax < wmin)
      return set_error(GAS_ERR_ARG, "gas_gen_road: bad arguments");
    const int64_t V = (int64_t)rows * cols;
    const int64_t H = (int64_t)rows * (cols - 1);   // horizontal links
    const int64_t L = H + (int64_t)(rows - 1) * cols;  // all grid links
    const int64_t n_drop = (int64_t)std::floor(drop_frac * (double)L);
    const int64_t n_long = (int64_t)std::floor(long_frac * (double)L);
    if (n_long > V * (V - 1) / 2 - L)
      return set_error(GAS_ERR_ARG, "gas_gen_road: too many long links");
    gas::Xoshiro256 rng(seed);
    // Partial Fisher-Yates: the first n_drop slots are the dropped links.
    std::vector<int64_t> perm(L);
    for (int64_t i = 0; i < L; ++i) perm[i] = i;
    for (int64_t i = 0; i < n_drop; ++i) {
      int64_t j = i + (int64_t)rng.below((uint64_t)(L - i));
      std::swap(perm[i], perm[j]);
    }
    std::vector<uint8_t> dropped(L, 0);
    for (int64_t i = 0; i < n_drop; ++i) dropped[perm[i]] = 1;
    perm.clear();
    perm.shrink_to_fit();

    auto link_ends = [&](int64_t l, int32_t &u, int32_t &v) {
      if (l < H) {
        int64_t r = l / (cols - 1), c = l % (cols - 1);
        u = (int32_t)(r * cols + c);
        v = u + 1;
      } else {
        int64_t k = l - H;
        u = (int32_t)k;  // (r, c) with r < rows-1, linear id r*cols + c
        v = u + cols;
      }
    };
    auto grid_adjacent = [&](int64_t a, int64_t b) {
      if (a > b) std::swap(a, b);
      if (b - a == cols) return true;
      return b - a == 1 && (a % cols) != cols - 1;
    };

    auto *e = new gas_edges;
    e->nV = (int32_t)V;
    size_t m = 2 * (size_t)(L - n_drop + n_long);
    e->src.reserve(m);
    e->dst.reserve(m);
    e->w.reserve(m);
    const int32_t span = wmax - wmin + 1;
    auto emit = [&](int32_t u, int32_t v) {
      uint32_t a = (uint32_t)std::min(u, v), b = (uint32_t)std::max(u, v);
      int32_t w = wmin + (int32_t)(gas::pair_hash(seed, a, b) % (uint64_t)span);
      e->src.push_back(u); e->dst.push_back(v); e->w.push_back(w);
      e->src.push_back(v); e->dst.push_back(u); e->w.push_back(w);
    };
    for (int64_t l = 0; l < L; ++l) {
      if (dropped[l]) continue;
      int32_t u, v;
      link_ends(l, u, v);
      emit(u, v);
    }
    std::unordered_set<uint64_t> added;
    added.reserve((size_t)n_long * 2);
    int64_t n_added = 0;
    while (n_added < n_long) {
      int64_t u = (int64_t)rng.below((uint64_t)V);
      int64_t v = (int64_t)rng.below((uint64_t)V);
      if (u == v || grid_adjacent(u, v)) continue;
      uint64_t key = ((uint64_t)std::min(u, v) << 32) | (uint64_t)std::max(u, v);
      if (!added.insert(key).second) continue;
      emit((int32_t)u, (int32_t)v);
      ++n_added;
    }
    *out = e;
    return GAS_OK;
  });
}

int gas_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                 double c, uint64_t seed, int symmetrize, int32_t wmin,
                 int32_t wmax, gas_edges_t *out) {
  return gas::host_guard("gas_gen_rmat", [&]() -> int {
    if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 ||
        b < 0 || c < 0 || a + b + c > 1.0 + 1e-12 ||
        (wmax != 0 && (wmin < 1 || wmax < wmin)))
      return set_error(GAS_ERR_ARG, "gas_gen_rmat: bad arguments");
    const int64_t nV = (int64_t)1 << scale;
    const int64_t m = (int64_t)edge_factor * nV;
    if (m * (symmetrize ? 2 : 1) > INT32_MAX)
      return set_error(GAS_ERR_ARG, "gas_gen_rmat: too many edges for int32");
    std::vector<int32_t> us(m), vs(m);
    const int64_t CH = 1 << 16;
    const int64_t nch = (m + CH - 1) / CH;
    const double ab = a + b, abc = a + b + c;
    gas::parallel_for(nch, [&](int64_t cb, int64_t ce) {
      for (int64_t ch = cb; ch < ce; ++ch) {
        gas::Xoshiro256 rng(seed * 0x100000001B3ull + (uint64_t)ch);
        int64_t e0 = ch * CH, e1 = std::min(m, e0 + CH);
        for (int64_t k = e0; k < e1; ++k) {
          uint32_t u = 0, v = 0;
          for (int32_t lvl = 0; lvl < scale; ++lvl) {
            uint32_t bit = 1u << (scale - 1 - lvl);
            double r = rng.uniform();
            if (r < a) {
            } else if (r < ab) {
              v |= bit;
            } else if (r < abc) {
              u |= bit;
            } else {
              u |= bit;
              v |= bit;
            }
          }
          us[k] = (int32_t)u;
          vs[k] = (int32_t)v;
        }
      }
    });
    // Drop self loops; optionally add reverse edges.
    std::vector<int32_t> ks, vals;
    ks.reserve(symmetrize ? 2 * m : m);
    vals.reserve(symmetrize ? 2 * m : m);
    for (int64_t k = 0; k < m; ++k) {
      if (us[k] == vs[k]) continue;
      ks.push_back(us[k]);
      vals.push_back(vs[k]);
      if (symmetrize) {
        ks.push_back(vs[k]);
        vals.push_back(us[k]);
      }
    }
    us.clear(); us.shrink_to_fit();
    vs.clear(); vs.shrink_to_fit();
    std::vector<int32_t> off, adj, eid;
    gas::build_compressed((int32_t)nV, (int32_t)ks.size(), ks.data(),
                          vals.data(), off, adj, eid);
    ks.clear(); ks.shrink_to_fit();
    vals.clear(); vals.shrink_to_fit();
    eid.clear(); eid.shrink_to_fit();
    gas::dedupe_csr((int32_t)nV, off, adj);
    auto *e = new gas_edges;
    e->nV = (int32_t)nV;
    int64_t E = off[nV];
    e->src.resize(E);
    e->dst.swap(adj);
    gas::parallel_for(nV, [&](int64_t vb, int64_t ve) {
      for (int64_t v = vb; v < ve; ++v)
        for (int32_t k = off[v]; k < off[v + 1]; ++k) e->src[k] = (int32_t)v;
    });
    if (wmax != 0) {
      e->w.resize(E);
      const uint64_t span = (uint64_t)(wmax - wmin + 1);
      gas::parallel_for(E, [&](int64_t kb, int64_t ke) {
        for (int64_t k = kb; k < ke; ++k) {
          uint32_t x = (uint32_t)e->src[k], y = (uint32_t)e->dst[k];
          uint32_t lo = std::min(x, y), hi = std::max(x, y);
          e->w[k] = wmin + (int32_t)(gas::pair_hash(seed, lo, hi) % span);
        }
      });
    }
    *out = e;
    return GAS_OK;
  });
}

int gas_edges_max_outdegree_vertex(gas_edges_t e, int32_t *vertex,
                                   int32_t *degree) {
  return gas::host_guard("gas_edges_max_outdegree_vertex", [&]() -> int {
    if (!e || !vertex)
      return set_error(GAS_ERR_ARG, "gas_edges_max_outdegree_vertex: null");
    std::vector<int32_t> deg(e->nV, 0);
    for (int32_t s : e->src) deg[s]++;
    int32_t best = -1, bv = -1;
    for (int32_t v = 0; v < e->nV; ++v)
      if (deg[v] > best) {
        best = deg[v];
        bv = v;
      }
    *vertex = bv;
    if (degree) *degree = best;
    return GAS_OK;
  });
}

static void graph_release(gas_graph *g) {
  for (int k = 0; k < gas_graph::kEngineSets; ++k)
    for (int i = 0; i < 8; ++i)
      if (g->engine_cache[k][i] && g->engine_release[k][i])
        g->engine_release[k][i](g->engine_cache[k][i]);
  if (g->ordered) gas_graph_free(g->ordered);
  g->ordered = nullptr;
  gasrt_free(g->d_new_of);
  gasrt_free(g->d_old_of);
  g->d_new_of = g->d_old_of = nullptr;
  gasrt_free(g->ell_dst);
  gasrt_free(g->ell_w);
  g->ell_dst = g->ell_w = nullptr;
  gasrt_free(g->csc_off);
  gasrt_free(g->csc_src);
  gasrt_free(g->csr_off);
  gasrt_free(g->csr_dst);
  gasrt_free(g->csc_w);
  gasrt_free(g->csr_w);
}

static int64_t g_device_build_min_edges = (int64_t)1 << 20;

int gas_graph_set_device_build(int64_t min_edges) {
  if (min_edges < 0)
    return set_error(GAS_ERR_ARG, "gas_graph_set_device_build: negative edge count");
  g_device_build_min_edges = min_edges;
  return GAS_OK;
}

int gas_graph_ordered_view(gas_graph_t g, gas_graph_t *ordered,
                           const int32_t **new_of_dev, const int32_t **new_of_host) {
  if (!g || !ordered || !new_of_dev || !new_of_host)
    return set_error(GAS_ERR_ARG, "gas_graph_ordered_view: null");
  return gas::host_guard("gas_graph_ordered_view", [&]() -> int {
    gas_graph *og = gas::ordered_graph(g);
    *ordered = og;
    *new_of_dev = og == g ? nullptr : g->d_new_of;
    *new_of_host = og == g ? nullptr : g->new_of.data();
    return GAS_OK;
  });
}

int gas_graph_set_vertex_order(gas_graph_t g, int32_t mode) {
  if (!g || mode < 0 || mode > 1)
    return set_error(GAS_ERR_ARG, "gas_graph_set_vertex_order: bad arguments");
  std::lock_guard<std::mutex> lock(g->order_mu);
  if (mode != g->vertex_order && g->ordered) {
    gas_graph_free(g->ordered);
    g->ordered = nullptr;
    gasrt_free(g->d_new_of);
    gasrt_free(g->d_old_of);
    g->d_new_of = g->d_old_of = nullptr;
    g->new_of.clear();
    g->old_of.clear();
  }
  g->vertex_order = mode;
  return GAS_OK;
}

// CSC and CSR built on the device (graph_build.hip), then the host copies
// of the layout downloaded.
static int device_create(gas_edges_t edges, gas_graph *g) {
  const int64_t E = g->nE;
  const int32_t V = g->nV;
  const size_t vb = ((size_t)V + 1) * 4, eb = (size_t)std::max<int64_t>(E, 1) * 4;
  const bool wt = !edges->w.empty();
  int32_t *d_src = nullptr, *d_dst = nullptr, *d_w = nullptr, *d_eid = nullptr;
  uint32_t *d_keys = nullptr;
  void *d_temp = nullptr;
  const size_t temp_bytes = gas::device_build_temp_bytes(E);
  gasrt_stream_t stream = nullptr;
  int rc = gasrt_stream_create(&stream);
  auto dmalloc = [&](void **p, size_t bytes) {
    if (!rc) rc = gasrt_malloc(p, bytes);
  };
  dmalloc((void **)&g->csc_off, vb);
  dmalloc((void **)&g->csc_src, eb);
  dmalloc((void **)&g->csr_off, vb);
  dmalloc((void **)&g->csr_dst, eb);
  if (wt) {
    dmalloc((void **)&g->csc_w, eb);
    dmalloc((void **)&g->csr_w, eb);
  }
  g->device_bytes = 2 * vb + (wt ? 4 : 2) * (size_t)E * 4;
  dmalloc((void **)&d_src, eb);
  dmalloc((void **)&d_dst, eb);
  if (wt) dmalloc((void **)&d_w, eb);
  dmalloc((void **)&d_eid, eb);
  dmalloc((void **)&d_keys, eb);
  dmalloc(&d_temp, std::max<size_t>(temp_bytes, 4));
  if (!rc && E) rc = gasrt_memcpy_h2d(d_src, edges->src.data(), (size_t)E * 4, stream);
  if (!rc && E) rc = gasrt_memcpy_h2d(d_dst, edges->dst.data(), (size_t)E * 4, stream);
  if (!rc && E && wt) rc = gasrt_memcpy_h2d(d_w, edges->w.data(), (size_t)E * 4, stream);
  auto layout = [&](const int32_t *keys, const int32_t *vals, int32_t *off,
                    int32_t *out_vals, int32_t *out_w, std::vector<int32_t> &h_off,
                    std::vector<int32_t> &h_vals, std::vector<int32_t> &h_eid) {
    if (rc) return;
    rc = gas::device_build_layout(V, E, keys, vals, d_w, off, out_vals, d_eid, out_w,
                                  d_keys, d_temp, temp_bytes, stream);
    h_off.resize((size_t)V + 1);
    h_vals.resize((size_t)E);
    h_eid.resize((size_t)E);
    if (!rc) rc = gasrt_memcpy_d2h(h_off.data(), off, vb, stream);
    if (!rc && E) rc = gasrt_memcpy_d2h(h_vals.data(), out_vals, (size_t)E * 4, stream);
    if (!rc && E) rc = gasrt_memcpy_d2h(h_eid.data(), d_eid, (size_t)E * 4, stream);
  };
  layout(d_dst, d_src, g->csc_off, g->csc_src, g->csc_w, g->h_csc_off, g->h_csc_src,
         g->h_csc_eid);
  layout(d_src, d_dst, g->csr_off, g->csr_dst, g->csr_w, g->h_csr_off, g->h_csr_dst,
         g->h_csr_eid);
  if (stream) {
    // the stream drains before its buffers are freed, also after a failure
    const int rs = gasrt_stream_synchronize(stream);
    if (!rc) rc = rs;
  }
  gasrt_free(d_src);
  gasrt_free(d_dst);
  gasrt_free(d_w);
  gasrt_free(d_eid);
  gasrt_free(d_keys);
  gasrt_free(d_temp);
  if (stream) gasrt_stream_destroy(stream);
  if (!rc && wt)
    for (int32_t x : edges->w) g->max_w = std::max(g->max_w, x);
  return rc;
}

int gas_graph_create(gas_edges_t edges, gas_graph_t *out) {
  return gas::host_guard("gas_graph_create", [&]() -> int {
    if (!edges || !out) return set_error(GAS_ERR_ARG, "gas_graph_create: null");
    if (edges->src.size() > (size_t)INT32_MAX)
      return set_error(GAS_ERR_ARG, "gas_graph_create: %zu edges exceed int32",
                       edges->src.size());
    auto *g = new gas_graph;
    g->nV = edges->nV;
    g->nE = (int32_t)edges->src.size();
    if ((int64_t)g->nE >= g_device_build_min_edges) {
      const int rc = device_create(edges, g);
      if (rc) {
        graph_release(g);
        delete g;
        return rc;
      }
      *out = g;
      return GAS_OK;
    }
    gas::build_compressed(g->nV, g->nE, edges->dst.data(), edges->src.data(),
                          g->h_csc_off, g->h_csc_src, g->h_csc_eid);
    gas::build_compressed(g->nV, g->nE, edges->src.data(), edges->dst.data(),
                          g->h_csr_off, g->h_csr_dst, g->h_csr_eid);
    const size_t vb = ((size_t)g->nV + 1) * 4, eb = (size_t)g->nE * 4;
    int rc = 0;
    auto up = [&](int32_t *&dptr, const void *src, size_t bytes) {
      if (rc) return;
      rc = gasrt_malloc((void **)&dptr, bytes ? bytes : 4);
      if (!rc) rc = gasrt_memcpy_h2d(dptr, src, bytes, nullptr);
      g->device_bytes += bytes;
    };
    up(g->csc_off, g->h_csc_off.data(), vb);
    up(g->csc_src, g->h_csc_src.data(), eb);
    up(g->csr_off, g->h_csr_off.data(), vb);
    up(g->csr_dst, g->h_csr_dst.data(), eb);
    if (!edges->w.empty()) {
      for (int32_t x : edges->w) g->max_w = std::max(g->max_w, x);
      std::vector<int32_t> wc(g->nE);
      for (int32_t k = 0; k < g->nE; ++k) wc[k] = edges->w[g->h_csc_eid[k]];
      up(g->csc_w, wc.data(), eb);
      for (int32_t k = 0; k < g->nE; ++k) wc[k] = edges->w[g->h_csr_eid[k]];
      up(g->csr_w, wc.data(), eb);
    }
    if (rc) {
      graph_release(g);
      delete g;
      return rc;
    }
    *out = g;
    return GAS_OK;
  });
}

int gas_graph_ell(gas_graph_t g, const int32_t **ell_dst, const int32_t **ell_w) {
  if (!g || !ell_dst || !ell_w) return set_error(GAS_ERR_ARG, "gas_graph_ell: null");
  return gas::host_guard("gas_graph_ell", [&]() -> int {
    std::lock_guard<std::mutex> lock(g->order_mu);
    if (g->ell_state == 0) {
      int32_t max_out = 0;
      for (int32_t v = 0; v < g->nV; ++v)
        max_out = std::max(max_out, g->h_csr_off[(size_t)v + 1] - g->h_csr_off[(size_t)v]);
      g->ell_state = -1;
      if (max_out <= gas::kEllWidth && g->nV > 0) {
        const size_t b = sizeof(int32_t) * gas::kEllWidth * (size_t)g->nV;
        int32_t *dst = nullptr, *w = nullptr;
        gasrt_stream_t s = nullptr;
        int rc = gasrt_malloc((void **)&dst, b);
        if (!rc && g->csr_w) rc = gasrt_malloc((void **)&w, b);
        if (!rc) rc = gasrt_stream_create(&s);
        if (!rc)
          rc = gas::device_build_ell(g->nV, g->csr_off, g->csr_dst, g->csr_w, dst, w, s);
        if (s) {
          const int rs = gasrt_stream_synchronize(s);
          if (!rc) rc = rs;
          gasrt_stream_destroy(s);
        }
        if (rc) {
          gasrt_free(dst);
          gasrt_free(w);
          g->ell_state = 0;
          return rc;
        }
        g->ell_dst = dst;
        g->ell_w = w;
        g->device_bytes += b * (w ? 2 : 1);
        g->ell_state = 1;
      }
    }
    *ell_dst = g->ell_dst;
    *ell_w = g->ell_w;
    return GAS_OK;
  });
}

int gas_graph_view_get(gas_graph_t g, gas_graph_view *v) {
  if (!g || !v) return set_error(GAS_ERR_ARG, "gas_graph_view_get: null");
  v->n_vertices = g->nV;
  v->n_edges = g->nE;
  v->csc_offsets = g->csc_off;
  v->csc_srcs = g->csc_src;
  v->csr_offsets = g->csr_off;
  v->csr_dsts = g->csr_dst;
  v->csc_weights = g->csc_w;
  v->csr_weights = g->csr_w;
  v->device_bytes = g->device_bytes;
  return GAS_OK;
}

int gas_graph_host_csc(gas_graph_t g, int32_t *offsets, int32_t *srcs,
                       int32_t *edge_index) {
  if (!g) return set_error(GAS_ERR_ARG, "gas_graph_host_csc: null");
  if (offsets) std::copy(g->h_csc_off.begin(), g->h_csc_off.end(), offsets);
  if (srcs) std::copy(g->h_csc_src.begin(), g->h_csc_src.end(), srcs);
  if (edge_index)
    std::copy(g->h_csc_eid.begin(), g->h_csc_eid.end(), edge_index);
  return GAS_OK;
}

int gas_graph_host_csr(gas_graph_t g, int32_t *offsets, int32_t *dsts,
                       int32_t *edge_index) {
  if (!g) return set_error(GAS_ERR_ARG, "gas_graph_host_csr: null");
  if (offsets) std::copy(g->h_csr_off.begin(), g->h_csr_off.end(), offsets);
  if (dsts) std::copy(g->h_csr_dst.begin(), g->h_csr_dst.end(), dsts);
  if (edge_index)
    std::copy(g->h_csr_eid.begin(), g->h_csr_eid.end(), edge_index);
  return GAS_OK;
}

int gas_graph_permute_to_csc(gas_graph_t g, const void *in, void *out,
                             size_t elem_bytes) {
  if (!g || (!in && g->nE) || (!out && g->nE) || !elem_bytes)
    return set_error(GAS_ERR_ARG, "gas_graph_permute_to_csc: bad arguments");
  const char *src = static_cast<const char *>(in);
  char *dst = static_cast<char *>(out);
  for (int32_t k = 0; k < g->nE; ++k)
    memcpy(dst + (size_t)k * elem_bytes,
           src + (size_t)g->h_csc_eid[k] * elem_bytes, elem_bytes);
  return GAS_OK;
}

int gas_graph_free(gas_graph_t g) {
  if (!g) return GAS_OK;
  graph_release(g);
  delete g;
  return GAS_OK;
}

}  // extern "C"
