// Batched multi-source traversal: up to 64 BFS or SSSP queries on one
// resident graph share every edge read.  Lane s of a batch is query s; a
// wave64 is exactly one vertex's row of lanes.
//
//   BFS   per-vertex uint64 masks: visited, frontier, next;
//         depth stored as 16 bits per (lane, vertex) (0xffff = unreached),
//         levels up to 65534.  A round is either a push from the frontier list
//         (items of up to 16 out-edges carrying the frontier mask, 4 items
//         per wave and 8 in flight, one lane per edge, 64-bit atomicOr of the
//         new bits) or a pull over every
//         vertex still missing lanes (in-edges with early exit once all
//         missing lanes are found); the direction follows the frontier's
//         out-edge count.
//   SSSP  int32 distance lanes dist[v][64] and a per-vertex mask of lanes
//         that changed in the last round.  A round pushes from the changed
//         list: one wave per item of out-edges, lane = query, candidate
//         dist[u][s] + w compared against dist[w][s] and lowered with
//         atomicMin only where it improves; improved lanes are OR-ed into the
//         destination's next mask.  Distances only decrease, so reading a
//         value another wave is lowering is safe: the lowering wave marks
//         the lane changed and it is pushed again next round.  The fixpoint
//         is the unique shortest-path distance of each lane, unreached lanes
//         keep 2147383647 (sssp.cu:32-33).
//
// Every round ends with a deterministic compaction (two passes, no atomics
// except one per hub) of the vertices whose mask is nonzero, which also cuts
// their out-edges into items; vertices are dealt to workgroups in 64-vertex
// slabs so heavy id ranges spread out, and a hub's items are written by a
// whole workgroup.  Lanes are reset and seeded between rounds, so queries
// can join a running batch (continuous batching) and leave when their lane
// empties.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "engine.hpp"

namespace gas {

constexpr int BW = 64;            // lanes per batch (= wave size)
constexpr int BFS_ITEM_EDGES = 16;   // BFS push: out-edges per item, 4 items per wave
constexpr int SSSP_ITEM_EDGES = 64;  // SSSP push: out-edges per item (one wave)
// Items may be cut shorter (gas_batch_set_item_edges): down to these sizes,
// which size the item and hub lists.
constexpr int BFS_ITEM_EDGES_MIN = 8;
constexpr int SSSP_ITEM_EDGES_MIN = 32;
constexpr int BFS_ITEMS_PER_TRIP = 8;  // BFS push: items in flight per wave
constexpr int HUB_ITEMS = 32;     // a vertex with more items is written by
                                  // a whole workgroup of kb_hub_items (the
                                  // default; gas_batch_set_hub_items)
constexpr int HUB_ITEMS_MIN = 16; // smallest hub threshold (sizes the hub list)
// kb_list rounds whose vertices have at most LIST_FEW push items each write
// them per vertex thread (no hubs among them: LIST_FEW < HUB_ITEMS_MIN)
constexpr int LIST_FEW = 8;
static_assert(LIST_FEW < HUB_ITEMS_MIN, "a few-item round holds no hub");
constexpr int SSSP_UNROLL = 8;    // edges per item per wave trip in the SSSP push
constexpr int SSSP_ITEMS_PER_TRIP = 1;  // SSSP push: items a wave walks together
// Workgroups of the SSSP push and the PPR item pass: items differ widely in
// cost (R-MAT hubs' rows stay hot, others miss), and many short workgroups
// let the dispatcher balance them (R-MAT-22 SSSP batch 39.2 -> 33.3 ms, PPR
// 70.8 -> 64.3 ms against 4 x 1024).
constexpr int PUSH_GRID = 32 * 1024;
// Graphs of average out-degree <= LOW_DEGREE (road-like meshes): list
// vertices of a few out-edges with a few changed lanes each; the SSSP push
// takes (vertex, octet) items (kb_sssp_push_oct) on PUSH_GRID_LOW workgroups.
constexpr int LOW_DEGREE = 8;
constexpr int PUSH_GRID_LOW = 2048;
constexpr int32_t SSSP_INF = INT32_MAX - 100000;  // sssp.cu:32-33

struct BatchCtl {
  int32_t n_list;       // vertices in the current list
  int32_t n_items;      // push items over the list
  int32_t pull;         // BFS: next round pulls
  int32_t n_hubs;       // list vertices whose items kb_hub_items writes
  int64_t list_edges;   // summed out-degree of the list
  uint64_t lanes_alive; // OR of the masks of the current list
  int64_t stat_edges;   // edges examined so far
  int64_t stat_rounds;
  // Round state read by the kernels (not by-value arguments, so a captured
  // graph of rounds replays unchanged): the rounds done so far, moved on by
  // the list pass that ends each round (its workgroup 0, the only reader of
  // the round in that launch), and the lanes in use.
  int32_t round;
  int32_t pad_;
  uint64_t used;
};

struct BatchScan {
  int32_t n_list, n_items, n_hubs, pad_;
  int64_t list_edges;
  uint64_t lanes_alive;
};

struct BatchDev {
  int32_t nV, nE;
  const int32_t *csc_off, *csc_src, *csr_off, *csr_dst, *csr_w;
  const int32_t *new_of;  // caller's vertex -> the engine graph's, or NULL
  BatchCtl *ctl;
  uint64_t *vn;                         // BFS [V][2]: visited, next bits
                                        // (one 16-byte pair per vertex)
  uint64_t *front;                      // BFS frontier masks (the finalize
                                        // pass overwrites them in place:
                                        // push reads the items instead)
  uint16_t *depth16;                    // BFS depths in tiles (depth_at)
  int32_t rowV;                         // V rounded up to DT
  int32_t *lane_start;                  // [BW]: round the lane began
  int32_t *dist;                        // SSSP [V][BW]
  uint64_t *chg, *chg_next;             // SSSP
  int32_t *list;                        // current list (slab order)
  int4 *item;                           // push items: BFS {e0, e1, mask};
                                        // SSSP {u, e0, e1, octet (0 unless
                                        // octets)}
  int32_t item_edges;                   // out-edges per item
  int32_t hub_items;                    // hub threshold in items
  int4 *hub;                            // vertices with > hub_items items:
                                        // {v, first item, e0, e1}
  uint64_t *hub_mask;                   // their masks
  int32_t *bsum;                        // [3][SCAN_BLOCKS] count | items | hubs
  int32_t *bpre;                        // [3][SCAN_BLOCKS] their exclusive prefixes
  struct BatchScan *scan;               // totals of the count pass (kb_scan)
  int64_t *esum;                        // [SCAN_BLOCKS] out-edges
  uint64_t *lor;                        // [SCAN_BLOCKS] OR of masks
  int32_t pull_div;                     // pull when list_edges*div >= E
                                        // (0: never, SSSP)
  int32_t sssp_delta;                   // SSSP distance bound step (0: none)
  int32_t octets;                       // SSSP items per (vertex, octet of
                                        // lanes with a change): see
                                        // kb_sssp_push_oct
  // Activity of 64-vertex slabs (SLAB): a slab whose flag is 0 holds no
  // nonzero mask of that array, so the O(V) passes skip it.  sl_next: some
  // vertex of the slab got bits in next (BFS) / chg_next (SSSP) this round;
  // sl_cur: front (BFS) / chg (SSSP) may be nonzero there.
  uint8_t *sl_next;
  uint8_t *sl_cur;
  // host-mapped ring of BSTATUS (tag, lanes alive) pairs: the list pass of
  // round r stores the lanes still alive after it, then tag r + 1
  // (system-scope stores; null: no report)
  int64_t *status;
};
constexpr int BSTATUS = 16;

// SSSP: a lane relaxes from a vertex in a round only while its distance
// there is below (rounds since the lane began) * delta; changed lanes above
// the bound stay marked for a later round.  delta = 0: no bound
// (Bellman-Ford rounds).
__device__ inline int32_t sssp_bound(const BatchDev &d, int lane, int round) {
  if (d.sssp_delta <= 0) return INT32_MAX;
  const int64_t b =
      (int64_t)(round + 1 - d.lane_start[lane]) * (int64_t)d.sssp_delta;
  return b > INT32_MAX ? INT32_MAX : (int32_t)b;
}

__device__ inline uint64_t wave_or(uint64_t x) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    uint32_t lo = __shfl_xor((uint32_t)x, o, WAVE);
    uint32_t hi = __shfl_xor((uint32_t)(x >> 32), o, WAVE);
    x |= ((uint64_t)hi << 32) | lo;
  }
  return x;
}

__device__ inline uint64_t block_or(uint64_t x, uint64_t *lds) {
  x = wave_or(x);
  if (lane_id() == 0) lds[threadIdx.x / WAVE] = x;
  __syncthreads();
  uint64_t r = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; ++i) r |= lds[i];
  __syncthreads();
  return r;
}

__device__ inline int64_t block_sum64(int64_t x, int64_t *lds) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
  if (lane_id() == 0) lds[threadIdx.x / WAVE] = x;
  __syncthreads();
  int64_t r = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; ++i) r += lds[i];
  __syncthreads();
  return r;
}

// Exclusive workgroup scans of two counters at once (one LDS exchange):
// a (int) and b (int64); totals in ta / tb.
__device__ inline void block_excl_scan2(int a, int64_t b, int &pa, int64_t &pb,
                                        int &ta, int64_t &tb, int *lds,
                                        int64_t *lds64) {
  const int lane = lane_id(), w = threadIdx.x / WAVE;
  int ia = a;
  int64_t ib = b;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const int ya = __shfl_up(ia, o, WAVE);
    const int64_t yb = __shfl_up(ib, o, WAVE);
    if (lane >= o) {
      ia += ya;
      ib += yb;
    }
  }
  if (lane == WAVE - 1) {
    lds[w] = ia;
    lds64[w] = ib;
  }
  __syncthreads();
  int qa = 0, sa = 0;
  int64_t qb = 0, sb = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; ++i) {
    const int xa = lds[i];
    const int64_t xb = lds64[i];
    qa += (i < w) ? xa : 0;
    qb += (i < w) ? xb : 0;
    sa += xa;
    sb += xb;
  }
  __syncthreads();
  pa = qa + ia - a;
  pb = qb + ib - b;
  ta = sa;
  tb = sb;
}

// ---- list construction -------------------------------------------------------
// Vertices of a workgroup in the list passes: 64-vertex slabs dealt
// round-robin to the SCAN_BLOCKS workgroups, NT / 64 slabs per round, so
// runs of heavy vertices (R-MAT's low ids) spread over many workgroups.
// Every pass that counts for kb_list walks the same assignment.
constexpr int SLAB = 64;
__device__ inline int slab_rounds(int nV) {
  const int nslabs = (nV + SLAB - 1) / SLAB;
  const int mine = (nslabs - (int)blockIdx.x + SCAN_BLOCKS - 1) / SCAN_BLOCKS;
  return (mine * SLAB + NT - 1) / NT;
}
__device__ inline int slab_vertex(int r) {
  const int idx = r * NT + (int)threadIdx.x;
  return ((int)blockIdx.x + (idx / SLAB) * SCAN_BLOCKS) * SLAB + idx % SLAB;
}

// BFS depths (uint16, 0xffff unreached) in tiles of DT = SLAB vertices:
// tile t holds lane s's depths of vertices DT t .. DT t + DT - 1 at
// [(t BW + s) DT, + DT), one 128-byte line per (tile, lane).  A lane's
// depths of a slab share a line (extraction reads 2 bytes per vertex and
// lane; the finalize pass's store for one lane writes one line), and the
// finalize pass may instead write a vertex's new lanes from the vertex's
// own thread (as many stores as the most new lanes of a vertex in the
// wave, not one per lane new anywhere in it).
constexpr int DT = SLAB;
__device__ inline size_t depth_at(int v, int s) {
  return ((size_t)(v / DT) * BW + s) * DT + (v % DT);
}

// Octets of a 64-lane mask that hold some set lane, as 8 bits.
__device__ inline uint32_t octet_mask(uint64_t m) {
  uint32_t om = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) om |= ((m >> (8 * i)) & 0xffull) ? 1u << i : 0u;
  return om;
}
__device__ inline int nth_set_bit(uint32_t m, int n) {
  for (int i = 0; i < n; ++i) m &= m - 1;
  return __ffs(m) - 1;
}

// Push items of a vertex with mask m and deg out-edges: chunks of
// item_edges out-edges, each once per octet of m with a set lane when the
// engine has octets.
__device__ inline int item_count(const BatchDev &d, uint64_t m, int deg) {
  const int c = (deg + d.item_edges - 1) / d.item_edges;
  return d.octets ? c * __popc(octet_mask(m)) : c;
}
// Push item k of vertex v (mask m, out-edges [o, o1)): octet-major.
__device__ inline int4 push_item(const BatchDev &d, int v, uint64_t m, int o,
                                 int o1, int k) {
  const int ie = d.item_edges;
  if (!d.dist) {
    const int e0 = o + k * ie;
    return make_int4(e0, min(e0 + ie, o1), (int)(uint32_t)m,
                     (int)(uint32_t)(m >> 32));
  }
  int oct = 0;
  if (d.octets) {
    const int c = (o1 - o + ie - 1) / ie;
    const int oi = k / c;
    k -= oi * c;
    oct = nth_set_bit(octet_mask(m), oi);
  }
  const int e0 = o + k * ie;
  return make_int4(v, e0, min(e0 + ie, o1), oct);
}

// Pass 1 over contiguous chunks of vertices with mask array m: per chunk the
// number of nonzero masks, their push items and out-edges, and the OR.
static __global__ __launch_bounds__(NT) void kb_count(BatchDev d, const uint64_t *m,
                                                       const uint8_t *slm) {
  __shared__ int lds[NT / WAVE];
  __shared__ int64_t lds64[NT / WAVE];
  __shared__ uint64_t ldsu[NT / WAVE];
  int cnt = 0, items = 0, hubs = 0;
  int64_t edges = 0;
  uint64_t any = 0;
  const int nr = slab_rounds(d.nV);
  for (int r = 0; r < nr; ++r) {
    const int v = slab_vertex(r);
    if (v >= d.nV || !slm[v / SLAB]) continue;
    const uint64_t x = m[v];
    if (x) {
      const int deg = d.csr_off[v + 1] - d.csr_off[v];
      ++cnt;
      const int nit = item_count(d, x, deg);
      items += nit;
      hubs += nit > d.hub_items;
      edges += deg;
      any |= x;
    }
  }
  cnt = block_sum(cnt, lds);
  items = block_sum(items, lds);
  hubs = block_sum(hubs, lds);
  edges = block_sum64(edges, lds64);
  any = block_or(any, ldsu);
  if (threadIdx.x == 0) {
    d.bsum[2 * SCAN_BLOCKS + blockIdx.x] = hubs;
    d.bsum[blockIdx.x] = cnt;
    d.bsum[SCAN_BLOCKS + blockIdx.x] = items;
    d.esum[blockIdx.x] = edges;
    d.lor[blockIdx.x] = any;
  }
}

// Between the count pass and the list pass: exclusive prefixes of the
// per-workgroup counts (list slots, items, hubs) and the totals, once, in
// one workgroup of four waves (thread t: chunks 4t .. 4t + 3) -- small
// enough to find room on a busy GPU (the server runs many engines at once).
constexpr int SCAN_NT = 256;
constexpr int SCAN_PER = SCAN_BLOCKS / SCAN_NT;
static_assert(SCAN_PER * SCAN_NT == SCAN_BLOCKS, "scan chunks per thread");
static __global__ __launch_bounds__(SCAN_NT) void kb_scan(BatchDev d) {
  __shared__ int64_t w_a[SCAN_NT / WAVE], w_b[SCAN_NT / WAVE], w_e[SCAN_NT / WAVE];
  __shared__ uint64_t w_o[SCAN_NT / WAVE];
  const int t = threadIdx.x, lane = lane_id(), w = t / WAVE;
  // list slots and hubs in one 64-bit sum (each total < 2^31)
  int64_t ca[SCAN_PER], cb[SCAN_PER];
  int64_t a0 = 0, b0 = 0, e = 0;
  uint64_t o = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    const int i = t * SCAN_PER + k;
    ca[k] = ((int64_t)d.bsum[2 * SCAN_BLOCKS + i] << 32) | (int64_t)d.bsum[i];
    cb[k] = d.bsum[SCAN_BLOCKS + i];
    a0 += ca[k];
    b0 += cb[k];
    e += d.esum[i];
    o |= d.lor[i];
  }
  int64_t a = a0, b = b0;
#pragma unroll
  for (int k = 1; k < WAVE; k <<= 1) {
    const int64_t ya = __shfl_up(a, k, WAVE), yb = __shfl_up(b, k, WAVE);
    if (lane >= k) {
      a += ya;
      b += yb;
    }
  }
#pragma unroll
  for (int k = WAVE / 2; k > 0; k >>= 1) {
    e += __shfl_xor(e, k, WAVE);
    o |= ((uint64_t)__shfl_xor((uint32_t)(o >> 32), k, WAVE) << 32) |
         __shfl_xor((uint32_t)o, k, WAVE);
  }
  if (lane == WAVE - 1) {
    w_a[w] = a;
    w_b[w] = b;
  }
  if (lane == 0) {
    w_e[w] = e;
    w_o[w] = o;
  }
  __syncthreads();
  int64_t pa = 0, pb = 0, ta = 0, tb = 0, te = 0;
  uint64_t to = 0;
  for (int k = 0; k < SCAN_NT / WAVE; ++k) {
    pa += k < w ? w_a[k] : 0;
    pb += k < w ? w_b[k] : 0;
    ta += w_a[k];
    tb += w_b[k];
    te += w_e[k];
    to |= w_o[k];
  }
  // exclusive prefix of this thread's first chunk, then its chunks in order
  int64_t xa = pa + a - a0, xb = pb + b - b0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    const int i = t * SCAN_PER + k;
    d.bpre[i] = (int32_t)(xa & 0xffffffff);
    d.bpre[SCAN_BLOCKS + i] = (int32_t)xb;
    d.bpre[2 * SCAN_BLOCKS + i] = (int32_t)(xa >> 32);
    xa += ca[k];
    xb += cb[k];
  }
  if (t == 0) {
    BatchScan *sc = d.scan;
    sc->n_list = (int32_t)(ta & 0xffffffff);
    sc->n_hubs = (int32_t)(ta >> 32);
    sc->n_items = (int32_t)tb;
    sc->list_edges = te;
    sc->lanes_alive = to;
  }
}

// Pass 2: ordered list of vertices with nonzero masks (the BFS frontier or
// the SSSP changes), item offsets and the owner of every item; workgroup 0
// publishes totals and the direction and, at the end of a round
// (advance = 1), reports the round and moves the round on.
static __global__ __launch_bounds__(NT) void kb_list(BatchDev d, int advance) {
  __shared__ int lds[NT / WAVE];
  __shared__ int64_t lds64[NT / WAVE];
  __shared__ uint64_t ldsu[NT / WAVE];
  __shared__ int32_t s_ipos[NT], s_jpos[NT + 1], s_v[NT], s_o[NT], s_o1[NT];
  __shared__ uint64_t s_m[NT];
  const uint64_t *m = d.dist ? d.chg : d.front;
  const uint8_t *slm = d.sl_cur;
  const BatchScan sc = *d.scan;
  const int pre = d.bpre[blockIdx.x], ipre = d.bpre[SCAN_BLOCKS + blockIdx.x],
            hpre = d.bpre[2 * SCAN_BLOCKS + blockIdx.x];
  const int total = sc.n_list, itotal = sc.n_items, htotal = sc.n_hubs;
  // every workgroup derives the direction from the edge totals; a round
  // that pulls needs no list
  const int64_t ed = sc.list_edges;
  const bool pull = d.pull_div > 0 && ed * (int64_t)d.pull_div >= (int64_t)d.nE;
  int carry = pre, icarry = ipre, hcarry = hpre;
  // a workgroup whose slabs the count pass found empty has nothing to list
  const int nr = (pull || d.bsum[blockIdx.x] == 0) ? 0 : slab_rounds(d.nV);
  for (int r = 0; r < nr; ++r) {
    const int v = slab_vertex(r);
    const bool act = v < d.nV && slm[v / SLAB];
    // a round of the workgroup's slabs with none active adds nothing
    if (!__syncthreads_or(act)) continue;
    const uint64_t mv = act ? m[v] : 0ull;
    const bool on = mv != 0;
    int nit = 0, o = 0, o1 = 0;
    if (on) {
      o = d.csr_off[v];
      o1 = d.csr_off[v + 1];
      nit = item_count(d, mv, o1 - o);
    }
    // one scan for four counters: list slots | hub slots (16 bits each,
    // at most NT per round) and items | non-hub items (32 bits each)
    const bool hub = nit > d.hub_items;
    int pk, tk;
    int64_t pi, ti;
    block_excl_scan2((on ? 1 : 0) | (hub ? 1 << 16 : 0),
                     ((int64_t)nit << 32) | (int64_t)(hub ? 0 : nit), pk, pi, tk,
                     ti, lds, lds64);
    const int tot = tk & 0xffff, htot = tk >> 16;
    const int itot = (int)(ti >> 32), jtot = (int)(ti & 0xffffffff);
    const int pos = carry + (pk & 0xffff);
    const int iex = (int)(pi >> 32), jex = (int)(pi & 0xffffffff);
    if (on) d.list[pos] = v;
    // a round whose vertices have at most LIST_FEW items each (no hub among
    // them): each thread writes its own vertex's items
    if (!__syncthreads_or(nit > LIST_FEW)) {
      for (int k = 0; k < nit; ++k)
        d.item[icarry + iex + k] = push_item(d, v, mv, o, o1, k);
      carry += tot;
      icarry += itot;
      hcarry += htot;
      continue;
    }
    // otherwise the workgroup writes the round's non-hub items together:
    // the j-th of them belongs to the last thread whose exclusive count of
    // non-hub items (s_jpos) is <= j; hubs' items are left to kb_hub_items
    s_ipos[threadIdx.x] = iex;
    s_jpos[threadIdx.x] = jex;
    s_v[threadIdx.x] = v;
    s_o[threadIdx.x] = o;
    s_o1[threadIdx.x] = o1;
    s_m[threadIdx.x] = mv;
    if (threadIdx.x == 0) s_jpos[NT] = jtot;
    const int h = hcarry + (pk >> 16);
    if (nit > d.hub_items) {
      d.hub[h] = make_int4(v, icarry + iex, o, o1);
      d.hub_mask[h] = mv;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < jtot; j += NT) {
      int lo = 0, hi = NT;  // largest t with s_jpos[t] <= j
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_jpos[mid] <= j) lo = mid; else hi = mid;
      }
      const int k = j - s_jpos[lo];
      // BFS items carry the frontier mask, SSSP items the vertex
      d.item[icarry + s_ipos[lo] + k] =
          push_item(d, s_v[lo], s_m[lo], s_o[lo], s_o1[lo], k);
    }
    __syncthreads();
    carry += tot;
    icarry += itot;
    hcarry += htot;
  }
  if (blockIdx.x == 0) {
    const uint64_t any = sc.lanes_alive;
    if (threadIdx.x == 0) {
      BatchCtl *c = d.ctl;
      c->n_list = total;
      c->n_items = itotal;
      c->list_edges = ed;
      c->lanes_alive = any;
      c->pull = pull ? 1 : 0;
      c->n_hubs = pull ? 0 : htotal;
      if (advance) {
        const int32_t round = c->round;
        c->round = round + 1;
        if (d.status) {
          int64_t *e = d.status + 2 * (round & (BSTATUS - 1));
          __hip_atomic_store(e + 1, (int64_t)any, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_SYSTEM);
          __hip_atomic_store(e, (int64_t)round + 1, __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_SYSTEM);
        }
      }
    }
  }
}

// Items of the list's hubs: one workgroup per hub.
static __global__ __launch_bounds__(NT) void kb_hub_items(BatchDev d) {
  const int n = d.ctl->n_hubs;
  for (int h = blockIdx.x; h < n; h += gridDim.x) {
    const int4 hb = d.hub[h];
    const uint64_t mk = d.hub_mask[h];
    const int nit = item_count(d, mk, hb.w - hb.z);
    for (int k = threadIdx.x; k < nit; k += NT)
      d.item[hb.y + k] = push_item(d, hb.x, mk, hb.z, hb.w, k);
  }
}

// ---- BFS ---------------------------------------------------------------------
// Clears lanes `reset` (visited bits, frontier bits, depth rows).
static __global__ __launch_bounds__(NT) void kb_bfs_reset(BatchDev d,
                                                   uint64_t reset) {
  const int stride = gridDim.x * NT;
  for (int v = blockIdx.x * NT + threadIdx.x; v < d.nV; v += stride) {
    d.vn[2 * (size_t)v] &= ~reset;
    d.front[v] &= ~reset;
  }
  // depths of the reset lanes: their sector of every tile, two depths per
  // thread
  uint32_t *w = reinterpret_cast<uint32_t *>(d.depth16);
  const int per = DT / 2, words = d.rowV / DT * per;
  for (int s = 0; s < BW; ++s) {
    if (!((reset >> s) & 1ull)) continue;
    for (int i = blockIdx.x * NT + threadIdx.x; i < words; i += stride)
      w[((size_t)(i / per) * BW + s) * per + i % per] = 0xffffffffu;
  }
}

// One thread per new lane: seed its source.
static __global__ void kb_bfs_seed(BatchDev d, const int32_t *lanes,
                            const int32_t *sources, int n, uint64_t used) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) d.ctl->used = used;
  if (i >= n) return;
  const int s = lanes[i], v = sources[i];
  const uint64_t bit = 1ull << s;
  atomicOr((unsigned long long *)&d.vn[2 * (size_t)v], (unsigned long long)bit);
  atomicOr((unsigned long long *)&d.front[v], (unsigned long long)bit);
  d.sl_cur[v / SLAB] = 1;
  d.depth16[depth_at(v, s)] = 0;
}

// Push: a wave takes BFS_ITEMS_PER_TRIP items at a time, 16 lanes per item
// and one lane per out-edge; all their loads are issued before any use.
// New bits (frontier mask minus the destination's visited bits) are OR-ed
// into the destination's next mask.
__device__ __forceinline__ void bfs_push(const BatchDev &d) {
  const BatchCtl *c = d.ctl;
  const int n_items = c->n_items;
  const uint64_t used = c->used;
  const int lane = lane_id();
  const int sub = lane / BFS_ITEM_EDGES, k = lane % BFS_ITEM_EDGES;
  constexpr int G = BFS_ITEMS_PER_TRIP / (WAVE / BFS_ITEM_EDGES);
  const int waves = gridDim.x * (NT / WAVE);
  const int wid = blockIdx.x * (NT / WAVE) + threadIdx.x / WAVE;
  for (int base = wid * BFS_ITEMS_PER_TRIP; base < n_items;
       base += waves * BFS_ITEMS_PER_TRIP) {
    int4 it[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int i = base + g * (WAVE / BFS_ITEM_EDGES) + sub;
      it[g] = i < n_items ? d.item[i] : make_int4(0, 0, 0, 0);
    }
    int w[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int ei = it[g].x + k;
      w[g] = ei < it[g].y ? d.csr_dst[ei] : -1;
    }
    ulonglong2 vn[G];  // visited, next of the destination: one 16-byte load
#pragma unroll
    for (int g = 0; g < G; ++g)
      vn[g] = w[g] >= 0 ? reinterpret_cast<const ulonglong2 *>(d.vn)[w[g]]
                        : make_ulonglong2(~0ull, ~0ull);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const uint64_t f =
          ((uint64_t)(uint32_t)it[g].w << 32) | (uint32_t)it[g].z;
      // bits already in next (possibly a stale read: then the atomic is
      // merely redundant) need no atomic
      const uint64_t nb = f & used & ~vn[g].x & ~vn[g].y;
      if (nb) {
        atomicOr((unsigned long long *)&d.vn[2 * (size_t)w[g] + 1],
                 (unsigned long long)nb);
        d.sl_next[w[g] / SLAB] = 1;
      }
    }
  }
}
__device__ __forceinline__ void bfs_pull(const BatchDev &d) {
  const BatchCtl *c = d.ctl;
  // a finished lane cannot reach more vertices
  const uint64_t used = c->used & c->lanes_alive;
  const uint64_t *fr = d.front;
  const int lane = lane_id();
  const int stride = gridDim.x * NT;
  const int rounds = (d.nV + stride - 1) / stride;
  for (int r = 0; r < rounds; ++r) {
    const int v = r * stride + blockIdx.x * NT + threadIdx.x;
    uint64_t need = 0;
    int o = 0, deg = 0;
    if (v < d.nV) {
      need = used & ~d.vn[2 * (size_t)v];
      if (need) {
        o = d.csc_off[v];
        deg = d.csc_off[v + 1] - o;
      }
    }
    const bool wide = need && deg > 32;
    if (need && !wide) {
      // in-edges in chunks of 8: all loads of a chunk issued before use
      uint64_t acc = 0;
      for (int k0 = 0; k0 < deg; k0 += 8) {
        int src[8];
#pragma unroll
        for (int x = 0; x < 8; ++x)
          src[x] = k0 + x < deg ? d.csc_src[o + k0 + x] : -1;
        uint64_t f[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) f[x] = src[x] >= 0 ? fr[src[x]] : 0ull;
#pragma unroll
        for (int x = 0; x < 8; ++x) acc |= f[x];
        if ((acc & need) == need) break;
      }
      if (acc & need) {
        d.vn[2 * (size_t)v + 1] = acc & need;
        d.sl_next[v / SLAB] = 1;
      }
    }
    uint64_t m = __ballot(wide);
    while (m) {
      const int j = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const int oj = __shfl(o, j, WAVE);
      const int dj = __shfl(deg, j, WAVE);
      const uint64_t nj =
          ((uint64_t)__shfl((uint32_t)(need >> 32), j, WAVE) << 32) |
          __shfl((uint32_t)need, j, WAVE);
      uint64_t acc = 0;
      for (int base = 0; base < dj; base += WAVE) {
        const int k = base + lane;
        uint64_t x = (k < dj) ? fr[d.csc_src[oj + k]] : 0ull;
        acc |= wave_or(x);
        if ((acc & nj) == nj) break;
      }
      if (lane == j && (acc & nj)) {
        d.vn[2 * (size_t)v + 1] = acc & nj;
        d.sl_next[v / SLAB] = 1;
      }
    }
  }
}

// A round's traversal: push from the list, or (engines that may pull) the
// direction the list pass chose, in one launch.
static __global__ __launch_bounds__(NT) void kb_bfs_push(BatchDev d) {
  if (!d.ctl->pull) bfs_push(d);
}
static __global__ __launch_bounds__(NT) void kb_bfs_step(BatchDev d) {
  if (d.ctl->pull)
    bfs_pull(d);
  else
    bfs_push(d);
}

// New bits become visited and form the next frontier; depth rows get
// round + 1 - lane_start.  Contiguous chunks, with the list
// counters of the new frontier (pass 1 of kb_list).
static __global__ __launch_bounds__(NT) void kb_bfs_finalize(BatchDev d) {
  __shared__ int lds[NT / WAVE];
  __shared__ int64_t lds64[NT / WAVE];
  __shared__ uint64_t ldsu[NT / WAVE];
  __shared__ int32_t s_start[BW];
  if (threadIdx.x < BW) s_start[threadIdx.x] = d.lane_start[threadIdx.x];
  __syncthreads();
  const int round = d.ctl->round;
  const uint64_t used = d.ctl->used;
  uint64_t *nf = d.front;
  uint8_t *slf = d.sl_cur;
  int cnt = 0, items = 0, hubs = 0;
  int64_t edges = 0;
  uint64_t any = 0;
  const int nr = slab_rounds(d.nV);
  for (int r = 0; r < nr; ++r) {
    const int v = slab_vertex(r);
    // a slab (one wave) with no next bits and no stale frontier is left as
    // it is: both arrays are zero there
    const int sl = v / SLAB;
    const bool work = v < d.nV && (d.sl_next[sl] | slf[sl]);
    if (!__any(work)) continue;
    const bool valid = work;
    uint64_t nv = 0;
    if (valid) {
      const ulonglong2 vnv = reinterpret_cast<const ulonglong2 *>(d.vn)[v];
      const uint64_t n = vnv.y;
      if (n) {
        nv = n & used & ~vnv.x;
        reinterpret_cast<ulonglong2 *>(d.vn)[v] = make_ulonglong2(vnv.x | nv, 0ull);
      }
      nf[v] = nv;
    }
    // depths of the newly visited lanes: each thread over its own new lanes
    // when that loop is at most half as long (sparse masks: road-like
    // graphs), else the wave over the lanes new anywhere in it, a store
    // writing one line, that lane's depths of the slab (dense masks: R-MAT)
    const uint64_t wnv = wave_or(nv);
    int most = __popcll(nv);
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) most = max(most, __shfl_xor(most, o, WAVE));
    if (2 * most <= __popcll(wnv)) {
      for (uint64_t x = nv; x; x &= x - 1) {
        const int s = __ffsll((unsigned long long)x) - 1;
        d.depth16[depth_at(v, s)] = (uint16_t)(round + 1 - s_start[s]);
      }
    } else {
      for (uint64_t wx = wnv; wx; wx &= wx - 1) {
        const int s = __ffsll((unsigned long long)wx) - 1;
        if ((nv >> s) & 1ull)
          d.depth16[depth_at(v, s)] = (uint16_t)(round + 1 - s_start[s]);
      }
    }
    if (lane_id() == 0 && v < d.nV) {
      slf[sl] = wnv != 0;
      d.sl_next[sl] = 0;
    }
    if (nv) {
      const int deg = d.csr_off[v + 1] - d.csr_off[v];
      ++cnt;
      const int nit = item_count(d, nv, deg);
      items += nit;
      hubs += nit > d.hub_items;
      edges += deg;
      any |= nv;
    }
  }
  cnt = block_sum(cnt, lds);
  items = block_sum(items, lds);
  hubs = block_sum(hubs, lds);
  edges = block_sum64(edges, lds64);
  any = block_or(any, ldsu);
  if (threadIdx.x == 0) {
    d.bsum[2 * SCAN_BLOCKS + blockIdx.x] = hubs;
    d.bsum[blockIdx.x] = cnt;
    d.bsum[SCAN_BLOCKS + blockIdx.x] = items;
    d.esum[blockIdx.x] = edges;
    d.lor[blockIdx.x] = any;
  }
}

// Depth of lane s as int32 (-1 unreached), for the caller.
static __global__ __launch_bounds__(NT) void kb_bfs_extract(BatchDev d, int s,
                                                     int32_t *out) {
  for (int v = blockIdx.x * NT + threadIdx.x; v < d.nV; v += gridDim.x * NT) {
    const uint16_t x = d.depth16[depth_at(d.new_of ? d.new_of[v] : v, s)];
    out[v] = x == 0xffff ? -1 : (int32_t)x;
  }
}

// ---- SSSP --------------------------------------------------------------------
static __global__ __launch_bounds__(NT) void kb_sssp_reset(BatchDev d,
                                                    uint64_t reset) {
  const int64_t total = (int64_t)d.nV * BW;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < total;
       i += stride) {
    const int s = (int)(i & (BW - 1));
    if ((reset >> s) & 1ull) d.dist[i] = SSSP_INF;
    if (s == 0) d.chg[i / BW] &= ~reset;
  }
}

static __global__ void kb_sssp_seed(BatchDev d, const int32_t *lanes,
                             const int32_t *sources, int n, uint64_t used) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) d.ctl->used = used;
  if (i >= n) return;
  const int s = lanes[i], v = sources[i];
  d.dist[(size_t)v * BW + s] = 0;
  atomicOr((unsigned long long *)&d.chg[v], (unsigned long long)(1ull << s));
  d.sl_cur[v / SLAB] = 1;
}

// Push from the changed list: one wave per item, lane = query.  The item's
// edges are loaded lane-parallel once and broadcast by shuffles; each trip
// issues SSSP_UNROLL distance-row loads before any compare.  A lane is
// marked changed in the destination whenever its candidate beats the value
// it read: a superset of the lanes actually lowered (the read may be older
// than a concurrent atomicMin), which can only cause a redundant push.
static __device__ __forceinline__ void sssp_push(const BatchDev &d) {
  constexpr int U = SSSP_UNROLL;
  const BatchCtl *c = d.ctl;
  const int n_items = c->n_items;
  const int lane = lane_id();
  const int waves = gridDim.x * (NT / WAVE);
  const int32_t bound = sssp_bound(d, lane, c->round);
  constexpr int G = SSSP_ITEMS_PER_TRIP;
  for (int base = (blockIdx.x * (NT / WAVE) + threadIdx.x / WAVE) * G;
       base < n_items; base += waves * G) {
    int u[G], e0[G], ne[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int4 item =
          base + g < n_items ? d.item[base + g] : make_int4(0, 0, 0, 0);
      u[g] = item.x;
      e0[g] = item.y;
      ne[g] = item.z - item.y;
    }
    bool act[G];
    int du[G], my_w[G], my_wt[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const bool changed = ne[g] > 0 && ((d.chg[u[g]] >> lane) & 1ull);
      du[g] = ne[g] > 0 ? d.dist[(size_t)u[g] * BW + lane] : SSSP_INF;
      act[g] = changed && du[g] < bound;
      // changed lanes above the bound wait for a later round
      const uint64_t wait = __ballot(changed && !act[g]);
      if (wait && lane == 0) {
        atomicOr((unsigned long long *)&d.chg_next[u[g]],
                 (unsigned long long)wait);
        d.sl_next[u[g] / SLAB] = 1;
      }
      my_w[g] = 0;
      my_wt[g] = 0;
      if (lane < ne[g]) {
        my_w[g] = d.csr_dst[e0[g] + lane];
        my_wt[g] = d.csr_w[e0[g] + lane];
      }
    }
    int mx = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) mx = max(mx, ne[g]);
    for (int j0 = 0; j0 < mx; j0 += U) {
      int w[G][U], cand[G][U], cur[G][U];
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int x = 0; x < U; ++x) {
          w[g][x] = __shfl(my_w[g], j0 + x, WAVE);
          cand[g][x] = du[g] + __shfl(my_wt[g], j0 + x, WAVE);
        }
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int x = 0; x < U; ++x)
          cur[g][x] = (act[g] && j0 + x < ne[g])
                          ? d.dist[(size_t)w[g][x] * BW + lane]
                          : SSSP_INF;
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int x = 0; x < U; ++x) {
          const bool imp = act[g] && j0 + x < ne[g] && cand[g][x] < cur[g][x];
          if (imp) atomicMin(&d.dist[(size_t)w[g][x] * BW + lane], cand[g][x]);
          const uint64_t bm = __ballot(imp);
          if (bm && lane == 0) {
            atomicOr((unsigned long long *)&d.chg_next[w[g][x]],
                     (unsigned long long)bm);
            d.sl_next[w[g][x] / SLAB] = 1;
          }
        }
    }
  }
}

static __global__ __launch_bounds__(NT) void kb_sssp_push(BatchDev d) {
  sssp_push(d);
}

// Push from octet items (engines with octets: graphs of low degree, whose
// list vertices have a few changed lanes spread over the 64): a wave takes
// 8 items at a time, lanes 8g .. 8g + 7 serve item g's octet (queries
// 8 * octet .. + 7), so a trip relaxes 8 (vertex, octet) pairs where the
// vertex-wide form relaxes one vertex for all 64 lanes.  The trip is three
// dependent loads deep: the item; the change mask, the octet's distances
// (one 32-byte sector, read whatever the mask says) and the first 8
// out-edges (a lane each, walked by shuffles) together; then the
// destinations' sectors.  The next trip's item loads during this one.
// Change marks as in sssp_push, per group: its lanes' bits, shifted to the
// octet.
static __device__ __forceinline__ void sssp_push_oct(const BatchDev &d) {
  const BatchCtl *c = d.ctl;
  const int n_items = c->n_items;
  const int round = c->round;
  const int lane = lane_id();
  const int g = lane >> 3, sub = lane & 7, g0 = lane & ~7;
  const int waves = gridDim.x * (NT / WAVE);
  const int wid = blockIdx.x * (NT / WAVE) + threadIdx.x / WAVE;
  int base = wid * 8;
  int4 next = base + g < n_items ? d.item[base + g] : make_int4(0, 0, 0, 0);
  for (; base < n_items; base += waves * 8) {
    const int4 item = next;
    {
      const int i = base + waves * 8 + g;
      next = i < n_items ? d.item[i] : make_int4(0, 0, 0, 0);
    }
    const int u = item.x, e0 = item.y, ne = item.z - item.y;
    const int sh = 8 * item.w;  // the octet's first lane
    const int q = sh + sub;     // this lane's query
    uint64_t cm = 0;
    int du = SSSP_INF, my_w = 0, my_wt = 0;
    if (ne > 0) {
      cm = d.chg[u];
      du = d.dist[(size_t)u * BW + q];
    }
    if (sub < ne) {
      my_w = d.csr_dst[e0 + sub];
      my_wt = d.csr_w[e0 + sub];
    }
    const bool changed = (cm >> q) & 1ull;
    const bool act = changed && du < sssp_bound(d, q, round);
    // changed lanes above the bound wait for a later round
    const uint64_t wait = __ballot(changed && !act);
    const uint64_t wb = (wait >> (8 * g)) & 0xffull;
    if (sub == 0 && wb) {
      atomicOr((unsigned long long *)&d.chg_next[u],
               (unsigned long long)(wb << sh));
      d.sl_next[u / SLAB] = 1;
    }
    // the wave's longest item
    int mx = ne;
#pragma unroll
    for (int o = 8; o < WAVE; o <<= 1) mx = max(mx, __shfl_xor(mx, o, WAVE));
    for (int c0 = 0; c0 < mx; c0 += 8) {
      if (c0 > 0) {
        my_w = my_wt = 0;
        if (c0 + sub < ne) {
          my_w = d.csr_dst[e0 + c0 + sub];
          my_wt = d.csr_w[e0 + c0 + sub];
        }
      }
      const int nx = min(8, mx - c0);
      for (int x0 = 0; x0 < nx; x0 += 4) {
        int w[4], cand[4], cur[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          w[x] = __shfl(my_w, g0 + x0 + x, WAVE);
          cand[x] = du + __shfl(my_wt, g0 + x0 + x, WAVE);
        }
#pragma unroll
        for (int x = 0; x < 4; ++x)
          cur[x] = (act && c0 + x0 + x < ne) ? d.dist[(size_t)w[x] * BW + q]
                                              : SSSP_INF;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          const bool imp = act && c0 + x0 + x < ne && cand[x] < cur[x];
          if (imp) atomicMin(&d.dist[(size_t)w[x] * BW + q], cand[x]);
          const uint64_t bm = (__ballot(imp) >> (8 * g)) & 0xffull;
          if (sub == 0 && bm) {
            atomicOr((unsigned long long *)&d.chg_next[w[x]],
                     (unsigned long long)(bm << sh));
            d.sl_next[w[x] / SLAB] = 1;
          }
        }
      }
    }
  }
}
static __global__ __launch_bounds__(NT) void kb_sssp_push_oct(BatchDev d) {
  sssp_push_oct(d);
}

// chg <- chg_next, chg_next <- 0, with the list counters (pass 1 of kb_list).
static __global__ __launch_bounds__(NT) void kb_sssp_finalize(BatchDev d) {
  __shared__ int lds[NT / WAVE];
  __shared__ int64_t lds64[NT / WAVE];
  __shared__ uint64_t ldsu[NT / WAVE];
  int cnt = 0, items = 0, hubs = 0;
  int64_t edges = 0;
  uint64_t any = 0;
  const int nr = slab_rounds(d.nV);
  for (int r = 0; r < nr; ++r) {
    const int v = slab_vertex(r);
    // slabs with no new changes and no stale ones are zero in both arrays
    const int sl = v / SLAB;
    const bool work = v < d.nV && (d.sl_next[sl] | d.sl_cur[sl]);
    if (!__any(work)) continue;
    uint64_t x = 0;
    if (work) {
      x = d.chg_next[v];
      if (x) d.chg_next[v] = 0;
      d.chg[v] = x;
    }
    const uint64_t wx = wave_or(x);
    if (lane_id() == 0 && v < d.nV) {
      d.sl_cur[sl] = wx != 0;
      d.sl_next[sl] = 0;
    }
    if (x) {
      const int deg = d.csr_off[v + 1] - d.csr_off[v];
      ++cnt;
      const int nit = item_count(d, x, deg);
      items += nit;
      hubs += nit > d.hub_items;
      edges += deg;
      any |= x;
    }
  }
  cnt = block_sum(cnt, lds);
  items = block_sum(items, lds);
  hubs = block_sum(hubs, lds);
  edges = block_sum64(edges, lds64);
  any = block_or(any, ldsu);
  if (threadIdx.x == 0) {
    d.bsum[2 * SCAN_BLOCKS + blockIdx.x] = hubs;
    d.bsum[blockIdx.x] = cnt;
    d.bsum[SCAN_BLOCKS + blockIdx.x] = items;
    d.esum[blockIdx.x] = edges;
    d.lor[blockIdx.x] = any;
  }
}

// Lanes in use after a release (stream-ordered like the rounds).
static __global__ void kb_set_used(BatchDev d, uint64_t used) {
  if (blockIdx.x == 0 && threadIdx.x == 0) d.ctl->used = used;
}

static __global__ __launch_bounds__(NT) void kb_sssp_extract(BatchDev d, int s,
                                                      int32_t *out) {
  for (int v = blockIdx.x * NT + threadIdx.x; v < d.nV; v += gridDim.x * NT)
    out[v] = d.dist[(size_t)(d.new_of ? d.new_of[v] : v) * BW + s];
}

// Results of every lane in `lanes` at once, lane s into outs.p[s] (V int32).
struct LaneOutputs {
  int32_t *p[BW];
};

// BFS: a lane's depths of DT consecutive vertices share a sector; each
// output is a conversion of those.
static __global__ __launch_bounds__(NT) void kb_bfs_extract_many(
    BatchDev d, uint64_t lanes, LaneOutputs outs) {
  for (int v = blockIdx.x * NT + threadIdx.x; v < d.nV; v += gridDim.x * NT) {
    const int u = d.new_of ? d.new_of[v] : v;
    for (uint64_t m = lanes; m; m &= m - 1) {
      const int s = __ffsll((unsigned long long)m) - 1;
      const uint16_t x = d.depth16[depth_at(u, s)];
      outs.p[s][v] = x == 0xffff ? -1 : (int32_t)x;
    }
  }
}

// SSSP: distances are vertex-major (one 256-byte row of 64 lanes per
// vertex); a workgroup reads 64 rows coalesced into LDS and writes each
// requested lane's 64 consecutive vertices coalesced, so the array is read
// once for any number of lanes.
// Rows hold 64 four-byte lanes: 16 lanes per 64-byte sector.  The lanes of
// the sectors holding some requested lane.
__host__ __device__ inline uint64_t lane_sectors(uint64_t lanes) {
  uint64_t m = 0;
  for (int q = 0; q < 4; ++q)
    if ((lanes >> (16 * q)) & 0xffffull) m |= 0xffffull << (16 * q);
  return m;
}

// Lane-parallel rows of 64 vertices into LDS, only the sectors of requested
// lanes, then each requested lane's 64 consecutive vertices written
// coalesced: the rows are read once for any number of lanes.
template <class T, class Outs>
__device__ inline void extract_rows(const T *rows, const int32_t *new_of,
                                    int32_t nV, uint64_t lanes, const Outs &outs) {
  __shared__ T tile[64][BW + 1];
  const int w = threadIdx.x / WAVE, lane = lane_id();
  const uint64_t sect = lane_sectors(lanes);
  for (int v0 = blockIdx.x * 64; v0 < nV; v0 += gridDim.x * 64) {
    const int nv = min(64, nV - v0);
    for (int i = threadIdx.x; i < 64 * BW; i += NT) {
      const int j = i / BW, s = i % BW;
      if (j < nv && ((sect >> s) & 1ull)) {
        const int u = new_of ? new_of[v0 + j] : v0 + j;
        tile[j][s] = rows[(size_t)u * BW + s];
      }
    }
    __syncthreads();
    int r = 0;
    for (uint64_t m = lanes; m; m &= m - 1, ++r) {
      if (r % (NT / WAVE) != w) continue;
      const int s = __ffsll((unsigned long long)m) - 1;
      if (lane < nv) outs.p[s][v0 + lane] = tile[lane][s];
    }
    __syncthreads();
  }
}

static __global__ __launch_bounds__(NT) void kb_sssp_extract_many(
    BatchDev d, uint64_t lanes, LaneOutputs outs) {
  extract_rows(d.dist, d.new_of, d.nV, lanes, outs);
}

// ---- host side -----------------------------------------------------------------
// One batch of up to 64 lanes of one kind on a resident graph.  Lanes are
// added between rounds (reset + seed + list rebuild) and report themselves
// finished through lanesAlive().
class BatchEngine {
 public:
  enum Kind { BFS_KIND = 0, SSSP_KIND = 1 };

  // On a skewed graph the engine runs on its degree-ordered copy
  // (gas_graph_ordered_view); sources and results stay in the caller's
  // numbering.
  BatchEngine(gas_graph_t caller, Kind kind) : m_kind(kind) {
    check(gasrt_stream_create(&m_stream), "stream");
    check(gasrt_host_alloc((void **)&m_hctl, sizeof(BatchCtl) * 2), "pinned");
    void *dev_status = nullptr;
    check(gasrt_host_alloc_mapped((void **)&m_status, &dev_status,
                                  sizeof(int64_t) * 2 * BSTATUS),
          "mapped status");
    for (int i = 0; i < 2 * BSTATUS; ++i) m_status[i] = 0;
    m_dev.status = static_cast<int64_t *>(dev_status);
    gas_graph_t g = nullptr;
    const int32_t *dnew = nullptr;
    check(gas_graph_ordered_view(caller, &g, &dnew, &m_hnew), "vertex order");
    m_dev.new_of = dnew;
    gas_graph_view gv;
    check(gas_graph_view_get(g, &gv), "graph view");
    if (kind == SSSP_KIND && !gv.csr_weights && gv.n_edges)
      throw Error(GAS_ERR_ARG, "batched SSSP needs edge weights");
    const size_t V = (size_t)gv.n_vertices;
    m_nV = gv.n_vertices;
    m_nE = gv.n_edges;
    BatchDev &d = m_dev;
    d.nV = gv.n_vertices;
    d.nE = gv.n_edges;
    d.csc_off = gv.csc_offsets;
    d.csc_src = gv.csc_srcs;
    d.csr_off = gv.csr_offsets;
    d.csr_dst = gv.csr_dsts;
    d.csr_w = gv.csr_weights;
    d.pull_div = 0;  // push only: see setPullDivisor
    d.item_edges = kind == BFS_KIND ? BFS_ITEM_EDGES : SSSP_ITEM_EDGES;
    m_low_degree = (int64_t)gv.n_edges <= (int64_t)LOW_DEGREE * gv.n_vertices;
    const int ie_min = kind == BFS_KIND ? BFS_ITEM_EDGES_MIN : SSSP_ITEM_EDGES_MIN;
    // at most one item per ie_min out-edges and vertex, 8 times that with
    // octets (whose item counts must stay within int32)
    const int64_t max_items = (int64_t)gv.n_edges / ie_min + (int64_t)V + 1;
    d.octets = kind == SSSP_KIND && m_low_degree && 8 * max_items < INT32_MAX;
    // the hub list holds every vertex that can have more than HUB_ITEMS_MIN
    // items (8 times its chunks with octets)
    int64_t max_hubs = 1;
    {
      // kb_hub_items runs only where some vertex can have > hub_items items
      std::vector<int32_t> off((size_t)gv.n_vertices + 1);
      check(gas_graph_host_csr(g, off.data(), nullptr, nullptr), "csr");
      const int64_t mult = d.octets ? 8 : 1;
      for (int32_t v = 0; v < gv.n_vertices; ++v) {
        const int64_t deg = off[v + 1] - off[v];
        m_max_out = std::max<int64_t>(m_max_out, deg);
        max_hubs += (deg + ie_min - 1) / ie_min * mult > HUB_ITEMS_MIN;
      }
    }
    setHubItems(HUB_ITEMS);
    m_ctl.alloc(1);
    m_list.alloc(V + 1);
    m_item.alloc((size_t)max_items * (d.octets ? 8 : 1));
    m_bsum.alloc(3 * SCAN_BLOCKS);
    m_bpre.alloc(3 * SCAN_BLOCKS);
    m_scan.alloc(1);
    m_esum.alloc(SCAN_BLOCKS);
    m_lor.alloc(SCAN_BLOCKS);
    m_lanes.alloc(BW);
    m_sources.alloc(BW);
    if (kind == BFS_KIND) {
      d.rowV = (int32_t)((V + DT - 1) / DT * DT);
      m_vn.alloc(2 * V);
      m_front.alloc(V);
      m_depth.alloc((size_t)d.rowV * BW);
      check(gasrt_memset(m_vn.p, 0, 2 * V * 8, m_stream), "memset");
      check(gasrt_memset(m_front.p, 0, V * 8, m_stream), "memset");
      check(gasrt_memset(m_depth.p, 0xff, (size_t)d.rowV * BW * 2, m_stream),
            "memset");
      d.vn = m_vn.p;
      d.front = m_front.p;
      d.depth16 = m_depth.p;
    } else {
      m_dist.alloc(V * BW);
      m_chg.alloc(V);
      m_chg_next.alloc(V);
      check(gasrt_memset(m_chg.p, 0, V * 8, m_stream), "memset");
      check(gasrt_memset(m_chg_next.p, 0, V * 8, m_stream), "memset");
      d.dist = m_dist.p;
      d.chg = m_chg.p;
      d.chg_next = m_chg_next.p;
    }
    m_lane_start.alloc(BW);
    m_tmp.alloc(V);  // single-lane extraction (held from the start)
    check(gasrt_memset(m_lane_start.p, 0, BW * 4, m_stream), "memset");
    {
      const size_t nsl = (V + SLAB - 1) / SLAB;
      for (auto *b : {&m_sl_next, &m_sl_cur}) {
        b->alloc(nsl);
        check(gasrt_memset(b->p, 0, std::max<size_t>(nsl, 1), m_stream), "memset");
      }
      d.sl_next = m_sl_next.p;
      d.sl_cur = m_sl_cur.p;
    }
    d.lane_start = m_lane_start.p;
    d.sssp_delta = 0;
    d.ctl = m_ctl.p;
    d.list = m_list.p;
    d.item = m_item.p;
    // a hub has > HUB_ITEMS_MIN items, so at most
    // E / (HUB_ITEMS_MIN * item_edges) of them
    m_hub.alloc((size_t)max_hubs);
    m_hub_mask.alloc((size_t)max_hubs);
    d.hub = m_hub.p;
    d.hub_mask = m_hub_mask.p;
    d.bsum = m_bsum.p;
    d.bpre = m_bpre.p;
    d.scan = m_scan.p;
    d.esum = m_esum.p;
    d.lor = m_lor.p;
    BatchCtl c{};
    check(gasrt_memcpy_h2d(m_ctl.p, &c, sizeof(c), m_stream), "ctl");
    check(gasrt_stream_synchronize(m_stream), "sync");
    // the graph's tuned setting (gas_graph_autotune_batch)
    int32_t param = 0;
    check(gas_graph_batch_param(caller, kind == BFS_KIND ? GAS_BATCH_BFS : GAS_BATCH_SSSP,
                                &param),
          "batch param");
    if (kind == BFS_KIND)
      setPullDivisor(param);
    else
      setSsspDelta(param);
    int32_t hub = 0;
    check(gas_graph_batch_hub_items(caller, kind == BFS_KIND ? GAS_BATCH_BFS : GAS_BATCH_SSSP,
                                    &hub),
          "batch hub items");
    setHubItems(hub);
    int32_t ie = 0;
    check(gas_graph_batch_item_edges(caller, kind == BFS_KIND ? GAS_BATCH_BFS : GAS_BATCH_SSSP,
                                     &ie),
          "batch item edges");
    setItemEdges(ie);
    int32_t graph_rounds = 0;
    check(gas_graph_batch_graph_rounds(
              caller, kind == BFS_KIND ? GAS_BATCH_BFS : GAS_BATCH_SSSP, &graph_rounds),
          "batch graph rounds");
    setGraphRounds(graph_rounds);
  }

  ~BatchEngine() {
    if (m_stream) gasrt_stream_synchronize(m_stream);
    if (m_gexec) gasrt_graph_destroy(m_gexec);
    gasrt_host_free(m_hctl);
    gasrt_host_free((void *)m_status);
    gasrt_stream_destroy(m_stream);
  }

  Kind kind() const { return m_kind; }
  gasrt_stream_t stream() const { return m_stream; }
  uint64_t lanesUsed() const { return m_used; }
  // Device memory the engine holds (lane state for all 64 lanes, lists).
  size_t deviceBytes() const {
    return m_ctl.bytes() + m_item.bytes() + m_list.bytes() + m_bsum.bytes() +
           m_bpre.bytes() + m_scan.bytes() +
           m_lanes.bytes() + m_sources.bytes() + m_lane_start.bytes() + m_dist.bytes() +
           m_tmp.bytes() + m_esum.bytes() + m_lor.bytes() + m_vn.bytes() +
           m_front.bytes() + m_chg.bytes() + m_chg_next.bytes() +
           m_hub_mask.bytes() + m_hub.bytes() + m_depth.bytes() + m_sl_next.bytes() +
           m_sl_cur.bytes();
  }
  int round() const { return m_round; }
  // BFS rounds pull instead of push when the frontier's out-edges reach
  // E / div (0: never, the default).  With 64 lanes a pull cannot stop
  // early while any lane's frontier is still far from the vertex, so on
  // the R-MAT batches measured here every pull round cost more than the
  // push it replaced.
  void setPullDivisor(int div) {
    if (m_kind == BFS_KIND && m_dev.pull_div != div) {
      dropGraph();
      m_dev.pull_div = div;
    }
  }
  // SSSP distance bound step (0: Bellman-Ford rounds).
  void setSsspDelta(int delta) {
    if (m_kind == SSSP_KIND && m_dev.sssp_delta != delta) {
      dropGraph();
      m_dev.sssp_delta = delta;
    }
  }
  // Items above which a list vertex's push items are written by a whole
  // workgroup (kb_hub_items) instead of its list pass thread.
  void setHubItems(int items) {
    if (items < HUB_ITEMS_MIN || items > (1 << 24))
      throw Error(GAS_ERR_ARG, "hub threshold outside [16, 2^24] items");
    if (m_dev.hub_items != items) dropGraph();
    m_dev.hub_items = items;
    updateMayHub();
  }
  // kb_hub_items runs only where some vertex can have more than hub_items
  // items
  void updateMayHub() {
    const int64_t ie = m_dev.item_edges;
    const int64_t most = (m_max_out + ie - 1) / ie * (m_dev.octets ? 8 : 1);
    m_may_hub = most > (int64_t)m_dev.hub_items;
  }

  // Out-edges per push item (a power of two from the kind's minimum up to
  // its default: BFS 8..16, SSSP 32..64); takes effect from the next list.
  void setItemEdges(int edges) {
    const int lo = m_kind == BFS_KIND ? BFS_ITEM_EDGES_MIN : SSSP_ITEM_EDGES_MIN;
    const int hi = m_kind == BFS_KIND ? BFS_ITEM_EDGES : SSSP_ITEM_EDGES;
    if (edges < lo || edges > hi || (edges & (edges - 1)))
      throw Error(GAS_ERR_ARG, "item edges: a power of two from " + std::to_string(lo) +
                                   " to " + std::to_string(hi));
    if (m_dev.item_edges != edges) dropGraph();
    m_dev.item_edges = edges;
    updateMayHub();
  }
  int itemEdges() const { return m_dev.item_edges; }

  // Starts queries: lane[i] runs from source[i].  The lanes must be free.
  void addLanes(const int32_t *lanes, const int32_t *sources, int n) {
    if (n <= 0) return;
    if (n > BW) throw Error(GAS_ERR_ARG, "more than 64 lanes");
    uint64_t reset = 0;
    for (int i = 0; i < n; ++i) {
      if (lanes[i] < 0 || lanes[i] >= BW || sources[i] < 0 ||
          sources[i] >= m_nV)
        throw Error(GAS_ERR_ARG, "lane or source out of range");
      const uint64_t bit = 1ull << lanes[i];
      if ((m_used | reset) & bit) throw Error(GAS_ERR_STATE, "lane in use");
      reset |= bit;
    }
    int32_t mapped[BW];
    for (int i = 0; i < n; ++i) mapped[i] = m_hnew ? m_hnew[sources[i]] : sources[i];
    check(gasrt_memcpy_h2d(m_lanes.p, lanes, 4 * (size_t)n, m_stream), "lanes");
    check(gasrt_memcpy_h2d(m_sources.p, mapped, 4 * (size_t)n, m_stream),
          "sources");
    const unsigned grid = gridFor(m_nV);
    for (int i = 0; i < n; ++i) m_hstart[lanes[i]] = m_round;
    m_alive_hint |= reset;
    m_added_since_snap |= reset;
    check(gasrt_memcpy_h2d(m_lane_start.p, m_hstart, sizeof(m_hstart),
                           m_stream),
          "lane start");
    if (m_kind == BFS_KIND) {
      launch(kb_bfs_reset, grid, m_stream, m_dev, reset);
      launch(kb_bfs_seed, 1, m_stream, m_dev, m_lanes.p, m_sources.p, n,
             m_used | reset);
      m_used |= reset;
      launch(kb_count, SCAN_BLOCKS, m_stream, m_dev,
             (const uint64_t *)m_dev.front, (const uint8_t *)m_dev.sl_cur);
    } else {
      launch(kb_sssp_reset, gridFor((int64_t)m_nV * BW), m_stream, m_dev, reset);
      launch(kb_sssp_seed, 1, m_stream, m_dev, m_lanes.p, m_sources.p, n,
             m_used | reset);
      m_used |= reset;
      launch(kb_count, SCAN_BLOCKS, m_stream, m_dev,
             (const uint64_t *)m_dev.chg, (const uint8_t *)m_dev.sl_cur);
    }
    launch_wg(kb_scan, 1, SCAN_NT, m_stream, m_dev);
    launch(kb_list, SCAN_BLOCKS, m_stream, m_dev, 0);
    if (m_may_hub) launch(kb_hub_items, 256, m_stream, m_dev);
  }

  // Frees lanes (their results stay readable until the lane is reused).
  void releaseLanes(uint64_t lanes) {
    if (!(m_used & lanes)) return;
    m_used &= ~lanes;
    launch(kb_set_used, 1, m_stream, m_dev, m_used);
  }

  // One round for every used lane.
  void step() {
    checkDepth(1);
    enqueueRound();
    ++m_round;
  }

  // Rounds per replay of the captured round graph; 0 launches each round's
  // kernels directly.  The graph holds k consecutive rounds' launches;
  // their arguments do not change between rounds (the round and the lanes
  // in use are read from the control block), so one capture
  // serves every replay until a setting it was captured with changes.
  void setGraphRounds(int k) {
    if (k < 0 || k > kMaxGraphRounds)
      throw Error(GAS_ERR_ARG, "graph rounds outside [0, 8]");
    dropGraph();
    m_graph_rounds = k;
  }
  int graphRounds() const { return m_graph_rounds; }
  static constexpr int kMaxGraphRounds = 8;

  // At least `rounds` rounds: whole replays of the round graph while they
  // fit, single rounds for the rest.
  void advance(int rounds) {
    while (rounds > 0) {
      if (m_graph_rounds > 0 && rounds >= m_graph_rounds) {
        replay();
        rounds -= m_graph_rounds;
      } else {
        step();
        --rounds;
      }
    }
  }

  // `rounds` rounds without waiting; poll() tells when the last of them has
  // reported, aliveSnapshot() reads its report.  Each round's list pass
  // stores its report (tag, lanes alive) into the host-mapped status ring,
  // so a poll reads host memory: no copy or event joins the engine's stream.
  void stepAsync(int rounds = 1) {
    advance(rounds);
    m_added_since_snap = 0;
  }
  bool poll() {
    if (m_round == 0) return true;
    const int64_t *e = m_status + 2 * ((m_round - 1) & (BSTATUS - 1));
    if (__atomic_load_n(e, __ATOMIC_ACQUIRE) == (int64_t)m_round) return true;
    int idle = 0;
    check(gasrt_stream_query(m_stream, &idle), "stream query");
    if (idle && __atomic_load_n(e, __ATOMIC_ACQUIRE) != (int64_t)m_round)
      throw Error(GAS_ERR_STATE, "batch round status was not reported");
    return false;
  }
  // Lanes alive after the last round enqueued (poll() returned true).
  uint64_t aliveSnapshot() {
    const int64_t *e = m_status + 2 * ((m_round - 1) & (BSTATUS - 1));
    const uint64_t alive =
        m_round ? (uint64_t)__atomic_load_n(e + 1, __ATOMIC_RELAXED) : m_alive_hint;
    // lanes added after that round was enqueued are not in it
    m_alive_hint &= alive | m_added_since_snap;
    return alive & m_used;
  }

  // Lanes whose query still has work (reads the device; synchronises).
  uint64_t lanesAlive() {
    check(gasrt_memcpy_d2h(m_hctl, m_ctl.p, sizeof(BatchCtl), m_stream),
          "ctl");
    m_alive_hint = m_hctl[0].lanes_alive;
    return m_hctl[0].lanes_alive & m_used;
  }

  BatchCtl readCtl() {
    check(gasrt_memcpy_d2h(m_hctl, m_ctl.p, sizeof(BatchCtl), m_stream),
          "ctl");
    m_alive_hint = m_hctl[0].lanes_alive;
    return m_hctl[0];
  }

  // Rounds until every used lane is finished.  BFS levels of a batch that
  // started together are at most ~V, so the loop is bounded by the graph.
  // Rounds are enqueued up to kAhead ahead of the host, which reads each
  // round's report from the mapped ring in order and stops at the first
  // with no used lane alive; the rounds enqueued past it find empty lists
  // and change nothing.
  int runToCompletion() {
    if (!lanesAlive()) return 0;
    const int r0 = m_round;
    int next = r0;  // the round whose report is awaited
    for (;;) {
      while (m_round - next < kAhead) advance(std::max(1, m_graph_rounds));
      const int64_t *e = m_status + 2 * (next & (BSTATUS - 1));
      uint64_t alive = 0;
      for (int64_t spin = 0;; ++spin) {
        if (__atomic_load_n(e, __ATOMIC_ACQUIRE) == (int64_t)next + 1) {
          alive = (uint64_t)__atomic_load_n(e + 1, __ATOMIC_RELAXED);
          break;
        }
        if ((spin & 255) == 255) {
          int idle = 0;
          check(gasrt_stream_query(m_stream, &idle), "stream query");
          if (idle && __atomic_load_n(e, __ATOMIC_ACQUIRE) != (int64_t)next + 1)
            throw Error(GAS_ERR_STATE, "batch round status was not reported");
        }
        __builtin_ia32_pause();
      }
      if (!(alive & m_used)) {
        m_alive_hint = alive;
        return next - r0 + 1;
      }
      ++next;
    }
  }
  static constexpr int kAhead = 3;
  // rounds enqueued past the awaited report stay within the status ring
  static_assert(kAhead + kMaxGraphRounds < BSTATUS, "status ring too small");

  // Result of a lane into a device buffer of V int32 (BFS depth, -1
  // unreached; SSSP distance, 2147383647 unreached).
  void laneResultDevice(int lane, int32_t *dev_out) {
    if (lane < 0 || lane >= BW) throw Error(GAS_ERR_ARG, "lane out of range");
    if (m_kind == BFS_KIND)
      launch(kb_bfs_extract, gridFor(m_nV), m_stream, m_dev, lane, dev_out);
    else
      launch(kb_sssp_extract, gridFor(m_nV), m_stream, m_dev, lane, dev_out);
  }

  // Results of every lane in `lanes` into outs[lane] (device, V int32
  // each), in one pass over the lane state.
  void laneResultsDevice(uint64_t lanes, int32_t *const *outs) {
    if (!lanes) return;
    LaneOutputs o{};
    for (uint64_t m = lanes; m; m &= m - 1) {
      const int s = __builtin_ctzll(m);
      if (!outs[s]) throw Error(GAS_ERR_ARG, "null lane output");
      o.p[s] = outs[s];
    }
    if (m_kind == BFS_KIND)
      launch(kb_bfs_extract_many, gridFor(m_nV), m_stream, m_dev, lanes, o);
    else
      launch(kb_sssp_extract_many,
             (unsigned)std::min<int64_t>(2048, (m_nV + 63) / 64), m_stream,
             m_dev, lanes, o);
  }

  void laneResult(int lane, int32_t *host_out) {
    if (m_tmp.n < (size_t)m_nV) m_tmp.alloc((size_t)m_nV);
    laneResultDevice(lane, m_tmp.p);
    check(gasrt_memcpy_d2h(host_out, m_tmp.p, 4 * (size_t)m_nV, m_stream),
          "lane result");
  }

  void sync() { check(gasrt_stream_synchronize(m_stream), "sync"); }

 private:
  static unsigned gridFor(int64_t n) {
    return (unsigned)std::min<int64_t>(2048, std::max<int64_t>(1, (n + NT - 1) / NT));
  }

  // Depths are 16-bit per lane: only lanes still running can outgrow them
  // in the next `rounds` rounds (a finished lane kept for its result does
  // not); the host's view of which lanes run is refreshed when one
  // reaches the limit.
  void checkDepth(int rounds) {
    if (m_kind != BFS_KIND) return;
    auto too_deep = [&](uint64_t lanes) {
      for (int s = 0; s < BW; ++s)
        if (((lanes >> s) & 1ull) && m_round + rounds - m_hstart[s] > 65534) return true;
      return false;
    };
    if (too_deep(m_used & m_alive_hint) && too_deep(lanesAlive()))
      throw Error(GAS_ERR_STATE, "batched BFS deeper than 65534 levels");
  }

  // The launches of one round (what the round graph captures).
  void enqueueRound() {
    if (m_kind == BFS_KIND) {
      if (m_dev.pull_div > 0)
        launch(kb_bfs_step, 4 * SCAN_BLOCKS, m_stream, m_dev);
      else
        launch(kb_bfs_push, 4 * SCAN_BLOCKS, m_stream, m_dev);
      launch(kb_bfs_finalize, SCAN_BLOCKS, m_stream, m_dev);
    } else {
      if (m_dev.octets)
        launch(kb_sssp_push_oct, PUSH_GRID_LOW, m_stream, m_dev);
      else
        launch(kb_sssp_push, PUSH_GRID, m_stream, m_dev);
      launch(kb_sssp_finalize, SCAN_BLOCKS, m_stream, m_dev);
    }
    launch_wg(kb_scan, 1, SCAN_NT, m_stream, m_dev);
    launch(kb_list, SCAN_BLOCKS, m_stream, m_dev, 1);
    if (m_may_hub) launch(kb_hub_items, 256, m_stream, m_dev);
  }

  // m_graph_rounds rounds by one launch of the captured graph (captured at
  // first use).
  void replay() {
    checkDepth(m_graph_rounds);
    if (!m_gexec) {
      check(gasrt_capture_begin(m_stream), "capture");
      try {
        for (int i = 0; i < m_graph_rounds; ++i) enqueueRound();
      } catch (...) {
        gasrt_graph_exec_t x = nullptr;
        if (gasrt_capture_end(m_stream, &x) == GAS_OK) gasrt_graph_destroy(x);
        throw;
      }
      check(gasrt_capture_end(m_stream, &m_gexec), "capture end");
    }
    check(gasrt_graph_launch(m_gexec, m_stream), "graph launch");
    m_round += m_graph_rounds;
  }

  void dropGraph() {
    if (!m_gexec) return;
    check(gasrt_stream_synchronize(m_stream), "sync");
    check(gasrt_graph_destroy(m_gexec), "graph destroy");
    m_gexec = nullptr;
  }

  Kind m_kind;
  gasrt_stream_t m_stream = nullptr;
  gasrt_graph_exec_t m_gexec = nullptr;  // captured rounds (m_graph_rounds)
  int m_graph_rounds = 0;
  BatchCtl *m_hctl = nullptr;  // [0] synchronous reads, [1] snapshots
  int32_t m_nV = 0, m_nE = 0;
  const int32_t *m_hnew = nullptr;  // caller's vertex -> engine graph's (host)
  int m_round = 0;  // rounds enqueued; the control block's round follows it
  int64_t *m_status = nullptr;  // mapped round reports (BSTATUS pairs)
  uint64_t m_used = 0;
  int32_t m_hstart[BW] = {};
  uint64_t m_alive_hint = 0;        // lanes not yet seen finished by the host
  uint64_t m_added_since_snap = 0;  // lanes added after the last stepAsync
  BatchDev m_dev{};
  DevBuf<BatchCtl> m_ctl;
  DevBuf<int4> m_item;
  DevBuf<BatchScan> m_scan;
  DevBuf<int32_t> m_list, m_bsum, m_bpre, m_lanes, m_sources,
      m_lane_start, m_dist, m_tmp;
  DevBuf<int64_t> m_esum;
  DevBuf<uint64_t> m_lor, m_vn, m_front, m_chg, m_chg_next,
      m_hub_mask;
  DevBuf<int4> m_hub;
  DevBuf<uint16_t> m_depth;
  DevBuf<uint8_t> m_sl_next, m_sl_cur;  // slab activity flags
  bool m_may_hub = true;  // some vertex has more than hub_items items
  int64_t m_max_out = 0;  // largest out-degree
  bool m_low_degree = false;  // average out-degree <= LOW_DEGREE
};

// ---- personalised PageRank, 64 FP32 lanes ----------------------------------
// r_0 = 0.15 e_s, r_{k+1}(v) = 0.15 [v == s] + 0.85 sum_{u->v} r_k(u)/outdeg(u)
// per lane s.  Every vertex owns >= 1 fixed item of at most PPR_ITEM in-edges;
// one wave per item walks them in CSC order, reading a source's row (the
// running lanes) only where the source's byte says some lane is nonzero,
// with a compensated sum per lane: a lane adds +0.0 for a source whose own
// value is zero, which leaves its sum and compensation unchanged.  Vertices with several items fold their partials in item order.
// The additions a lane performs depend on that lane alone, so a query's
// result is bit-identical alone and inside any mix.
constexpr int PPR_ITEM = 256;
constexpr int PPR_UNROLL = 8;

struct PprDev {
  int32_t nV, nE;
  const int32_t *csc_off, *csc_src;
  const int32_t *outdeg;          // [V]
  const int32_t *itv, *ite0;      // item -> vertex, first in-edge
  const int32_t *pslot;           // item -> partial slot, -1: single item
  const int32_t *vitem0;          // [V+1] first item of each vertex
  const int32_t *mv;              // vertices with several items
  int32_t n_items, n_multi;
  float *rank[2];                 // [V][BW]
  float *contrib[2];              // [V][BW] rank / outdeg, what gathers read
  uint8_t *nz[3];                 // [V] some running lane is nonzero, by
                                  // sweep mod 3 (read / written / cleared)
  int32_t rot;                    // sweep mod 3: nz[rot] is this sweep's input
  int32_t *nzc;                   // [3] bytes set in nz[q] (vertices written
                                  // nonzero by a sweep; seeds not counted)
  float *part_s, *part_c;         // [slots][BW]
  const int32_t *src_lane;        // [BW]
  const int32_t *new_of;          // caller's vertex -> the engine graph's, or NULL
};

__device__ inline bool ppr_write(const PprDev &d, int p, int v, uint64_t run,
                                 float sum) {
  const int lane = lane_id();
  const bool on = (run >> lane) & 1ull;
  const float tele = (on && d.src_lane[lane] == v) ? 0.15f : 0.0f;
  const float val = on ? tele + 0.85f * sum : 0.0f;
  const int od = d.outdeg[v];
  d.rank[p ^ 1][(size_t)v * BW + lane] = val;
  d.contrib[p ^ 1][(size_t)v * BW + lane] = od > 0 ? val / (float)od : 0.0f;
  const uint64_t m = __ballot(val != 0.0f);
  if (lane == 0 && m) d.nz[(d.rot + 1) % 3][v] = 1;
  return m != 0;
}

static __device__ __forceinline__ void ppr_items(const PprDev &d, int p, uint64_t run) {
  const int lane = lane_id();
  const int waves = gridDim.x * (NT / WAVE);
  const float *ct = d.contrib[p];
  const uint8_t *nzin = d.nz[d.rot];
  // an input with most rows nonzero is read without the byte test (a zero
  // row adds +0.0 to a lane: its sum and compensation stay as they are)
  const bool dense = 2 * (int64_t)d.nzc[d.rot] > (int64_t)d.nV;
  int written = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) d.nzc[(d.rot + 2) % 3] = 0;
  // the byte array the next sweep writes, cleared while this one reads nz[rot]
  {
    uint32_t *clr = reinterpret_cast<uint32_t *>(d.nz[(d.rot + 2) % 3]);
    const int words = (d.nV + 3) / 4;
    for (int i = blockIdx.x * NT + threadIdx.x; i < words; i += gridDim.x * NT)
      clr[i] = 0u;
  }
  for (int it = blockIdx.x * (NT / WAVE) + threadIdx.x / WAVE; it < d.n_items;
       it += waves) {
    const int v = d.itv[it];
    const int e0 = d.ite0[it];
    const int e1 = min(e0 + PPR_ITEM, d.csc_off[v + 1]);
    float s = 0.f, c = 0.f;
    for (int base = e0; base < e1; base += WAVE) {
      const int k = base + lane;
      int u_l = 0;
      bool f_l = false;
      if (k < e1) {
        u_l = d.csc_src[k];
        f_l = dense || nzin[u_l] != 0;
      }
      uint64_t nz = __ballot(f_l);
      const bool mine = (run >> lane) & 1ull;
      while (nz) {
        int u[PPR_UNROLL];
        bool on[PPR_UNROLL];
        float r[PPR_UNROLL];
#pragma unroll
        for (int x = 0; x < PPR_UNROLL; ++x) {
          const int j = nz ? __ffsll((unsigned long long)nz) - 1 : 0;
          const bool have = nz != 0;
          nz &= nz - 1;
          u[x] = __shfl(u_l, j, WAVE);
          on[x] = have && mine;
        }
#pragma unroll
        for (int x = 0; x < PPR_UNROLL; ++x)
          r[x] = on[x] ? ct[(size_t)u[x] * BW + lane] : 0.f;
#pragma unroll
        for (int x = 0; x < PPR_UNROLL; ++x) {
          if (!on[x]) continue;
          const float y = r[x];  // rank(u) / outdeg(u), divided once by the writer
          const float t = s + y;
          const float bp = t - s;
          c += (s - (t - bp)) + (y - bp);
          s = t;
        }
      }
    }
    const int ps = d.pslot[it];
    if (ps < 0) {
      written += ppr_write(d, p, v, run, s + c);
    } else {
      d.part_s[(size_t)ps * BW + lane] = s;
      d.part_c[(size_t)ps * BW + lane] = c;
    }
  }
  if (lane_id() == 0 && written) atomicAdd(&d.nzc[(d.rot + 1) % 3], written);
}

static __global__ __launch_bounds__(NT) void kp_item(PprDev d, int p, uint64_t run) {
  ppr_items(d, p, run);
}

static __global__ __launch_bounds__(NT) void kp_fix(PprDev d, int p, uint64_t run) {
  const int lane = lane_id();
  const int waves = gridDim.x * (NT / WAVE);
  int written = 0;
  for (int i = blockIdx.x * (NT / WAVE) + threadIdx.x / WAVE; i < d.n_multi;
       i += waves) {
    const int v = d.mv[i];
    float s = 0.f, c = 0.f;
    for (int it = d.vitem0[v]; it < d.vitem0[v + 1]; ++it) {
      const int ps = d.pslot[it];
      const float bs = d.part_s[(size_t)ps * BW + lane];
      const float bc = d.part_c[(size_t)ps * BW + lane];
      const float t = s + bs;
      const float bp = t - s;
      c = c + bc + ((s - (t - bp)) + (bs - bp));
      s = t;
    }
    written += ppr_write(d, p, v, run, s + c);
  }
  if (lane == 0 && written) atomicAdd(&d.nzc[(d.rot + 1) % 3], written);
}

static __global__ __launch_bounds__(NT) void kp_reset(PprDev d, int p,
                                               uint64_t reset) {
  const int64_t total = (int64_t)d.nV * BW;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < total;
       i += stride) {
    const int s = (int)(i & (BW - 1));
    if ((reset >> s) & 1ull) {
      d.rank[p][i] = 0.f;
      d.contrib[p][i] = 0.f;
    }
  }
}

static __global__ void kp_seed(PprDev d, int p, const int32_t *lanes,
                        const int32_t *sources, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = lanes[i], v = sources[i];
  d.rank[p][(size_t)v * BW + s] = 0.15f;
  const int od = d.outdeg[v];
  d.contrib[p][(size_t)v * BW + s] = od > 0 ? 0.15f / (float)od : 0.0f;
  d.nz[d.rot][v] = 1;
}

static __global__ __launch_bounds__(NT) void kp_extract(PprDev d, int p, int s,
                                                 float *out) {
  for (int v = blockIdx.x * NT + threadIdx.x; v < d.nV; v += gridDim.x * NT)
    out[v] = d.rank[p][(size_t)(d.new_of ? d.new_of[v] : v) * BW + s];
}

struct LaneOutputsF {
  float *p[BW];
};

// Ranks of every lane in `lanes`, lane s into outs.p[s], in one pass.
static __global__ __launch_bounds__(NT) void kp_extract_many(PprDev d, int p,
                                                      uint64_t lanes,
                                                      LaneOutputsF outs) {
  extract_rows(d.rank[p], d.new_of, d.nV, lanes, outs);
}

class PprBatch {
 public:
  PprBatch(gas_graph_t caller, int iterations) : m_iters(iterations) {
    if (iterations < 1) throw Error(GAS_ERR_ARG, "PPR needs >= 1 iteration");
    check(gasrt_stream_create(&m_stream), "stream");
    gas_graph_t g = nullptr;
    const int32_t *dnew = nullptr;
    check(gas_graph_ordered_view(caller, &g, &dnew, &m_hnew), "vertex order");
    gas_graph_view gv;
    check(gas_graph_view_get(g, &gv), "graph view");
    m_nV = gv.n_vertices;
    const size_t V = (size_t)m_nV;
    std::vector<int32_t> cscoff(V + 1), csroff(V + 1);
    check(gas_graph_host_csc(g, cscoff.data(), nullptr, nullptr), "csc");
    check(gas_graph_host_csr(g, csroff.data(), nullptr, nullptr), "csr");
    std::vector<int32_t> outdeg(V), itv, ite0, pslot, vitem0(V + 1), mv;
    int32_t slots = 0;
    for (size_t v = 0; v < V; ++v) {
      outdeg[v] = csroff[v + 1] - csroff[v];
      const int32_t deg = cscoff[v + 1] - cscoff[v];
      const int32_t nit = deg == 0 ? 1 : (deg + PPR_ITEM - 1) / PPR_ITEM;
      vitem0[v] = (int32_t)itv.size();
      if (nit > 1) mv.push_back((int32_t)v);
      for (int32_t k = 0; k < nit; ++k) {
        itv.push_back((int32_t)v);
        ite0.push_back(cscoff[v] + k * PPR_ITEM);
        pslot.push_back(nit > 1 ? slots++ : -1);
      }
    }
    vitem0[V] = (int32_t)itv.size();
    m_outdeg.upload(outdeg, m_stream);
    m_itv.upload(itv, m_stream);
    m_ite0.upload(ite0, m_stream);
    m_pslot.upload(pslot, m_stream);
    m_vitem0.upload(vitem0, m_stream);
    m_mv.upload(mv, m_stream);
    m_rank0.alloc(V * BW);
    m_rank1.alloc(V * BW);
    m_contrib0.alloc(V * BW);
    m_contrib1.alloc(V * BW);
    for (auto &b : m_nz) b.alloc((V + 3) / 4 * 4);
    m_ps.alloc((size_t)std::max(slots, 1) * BW);
    m_pc.alloc((size_t)std::max(slots, 1) * BW);
    m_src.alloc(BW);
    m_lanes.alloc(BW);
    m_sources.alloc(BW);
    m_tmp.alloc(V);  // single-lane extraction (held from the start)
    check(gasrt_memset(m_rank0.p, 0, V * BW * 4, m_stream), "memset");
    check(gasrt_memset(m_rank1.p, 0, V * BW * 4, m_stream), "memset");
    check(gasrt_memset(m_contrib0.p, 0, V * BW * 4, m_stream), "memset");
    check(gasrt_memset(m_contrib1.p, 0, V * BW * 4, m_stream), "memset");
    for (auto &b : m_nz)
      check(gasrt_memset(b.p, 0, (V + 3) / 4 * 4, m_stream), "memset");
    for (int i = 0; i < BW; ++i) m_hsrc[i] = -1;
    check(gasrt_memcpy_h2d(m_src.p, m_hsrc, sizeof(m_hsrc), m_stream), "src");
    PprDev &d = m_dev;
    d.nV = m_nV;
    d.nE = gv.n_edges;
    d.new_of = dnew;
    d.csc_off = gv.csc_offsets;
    d.csc_src = gv.csc_srcs;
    d.outdeg = m_outdeg.p;
    d.itv = m_itv.p;
    d.ite0 = m_ite0.p;
    d.pslot = m_pslot.p;
    d.vitem0 = m_vitem0.p;
    d.mv = m_mv.p;
    d.n_items = (int32_t)itv.size();
    d.n_multi = (int32_t)mv.size();
    d.rank[0] = m_rank0.p;
    d.rank[1] = m_rank1.p;
    d.contrib[0] = m_contrib0.p;
    d.contrib[1] = m_contrib1.p;
    for (int i = 0; i < 3; ++i) d.nz[i] = m_nz[i].p;
    m_nzc.alloc(3);
    check(gasrt_memset(m_nzc.p, 0, 3 * sizeof(int32_t), m_stream), "memset");
    d.nzc = m_nzc.p;
    d.rot = 0;
    d.part_s = m_ps.p;
    d.part_c = m_pc.p;
    d.src_lane = m_src.p;
    check(gasrt_stream_synchronize(m_stream), "sync");
  }

  ~PprBatch() {
    if (m_stream) gasrt_stream_synchronize(m_stream);
    if (m_ev) gasrt_event_destroy(m_ev);
    gasrt_stream_destroy(m_stream);
  }

  gasrt_stream_t stream() const { return m_stream; }
  // Device memory the engine holds (lane state for all 64 lanes, items).
  size_t deviceBytes() const {
    size_t b = m_outdeg.bytes() + m_itv.bytes() + m_ite0.bytes() + m_pslot.bytes() +
               m_vitem0.bytes() + m_mv.bytes() + m_src.bytes() + m_lanes.bytes() +
               m_sources.bytes() + m_rank0.bytes() + m_rank1.bytes() +
               m_contrib0.bytes() + m_contrib1.bytes() + m_ps.bytes() + m_pc.bytes() +
               m_tmp.bytes();
    for (const auto &z : m_nz) b += z.bytes();
    return b + m_nzc.bytes();
  }
  uint64_t lanesUsed() const { return m_used; }
  int iterations() const { return m_iters; }

  void addLanes(const int32_t *lanes, const int32_t *sources, int n) {
    if (n <= 0) return;
    if (n > BW) throw Error(GAS_ERR_ARG, "more than 64 lanes");
    uint64_t reset = 0;
    for (int i = 0; i < n; ++i) {
      if (lanes[i] < 0 || lanes[i] >= BW || sources[i] < 0 ||
          sources[i] >= m_nV)
        throw Error(GAS_ERR_ARG, "lane or source out of range");
      const uint64_t bit = 1ull << lanes[i];
      if ((m_used | reset) & bit) throw Error(GAS_ERR_STATE, "lane in use");
      reset |= bit;
      m_hsrc[lanes[i]] = m_hnew ? m_hnew[sources[i]] : sources[i];
      m_done[lanes[i]] = 0;
    }
    int32_t mapped[BW];
    for (int i = 0; i < n; ++i) mapped[i] = m_hnew ? m_hnew[sources[i]] : sources[i];
    check(gasrt_memcpy_h2d(m_src.p, m_hsrc, sizeof(m_hsrc), m_stream), "src");
    check(gasrt_memcpy_h2d(m_lanes.p, lanes, 4 * (size_t)n, m_stream), "lanes");
    check(gasrt_memcpy_h2d(m_sources.p, mapped, 4 * (size_t)n, m_stream),
          "sources");
    launch(kp_reset, gridFor((int64_t)m_nV * BW), m_stream, m_dev, m_par, reset);
    launch(kp_seed, 1, m_stream, m_dev, m_par, m_lanes.p, m_sources.p, n);
    m_used |= reset;
  }

  // Lanes whose query has iterations left (host bookkeeping only).
  uint64_t lanesAlive() const {
    uint64_t a = 0;
    for (int s = 0; s < BW; ++s)
      if (((m_used >> s) & 1ull) && m_done[s] < m_iters) a |= 1ull << s;
    return a;
  }

  void step() {
    const uint64_t run = lanesAlive();
    if (!run) return;
    launch(kp_item, PUSH_GRID, m_stream, m_dev, m_par, run);
    if (m_dev.n_multi)
      launch(kp_fix, SCAN_BLOCKS, m_stream, m_dev, m_par, run);
    m_par ^= 1;
    m_dev.rot = (m_dev.rot + 1) % 3;
    for (int s = 0; s < BW; ++s)
      if ((run >> s) & 1ull) ++m_done[s];
  }

  // Rounds until the first running lane finishes (a finished lane must be
  // extracted before the next round overwrites its buffer).
  int roundsToFirstFinish() const {
    int r = 1 << 30;
    for (int s = 0; s < BW; ++s)
      if (((m_used >> s) & 1ull) && m_done[s] < m_iters)
        r = std::min(r, m_iters - m_done[s]);
    return r == (1 << 30) ? 1 : r;
  }
  void stepAsync(int rounds = 1) {
    rounds = std::max(1, std::min(rounds, roundsToFirstFinish()));
    for (int i = 0; i < rounds; ++i) step();
    if (!m_ev) check(gasrt_event_create(&m_ev), "event");
    check(gasrt_event_record(m_ev, m_stream), "event");
  }
  bool poll() {
    int done = 1;
    if (m_ev) check(gasrt_event_query(m_ev, &done), "event query");
    return done != 0;
  }

  // Finished lanes must be read out before the next step(): a step leaves
  // the rows of lanes that are not running at zero.
  void releaseLanes(uint64_t lanes) { m_used &= ~lanes; }

  int runToCompletion() {
    int rounds = 0;
    while (lanesAlive()) {
      step();
      ++rounds;
    }
    return rounds;
  }

  void laneResultDevice(int lane, float *dev_out) {
    if (lane < 0 || lane >= BW) throw Error(GAS_ERR_ARG, "lane out of range");
    launch(kp_extract, gridFor(m_nV), m_stream, m_dev, m_par, lane, dev_out);
  }

  // Ranks of every lane in `lanes` into outs[lane] (device, V floats
  // each), in one pass over the rank rows.
  void laneResultsDevice(uint64_t lanes, float *const *outs) {
    if (!lanes) return;
    LaneOutputsF o{};
    for (uint64_t m = lanes; m; m &= m - 1) {
      const int s = __builtin_ctzll(m);
      if (!outs[s]) throw Error(GAS_ERR_ARG, "null lane output");
      o.p[s] = outs[s];
    }
    launch(kp_extract_many, (unsigned)std::min<int64_t>(2048, (m_nV + 63) / 64),
           m_stream, m_dev, m_par, lanes, o);
  }

  void laneResult(int lane, float *host_out) {
    if (m_tmp.n < (size_t)m_nV) m_tmp.alloc((size_t)m_nV);
    laneResultDevice(lane, m_tmp.p);
    check(gasrt_memcpy_d2h(host_out, m_tmp.p, 4 * (size_t)m_nV, m_stream),
          "lane result");
  }

  void sync() { check(gasrt_stream_synchronize(m_stream), "sync"); }

 private:
  static unsigned gridFor(int64_t n) {
    return (unsigned)std::min<int64_t>(2048, std::max<int64_t>(1, (n + NT - 1) / NT));
  }
  int m_iters;
  gasrt_stream_t m_stream = nullptr;
  gasrt_event_t m_ev = nullptr;
  int32_t m_nV = 0;
  const int32_t *m_hnew = nullptr;  // caller's vertex -> engine graph's (host)
  int m_par = 0;
  uint64_t m_used = 0;
  int32_t m_hsrc[BW];
  int m_done[BW] = {};
  PprDev m_dev{};
  DevBuf<int32_t> m_outdeg, m_itv, m_ite0, m_pslot, m_vitem0, m_mv, m_src,
      m_lanes, m_sources;
  DevBuf<float> m_rank0, m_rank1, m_contrib0, m_contrib1, m_ps, m_pc, m_tmp;
  DevBuf<uint8_t> m_nz[3];
  DevBuf<int32_t> m_nzc;
};

}  // namespace gas
