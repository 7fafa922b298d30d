// gfx950 kernels of the GAS superstep, templated on the vertex program.
//
// One superstep of the reference engine (gpugas.h:421-751, semantics of
// refgas.h:143-210) runs here as a short fixed sequence of kernels whose
// sizes all come from device memory, so the host never waits for a frontier
// size before enqueuing the next superstep.  Supersteps alternate parity
// p = 0, 1 (input list q[p] / activation bytes fl[p], next ones q[p^1] /
// fl[p^1]); per-superstep counters live in a ring of four entries indexed by
// the superstep number, so no kernel reads a counter the same launch writes
// and no kernel needs to know it is the last workgroup to finish.
//
// The frontier of a superstep is one of
//   all      every vertex (setActive(0, V));
//   list     the ids in q[p] -- built by the previous superstep's scatter,
//            which claims each destination with an atomicOr on its
//            activation byte and appends the winners (O(|frontier|) work);
//   bitmap   the activation bytes fl[p] -- written by the previous
//            superstep's scatter with plain stores (no atomics), scanned
//            in O(V).
// The host picks, per superstep, whether the scatter builds a list or a
// bitmap from the last frontier size the kernels reported (sparse: list;
// dense: bitmap), the switch driven by frontier density.
//
// Work on an active vertex is split by degree:
//   degree <= t1        one thread walks the vertex's edges;
//   t1 < degree <= t2   one wave64 walks them 64 at a time (coalesced) and
//                       reduces with cross-lane shuffles;
//   degree > t2         "hubs": the vertex goes to a hub list whose edges
//                       are cut into equal merge-path tiles of MP_NV items
//                       staged through LDS, with a segmented reduce and a
//                       fold of the per-tile partials of segments that
//                       cross tiles.
// The host knows the graph's maximum in/out degree and launches the hub
// kernels only when some vertex exceeds t2.
//
//   k_gather_vp          thread/wave tiers of the gather; appends hubs
//   k_hscan1/2, k_hgather, k_hfix   hub gather (skipped without hubs)
//   k_apply              Program::apply; fused with the thread/wave tiers
//                        of the scatter inside run(), or recording the
//                        return flag for scatterActivate()
//   k_scatter_vp         scatter tiers when apply was not fused
//   k_hscan1/2, k_hscatter          hub scatter (skipped without hubs)
//   k_compact1/2         ordered compaction of activation bytes into a list
//                        (the step-by-step API, whose next frontier is in
//                        ascending vertex order as refgas.h:201-210)
//   k_fused              gather + apply + scatter of a superstep in one
//                        launch (no gather hubs, all / bitmap frontier)
//   k_small              consecutive supersteps of a small list frontier
//                        in one workgroup (graphs of degree <= 64)
//
// The first kernel of a superstep reports a frontier size to a host-mapped
// status word; the host reads it a few supersteps behind and stops once a
// reported frontier is empty (later supersteps are no-ops).  Wave size is
// 64.  No MFMA: there is no dense contraction on this path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace gas {

constexpr int NT = 256;            // threads per workgroup: 4 waves of 64
constexpr int WAVE = 64;
constexpr int SCAN_BLOCKS = 1024;  // chunks of per-vertex kernels and scans
constexpr int MP_VT = 6;           // merge-path items per thread
constexpr int MP_NV = NT * MP_VT;  // merge-path items per tile
constexpr int MP_BLOCKS = 4096;    // grid of the merge-path kernels
constexpr int HUB_MIN_T2 = 4;      // hubs have degree > t2 >= 4 ...
// ... so a segment spans >= 6 merge-path items and a tile touches at most
constexpr int HUB_SEG_MAX = MP_NV / (HUB_MIN_T2 + 2) + 2;  // segments
// span hub gather: source values of the lowest ids (the highest in-degree
// in a degree-ordered graph, whose sources are themselves skewed to low ids:
// 2,048 of R-MAT-21's 2.1 M vertices source 22% of its edges) staged in LDS
// per workgroup, so those gathers take no L2 request
constexpr int SPAN_HOT = 2048;
constexpr int SPAN_VT = 8;                // span hub gather: edges per thread
constexpr int SPAN_TE = NT * SPAN_VT;     // ... per tile
// offsets a span tile needs: the segments it overlaps (one holds its first
// edge, at most (SPAN_TE - 1) / (HUB_MIN_T2 + 1) more start in it) + 1
constexpr int SPAN_SEG = (SPAN_TE - 1) / (HUB_MIN_T2 + 1) + 3;
constexpr int T1_MAX = 16;         // largest thread tier (unrolled loads)
constexpr int MID_UNROLL = 8;      // wave tier: edges per lane per trip
constexpr int DEFAULT_T1 = 8;      // thread tier: degree <= t1
constexpr int DEFAULT_T2 = 8;      // wave tier: degree <= t2; above: hubs
constexpr int STAT_SLOTS = 6;      // run totals kept per workgroup (add_stat)
constexpr int ELL_W = 8;           // width of the out-neighbour table

// ---------------------------------------------------------------------------
// Superstep number visible to vertex programs.  The engine's apply kernel
// stores the device-resident superstep counter here before calling
// Program::apply; the host overload serves programs evaluated on the CPU.
static __shared__ int s_superstep;
inline int &host_superstep() {
  static thread_local int v = 0;
  return v;
}
__host__ inline int iteration() { return host_superstep(); }
__device__ inline int iteration() { return s_superstep; }

// ---------------------------------------------------------------------------
// Per-superstep counters, in a ring of four entries indexed by the
// superstep number s & 3.  Superstep s reads its own entry (list size, hub
// lists), appends the next active list into entry (s + 1) & 3, and its first
// kernel clears entry (s + 2) & 3 -- the one superstep s + 1 appends into,
// last used by superstep s - 2.  No kernel ever clears an entry another
// kernel of the same launch reads or writes.
struct StepCounts {
  int32_t n_active;    // |A| of a list input (appended by superstep s - 1)
  int32_t n_ghub;      // gather hubs appended by this superstep
  int32_t ghub_items;  // their summed in-degree
  int32_t n_shub;      // scatter hubs appended by this superstep
  int32_t shub_items;  // their summed out-degree
  int32_t pad[3];
};
constexpr int STEP_COUNT_WORDS = sizeof(StepCounts) / 4;

// Control block, one per engine, in device memory.
struct Ctl {
  StepCounts ring[4];
  // k_small's report: supersteps run, size of the list it left, and whether
  // it stopped on an empty frontier
  int32_t small_steps, small_n, small_empty, small_pad;
};

// How a superstep finds its active vertices.
enum : int32_t {
  IN_ALL = 0,     // every vertex, slot i is vertex i
  IN_LIST = 1,    // slots index the list q[p] of ring[s & 3].n_active ids
  IN_BITMAP = 2,  // slot i is vertex i, active iff its activation byte is set
};

// What the first kernel of superstep s reports to the host status ring:
// slot r & (STATUS_SLOTS - 1) receives ((r + 1) << 32) | |A(r)| for the
// reported superstep r.  Every superstep is reported exactly once: a list
// or all input by itself, a bitmap input by the next superstep (its size is
// the sum of per-workgroup counts, known only after its apply kernel).
constexpr int STATUS_SLOTS = 16;
enum : int32_t {
  PUB_OWN_LIST = 1,     // (s, list size)
  PUB_OWN_ALL = 2,      // (s, V)
  PUB_PREV_BITMAP = 4,  // (s - 1, sum of the per-workgroup counts of s - 1)
};

// Optional program trait: `static constexpr bool compensatedSum = true`
// selects TwoSum-compensated accumulation of a float sum.
template <class P, class = void>
struct is_compensated : std::false_type {};
template <class P>
struct is_compensated<P, std::void_t<decltype(P::compensatedSum)>>
    : std::integral_constant<bool, P::compensatedSum> {};

// Optional program trait: `static constexpr bool scatterEdges = true`
// declares a Program::scatter that writes its edge.  The engine then calls
// it for every out-edge of each vertex whose apply returned true, after
// every apply of the superstep, whenever scatterActivate(true) runs (run()
// included), as the reference CPU engine does (refgas.h:170-194).  Without
// it scatter is never called, as in the reference GPU engine
// (gpugas.h:666-668).
template <class P, class = void>
struct scatters_edges : std::false_type {};
template <class P>
struct scatters_edges<P, std::void_t<decltype(P::scatterEdges)>>
    : std::integral_constant<bool, P::scatterEdges> {};

// Optional program trait: `static constexpr bool gatherSourceOnly = true`
// declares that gatherMap reads only its source vertex (not the
// destination, not the edge).  The engine then keeps gatherMap(src) per
// vertex in an array refreshed by every apply, and gathers read that array
// (sizeof(GatherResult) per edge) instead of the source's whole state.
template <class P, class = void>
struct is_source_only : std::false_type {};
template <class P>
struct is_source_only<P, std::void_t<decltype(P::gatherSourceOnly)>>
    : std::integral_constant<bool, P::gatherSourceOnly> {};

// Optional program trait: `static constexpr bool pushMin = true` declares
// a min-program whose superstep can run as a push: VertexData and
// GatherResult are the same int, gatherMap(dst, src, edge) does not read
// dst, gatherReduce is min, and apply(v, g) lowers v to g when g < v and
// returns whether it did (SSSP, CC).  Then the superstep's new values are
// min(old, min over changed in-neighbours u of gatherMap(u, edge)), which
// the changed vertices can push along their out-edges with atomicMin
// (k_push) -- the same values, superstep by superstep, as the pull over
// the out-neighbours of the changed vertices.  An optional
// `static bool pushSkip(const VertexData &)` names values whose pushes
// cannot lower anything (SSSP's unreached distance).
template <class P, class = void>
struct is_push_min : std::false_type {};
template <class P>
struct is_push_min<P, std::void_t<decltype(P::pushMin)>>
    : std::integral_constant<bool, P::pushMin> {};
template <class P, class = void>
struct has_push_skip : std::false_type {};
template <class P>
struct has_push_skip<P, std::void_t<decltype(P::pushSkip(
                            std::declval<const typename P::VertexData &>()))>>
    : std::true_type {};

template <class T>
__device__ inline T shfl_xor_4b(T x, int m) {
  static_assert(sizeof(T) == 4, "cross-lane reduce needs a 4-byte result");
  int r = __shfl_xor(__builtin_bit_cast(int, x), m, WAVE);
  return __builtin_bit_cast(T, r);
}

// Partial gather result with an explicit "holds a value" flag, so partial
// results never need an identity element: the seed gatherZero is folded in
// exactly once when a vertex's gather is finished (refgas.h:148-158).
// Every program here has a commutative, associative gatherReduce, so the
// order partials are combined in does not change integer results.
template <class P, bool Comp = is_compensated<P>::value>
struct Acc {
  typedef typename P::GatherResult T;
  T v;
  int ok;
  __device__ static Acc empty() {
    Acc a;
    a.v = T();
    a.ok = 0;
    return a;
  }
  __device__ void add(const T &x) {
    if (ok) {
      v = P::gatherReduce(v, x);
    } else {
      v = x;
      ok = 1;
    }
  }
  // this holds the earlier items, b the later ones
  __device__ void merge(const Acc &b) {
    if (b.ok) add(b.v);
  }
  __device__ Acc xor_peer(int m) const {
    Acc r;
    r.v = shfl_xor_4b(v, m);
    r.ok = __shfl_xor(ok, m, WAVE);
    return r;
  }
  __device__ Acc up_peer(int o) const {
    Acc r;
    r.v = __builtin_bit_cast(T, __shfl_up(__builtin_bit_cast(int, v), o, WAVE));
    r.ok = __shfl_up(ok, o, WAVE);
    return r;
  }
  __device__ T finish() const {
    T z = P::gatherZero;
    return ok ? P::gatherReduce(z, v) : z;
  }
};

// Compensated float sum: (s, c) with TwoSum error terms (Knuth, TAOCP 2,
// 4.2.2; Neumaier's variant for merging two partial sums).  No "holds a
// value" flag: the empty sum is (0, 0), and TwoSum(s, 0) = (s, 0) and
// TwoSum(0, x) = (x, 0) exactly, so adding or merging an empty partial
// changes nothing, and finishing it gives gatherZero (0).
template <class P>
struct Acc<P, true> {
  typedef typename P::GatherResult T;
  static_assert(std::is_same<T, float>::value, "compensated sum needs float");
  static_assert(P::gatherZero == 0.0f, "compensated sum starts from 0");
  float s, c;
  __device__ static Acc empty() {
    Acc a;
    a.s = 0.f;
    a.c = 0.f;
    return a;
  }
  __device__ static float two_sum(float a, float b, float &err) {
    float t = a + b;
    float bp = t - a;
    err = (a - (t - bp)) + (b - bp);
    return t;
  }
  __device__ void add(float x) {
    float err;
    s = two_sum(s, x, err);
    c += err;
  }
  __device__ void merge(const Acc &b) {
    float err;
    s = two_sum(s, b.s, err);
    c = c + b.c + err;
  }
  __device__ Acc xor_peer(int m) const {
    Acc r;
    r.s = __shfl_xor(s, m, WAVE);
    r.c = __shfl_xor(c, m, WAVE);
    return r;
  }
  __device__ Acc up_peer(int o) const {
    Acc r;
    r.s = __shfl_up(s, o, WAVE);
    r.c = __shfl_up(c, o, WAVE);
    return r;
  }
  __device__ T finish() const { return s + c; }
};

template <class P>
struct SegItem {
  Acc<P> a;
  int flag;  // segment end seen
};

template <class P>
__device__ inline SegItem<P> seg_combine(const SegItem<P> &x,
                                         const SegItem<P> &y) {
  if (y.flag) return y;
  SegItem<P> r = x;
  r.a.merge(y.a);
  return r;
}

template <class P>
__device__ inline SegItem<P> seg_empty() {
  SegItem<P> r;
  r.a = Acc<P>::empty();
  r.flag = 0;
  return r;
}

// Segmented scan over the workgroup in thread order: returns the exclusive
// prefix of (segment end seen, partial) for this thread and leaves the
// workgroup's inclusive total in `total` (all threads).  Wave-level shuffle
// scan, then one LDS exchange of the NT/WAVE wave totals.
template <class P>
__device__ inline SegItem<P> block_seg_scan(SegItem<P> me, SegItem<P> &total,
                                            SegItem<P> *s_wtot) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    SegItem<P> up;
    up.a = me.a.up_peer(o);
    up.flag = __shfl_up(me.flag, o, WAVE);
    if (lane >= o) me = seg_combine(up, me);
  }
  SegItem<P> prev;
  prev.a = me.a.up_peer(1);
  prev.flag = __shfl_up(me.flag, 1, WAVE);
  if (lane == WAVE - 1) s_wtot[w] = me;
  __syncthreads();
  SegItem<P> wpre = seg_empty<P>();
  for (int k = 0; k < w; ++k) wpre = seg_combine(wpre, s_wtot[k]);
  SegItem<P> tot = seg_empty<P>();
#pragma unroll
  for (int k = 0; k < NT / WAVE; ++k) tot = seg_combine(tot, s_wtot[k]);
  total = tot;
  return lane == 0 ? wpre : seg_combine(wpre, prev);
}

// Device view of one engine.
template <class P>
struct Dev {
  typedef typename P::VertexData VD;
  typedef typename P::EdgeData ED;
  typedef typename P::GatherResult GR;
  int32_t nV, nE;
  const int32_t *csc_off, *csc_src, *csr_off, *csr_dst;
  const ED *edata;  // CSC order
  const ED *edata_csr;  // the same values in CSR order (push supersteps), or null
  // out-neighbour table (gas_graph_ell): ELL_W ids per vertex, -1 after the
  // last, and their edge values; null unless every out-degree <= ELL_W
  const int32_t *ell_dst;
  const ED *ell_w;
  ED *edata_rw;     // the same array, written by scatterEdges programs
  const int32_t *csr_csc;  // CSR slot -> CSC slot (scatterEdges programs)
  VD *vdata;        // current vertex state (read by gathers)
  VD *vnext;        // the other state buffer, written by k_fused's apply
  GR *srcval;       // gatherMap(v) per vertex (source-only programs)
  uint8_t *chg[2];  // k_fused: apply returned true, by parity (V rounded to 4)
  Ctl *ctl;
  int32_t *q[2];     // active lists by parity
  uint8_t *fl[2];    // activation bytes by parity (V rounded up to 4)
  int32_t nfwords;   // activation bytes as 32-bit words
  uint8_t *rflag;    // apply's return per active slot (unfused path)
  GR *gtmp;          // gather result per active slot
  int32_t gt1, gt2;  // gather degree tiers (thread <= gt1 < wave <= gt2 < hub)
  int32_t st1, st2;  // scatter degree tiers
  // gather hubs of the current superstep (list / bitmap input): slots
  int32_t *ghub, *ghoff, *ghpart;
  // gather hubs when every vertex is active (fixed per graph): vertex ids
  const int32_t *ahub, *ahoff, *ahpart;
  int32_t n_ahub, ahub_items;
  // the same hubs as one span of vertices [span_lo, span_hi) with in-edges
  // [span_e0, span_e1), and the vertex holding each span tile's first edge
  // (null: the hubs are not one span, or the program cannot use it)
  const int32_t *span_first;
  int32_t span_lo, span_hi, span_e0, span_e1;
  // scatter hubs of the current superstep: slots
  int32_t *shub, *shoff, *shpart;
  int32_t *bsum;     // per-chunk sums of the two-pass scans
  int32_t *bcnt;     // [2][SCAN_BLOCKS] active slots per workgroup, bitmap input
  int64_t *stat_acc; // [STAT_SLOTS][SCAN_BLOCKS] per-workgroup totals (add_stat)
  double *rpart;     // per-chunk FP64 residual sums
  int64_t *status;   // host-mapped status ring of STATUS_SLOTS words (PUB_*)
  // set by the host for every launch
  int32_t seq;       // superstep number (ring index seq & 3)
  int32_t iter;      // the program-visible superstep counter
  int32_t in_mode;   // IN_*
  int32_t out_list;  // 1: scatter claims and appends the next list; 0: bytes only
  int32_t first;     // 1: first kernel of the superstep (clears, publishes)
  int32_t pub;       // PUB_* bits
  Acc<P> *part_first, *part_last;  // per-tile partials of crossing segments
};

// gatherMap of in-edge e = (src -> dst): from the per-vertex array for
// source-only programs, from the states otherwise.
template <class P>
__device__ __forceinline__ typename P::GatherResult gather_value(
    const Dev<P> &d, int dst, int src, int e) {
  if constexpr (is_source_only<P>::value)
    return d.srcval[src];
  else
    return P::gatherMap(d.vdata + dst, d.vdata + src, d.edata + e);
}

// ---------------------------------------------------------------------------
// Wave / block helpers.
__device__ inline int lane_id() { return threadIdx.x & (WAVE - 1); }

__device__ inline int wave_incl_scan(int x) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    int y = __shfl_up(x, o, WAVE);
    if (lane >= o) x += y;
  }
  return x;
}

__device__ inline int wave_sum(int x) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
  return x;
}

__device__ inline double wave_sum(double x) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
  return x;
}

// Exclusive scan over the block; lds needs NT/WAVE ints.  Returns the
// exclusive prefix and writes the block total.
__device__ inline int block_excl_scan(int x, int &total, int *lds) {
  const int w = threadIdx.x / WAVE;
  int inc = wave_incl_scan(x);
  if (lane_id() == WAVE - 1) lds[w] = inc;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; ++i) {
    int t = lds[i];
    pre += (i < w) ? t : 0;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return pre + inc - x;
}

__device__ inline int block_sum(int x, int *lds) {
  int s = wave_sum(x);
  if (lane_id() == 0) lds[threadIdx.x / WAVE] = s;
  __syncthreads();
  int t = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; ++i) t += lds[i];
  __syncthreads();
  return t;
}

// Workgroup sums of three counters in one LDS exchange (lds: 3 * NT / WAVE).
__device__ inline void block_sum3(int &a, int &b, int &c, int *lds) {
  const int sa = wave_sum(a), sb = wave_sum(b), sc = wave_sum(c);
  const int w = threadIdx.x / WAVE;
  if (lane_id() == 0) {
    lds[w] = sa;
    lds[NT / WAVE + w] = sb;
    lds[2 * (NT / WAVE) + w] = sc;
  }
  __syncthreads();
  int ta = 0, tb = 0, tc = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; ++i) {
    ta += lds[i];
    tb += lds[NT / WAVE + i];
    tc += lds[2 * (NT / WAVE) + i];
  }
  __syncthreads();
  a = ta;
  b = tb;
  c = tc;
}

// Workgroup sums of four counters in one LDS exchange (lds: 4 * NT / WAVE).
__device__ inline void block_sum4(int &a, int &b, int &c, int &e, int *lds) {
  const int sa = wave_sum(a), sb = wave_sum(b), sc = wave_sum(c), se = wave_sum(e);
  const int w = threadIdx.x / WAVE;
  if (lane_id() == 0) {
    lds[w] = sa;
    lds[NT / WAVE + w] = sb;
    lds[2 * (NT / WAVE) + w] = sc;
    lds[3 * (NT / WAVE) + w] = se;
  }
  __syncthreads();
  int ta = 0, tb = 0, tc = 0, te = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; ++i) {
    ta += lds[i];
    tb += lds[NT / WAVE + i];
    tc += lds[2 * (NT / WAVE) + i];
    te += lds[3 * (NT / WAVE) + i];
  }
  __syncthreads();
  a = ta;
  b = tb;
  c = tc;
  e = te;
}

__device__ inline double block_sum(double x, double *lds) {
  double s = wave_sum(x);
  if (lane_id() == 0) lds[threadIdx.x / WAVE] = s;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; ++i) t += lds[i];
  __syncthreads();
  return t;
}

__device__ inline int lds_lower_bound(const int32_t *a, int n, int x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ inline int lanemask_lt_count(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

// Chunk [b, e) of n items handled by workgroup `blk`.
__device__ inline void scan_chunk(int n, int blk, int &b, int &e) {
  int c = (n + SCAN_BLOCKS - 1) / SCAN_BLOCKS;
  b = min(n, blk * c);
  e = min(n, b + c);
}

// Prefix of the chunk sums before chunk `blk`, and the grand total.
__device__ inline void chunk_prefix(const int32_t *bsum, int blk, int &pre,
                                    int &total, int *lds) {
  int p = 0, t = 0;
  for (int i = threadIdx.x; i < SCAN_BLOCKS; i += NT) {
    int x = bsum[i];
    t += x;
    p += (i < blk) ? x : 0;
  }
  pre = block_sum(p, lds);
  total = block_sum(t, lds);
}

// Merge-path layout: segment h (0 <= h < n) occupies the positions
// off[h]+h .. off[h+1]+h -- its edges, then its end item.  Tile t covers
// positions [t*MP_NV, (t+1)*MP_NV); part[t] is the segment holding the
// tile's first position and part[ntiles] = n.  The writer of segment h
// records itself for every tile that starts inside it.
__device__ inline void mark_tiles(int32_t *part, int h, int off_h, int deg) {
  const int start = off_h + h;
  const int end = off_h + deg + h;
  for (int t = (start + MP_NV - 1) / MP_NV; t * MP_NV <= end; ++t)
    part[t] = h;
}

// Appends `slot` to a global list for the lanes with `take`, one
// atomicAdd per wave-round into LDS and one per workgroup round globally.
// Must be reached by every thread of the workgroup.
__device__ inline void block_append(bool take, int slot, int32_t *list,
                                    int32_t *count, int32_t *s_buf,
                                    int *s_cnt, int *s_base) {
  if (threadIdx.x == 0) *s_cnt = 0;
  __syncthreads();
  uint64_t m = __ballot(take);
  if (m) {
    int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(s_cnt, __popcll(m));
    base = __shfl(base, leader, WAVE);
    if (take) s_buf[base + lanemask_lt_count(m)] = slot;
  }
  __syncthreads();
  const int n = *s_cnt;
  if (threadIdx.x == 0 && n) *s_base = atomicAdd(count, n);
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += NT) list[*s_base + i] = s_buf[i];
}

// Row v of the out-neighbour table: ELL_W ids (-1 after the last) in two
// 16-byte loads.
__device__ __forceinline__ void load_ell_row(const int32_t *ell, int v, int (&e)[ELL_W]) {
  const int4 *row = reinterpret_cast<const int4 *>(ell + (int64_t)v * ELL_W);
  const int4 r0 = row[0], r1 = row[1];
  e[0] = r0.x; e[1] = r0.y; e[2] = r0.z; e[3] = r0.w;
  e[4] = r1.x; e[5] = r1.y; e[6] = r1.z; e[7] = r1.w;
}

// ---------------------------------------------------------------------------
// Frontier of the current superstep.
template <class P>
__device__ inline int frontier_size(const Dev<P> &d) {
  return d.in_mode == IN_LIST ? d.ctl->ring[d.seq & 3].n_active : d.nV;
}

// Vertex of slot i (< frontier_size) and whether it is active.
template <class P>
__device__ inline bool frontier_vertex(const Dev<P> &d, int p, int i, int &v) {
  if (d.in_mode == IN_LIST) {
    v = d.q[p][i];
    return true;
  }
  v = i;
  return d.in_mode == IN_ALL || d.fl[p][i] != 0;
}

// Vertex of a slot stored in a hub list.
template <class P>
__device__ inline int slot_vertex(const Dev<P> &d, int p, int slot) {
  return d.in_mode == IN_LIST ? d.q[p][slot] : slot;
}

// Duties of the first kernel of a superstep, done by its workgroup 0: clear
// ring entry (s + 2) & 3 and report frontier sizes to the host status ring
// (system-scope stores into host-mapped memory; the host reads them while
// later supersteps run).  Must be reached by every thread of the workgroup.
template <class P>
__device__ inline void superstep_prologue(const Dev<P> &d, int p, int *lds) {
  if (!d.first || blockIdx.x != 0) return;
  Ctl *c = d.ctl;
  if (threadIdx.x < STEP_COUNT_WORDS)
    reinterpret_cast<int32_t *>(&c->ring[(d.seq + 2) & 3])[threadIdx.x] = 0;
  if (d.pub & PUB_PREV_BITMAP) {
    int t = 0;
    for (int i = threadIdx.x; i < SCAN_BLOCKS; i += NT)
      t += d.bcnt[(p ^ 1) * SCAN_BLOCKS + i];
    t = block_sum(t, lds);
    const int64_t r = d.seq - 1;
    if (threadIdx.x == 0)
      __hip_atomic_store(d.status + (r & (STATUS_SLOTS - 1)),
                         ((r + 1) << 32) | (int64_t)(uint32_t)t,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  if ((d.pub & (PUB_OWN_LIST | PUB_OWN_ALL)) && threadIdx.x == 0) {
    const int64_t r = d.seq;
    const int a = (d.pub & PUB_OWN_LIST) ? c->ring[d.seq & 3].n_active : d.nV;
    __hip_atomic_store(d.status + (r & (STATUS_SLOTS - 1)),
                       ((r + 1) << 32) | (int64_t)(uint32_t)a,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Claims vertex u for the next list: sets its activation byte and returns
// true for exactly one of the threads that activate u in this superstep.
__device__ inline bool claim_vertex(uint8_t *fl, int u) {
  if (fl[u]) return false;  // already set (bytes only go 0 -> 1 here)
  const uint32_t bit = 1u << ((u & 3) * 8);
  const uint32_t old = atomicOr(reinterpret_cast<uint32_t *>(fl) + (u >> 2), bit);
  return (old & bit) == 0;
}

// Per-workgroup run totals (0 gathered, 1 scattered, 2 active, 3 vertices
// whose apply returned true, 4 gather hubs, 5 their in-edges): each
// workgroup adds into its own slot (no contention); a non-returning atomic
// add, so the workgroup does not wait for a read before it ends.
template <class P>
__device__ inline void add_stat(const Dev<P> &d, int which, int64_t x) {
  if (threadIdx.x == 0 && x)
    atomicAdd(reinterpret_cast<unsigned long long *>(
                  d.stat_acc + which * SCAN_BLOCKS + blockIdx.x),
              (unsigned long long)x);
}

// Calls body(valid, slot, vertex) for every active slot of the superstep,
// in rounds of NT that every thread of the workgroup reaches (body may use
// workgroup barriers).  All / list inputs: slot i = round * stride + thread.
// Bitmap input: the workgroup reads its chunk of activation words
// (BM_WORDS at a time, 4 slots per word), compacts the active vertex ids
// into LDS in ascending order and processes them NT per round, so a sparse
// bitmap costs one read of its words plus one round; with clear = true the
// words read are zeroed (the last reader of this superstep's bitmap).
constexpr int BM_WORDS = 2 * NT;
struct NoWordHook {
  __device__ void operator()(int, uint32_t) const {}
};
template <class P, class Body, class WordHook = NoWordHook>
__device__ __forceinline__ void for_each_active(const Dev<P> &d, int p, bool clear,
                                                int32_t *s_ids, int *lds, Body body,
                                                WordHook hook = WordHook()) {
  if (d.in_mode != IN_BITMAP) {
    const int A = frontier_size(d);
    const int stride = gridDim.x * NT;
    const int rounds = (A + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {
      const int i = r * stride + blockIdx.x * NT + threadIdx.x;
      int v = 0;
      const bool valid = i < A && frontier_vertex(d, p, i, v);
      body(valid, i, v);
    }
    return;
  }
  uint32_t *fw = reinterpret_cast<uint32_t *>(d.fl[p]);
  int wb, we;
  scan_chunk(d.nfwords, blockIdx.x, wb, we);
  for (int w0 = wb; w0 < we; w0 += BM_WORDS) {
    uint32_t wd[BM_WORDS / NT];
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < BM_WORDS / NT; ++k) {
      const int wi = w0 + k * NT + threadIdx.x;
      wd[k] = wi < we ? fw[wi] : 0u;
      cnt += __popc(wd[k]);
    }
#pragma unroll
    for (int k = 0; k < BM_WORDS / NT; ++k) {
      const int wi = w0 + k * NT + threadIdx.x;
      if (wi < we) hook(wi, wd[k]);
    }
    int tot;
    int pos = block_excl_scan(cnt, tot, lds);
#pragma unroll
    for (int k = 0; k < BM_WORDS / NT; ++k) {
      if (wd[k]) {
        const int wi = w0 + k * NT + threadIdx.x;
        if (clear) fw[wi] = 0u;
#pragma unroll
        for (int b = 0; b < 4; ++b)
          if ((wd[k] >> (8 * b)) & 0xffu) s_ids[pos++] = wi * 4 + b;
      }
    }
    __syncthreads();
    for (int j0 = 0; j0 < tot; j0 += NT) {
      const int j = j0 + threadIdx.x;
      const bool valid = j < tot;
      const int v = valid ? s_ids[j] : 0;
      body(valid, v, v);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Frontier initialisation: setActive(start, end) (gpugas.h:342-349); clears
// the activation and changed bytes of both parities (unless the host knows
// no scatter ran since they were last cleared) and the counter ring, and
// lists [start, end) unless it is every vertex.
template <class P>
__global__ __launch_bounds__(NT) void k_set_active(Dev<P> d, int p, int start,
                                                   int end, int clear) {
  const bool all = (start == 0 && end == d.nV);
  const int n = end - start;
  int32_t *q = d.q[p];
  const int stride = gridDim.x * NT;
  uint32_t *f0 = reinterpret_cast<uint32_t *>(d.fl[0]);
  uint32_t *f1 = reinterpret_cast<uint32_t *>(d.fl[1]);
  uint32_t *c0 = reinterpret_cast<uint32_t *>(d.chg[0]);
  uint32_t *c1 = reinterpret_cast<uint32_t *>(d.chg[1]);
  for (int i = blockIdx.x * NT + threadIdx.x; clear && i < d.nfwords;
       i += stride) {
    f0[i] = 0u;
    f1[i] = 0u;
    c0[i] = 0u;
    c1[i] = 0u;
  }
  if (!all)
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += stride)
      q[i] = start + i;
  if (blockIdx.x == 0) {
    int32_t *ring = reinterpret_cast<int32_t *>(d.ctl->ring);
    if (threadIdx.x < 4 * STEP_COUNT_WORDS) ring[threadIdx.x] = 0;
    __syncthreads();
    if (threadIdx.x == 0) d.ctl->ring[d.seq & 3].n_active = n;
  }
}

// Thread and wave tiers of the gather; hub slots are appended to ghub
// (list / bitmap input) or skipped (all active: the fixed hub list covers
// them).
template <class P>
__global__ __launch_bounds__(NT) void k_gather_vp(Dev<P> d, int p) {
  typedef typename P::GatherResult GR;
  __shared__ int32_t s_buf[NT];
  __shared__ int32_t s_ids[BM_WORDS * 4];
  __shared__ int s_cnt, s_base;
  __shared__ int lds[NT / WAVE];
  superstep_prologue(d, p, lds);
  Ctl *c = d.ctl;
  const bool all = d.in_mode == IN_ALL;
  const int lane = lane_id();
  int gathered = 0, hubv = 0, hube = 0;
  for_each_active(d, p, false, s_ids, lds, [&](bool valid, int i, int v) __attribute__((always_inline)) {
    int o = 0, deg = 0;
    if (valid) {
      o = d.csc_off[v];
      deg = d.csc_off[v + 1] - o;
      gathered += deg;
    }
    const bool small = valid && deg <= d.gt1;
    const bool mid = valid && deg > d.gt1 && deg <= d.gt2;
    const bool hub = valid && deg > d.gt2;
    hubv += hub ? 1 : 0;
    hube += hub ? deg : 0;
    if (small) {
      // every load of the vertex issued before the first use
      int src[T1_MAX];
#pragma unroll
      for (int k = 0; k < T1_MAX; ++k)
        src[k] = (k < deg) ? d.csc_src[o + k] : 0;
      GR val[T1_MAX];
#pragma unroll
      for (int k = 0; k < T1_MAX; ++k)
        if (k < deg)
          val[k] = gather_value(d, v, src[k], o + k);
      Acc<P> acc = Acc<P>::empty();
#pragma unroll
      for (int k = 0; k < T1_MAX; ++k)
        if (k < deg) acc.add(val[k]);
      d.gtmp[i] = acc.finish();
    }
    uint64_t m = __ballot(mid);
    while (m) {
      const int j = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const int vj = __shfl(v, j, WAVE);
      const int oj = __shfl(o, j, WAVE);
      const int dj = __shfl(deg, j, WAVE);
      const int ij = __shfl(i, j, WAVE);
      Acc<P> acc = Acc<P>::empty();
      for (int base = 0; base < dj; base += WAVE * MID_UNROLL) {
        int src[MID_UNROLL];
#pragma unroll
        for (int u = 0; u < MID_UNROLL; ++u) {
          const int k = base + u * WAVE + lane;
          src[u] = (k < dj) ? d.csc_src[oj + k] : 0;
        }
        GR val[MID_UNROLL];
#pragma unroll
        for (int u = 0; u < MID_UNROLL; ++u) {
          const int k = base + u * WAVE + lane;
          if (k < dj)
            val[u] = gather_value(d, vj, src[u], oj + k);
        }
#pragma unroll
        for (int u = 0; u < MID_UNROLL; ++u)
          if (base + u * WAVE + lane < dj) acc.add(val[u]);
      }
#pragma unroll
      for (int x = WAVE / 2; x > 0; x >>= 1) acc.merge(acc.xor_peer(x));
      if (lane == 0) d.gtmp[ij] = acc.finish();
    }
    if (!all)
      block_append(hub, i, d.ghub, &c->ring[d.seq & 3].n_ghub, s_buf, &s_cnt,
                   &s_base);
  });
  {
    __shared__ int lds3[3 * (NT / WAVE)];
    block_sum3(gathered, hubv, hube, lds3);
  }
  add_stat(d, 0, gathered);
  add_stat(d, 4, hubv);
  add_stat(d, 5, hube);
}

// Hub list of a superstep: all-active supersteps use the fixed list.
template <class P>
struct HubList {
  const int32_t *list, *off, *part;
  int n, items;
};

template <class P>
__device__ inline HubList<P> gather_hubs(const Dev<P> &d) {
  HubList<P> h;
  if (d.in_mode == IN_ALL) {
    h.list = d.ahub;
    h.off = d.ahoff;
    h.part = d.ahpart;
    h.n = d.n_ahub;
    h.items = d.ahub_items;
  } else {
    const StepCounts &sc = d.ctl->ring[d.seq & 3];
    h.list = d.ghub;
    h.off = d.ghoff;
    h.part = d.ghpart;
    h.n = sc.n_ghub;
    h.items = sc.ghub_items;
  }
  return h;
}

// Degree scan over a hub list (pass 1: chunk sums).  Which = 0: gather hubs
// (in-degrees), 1: scatter hubs (out-degrees).
template <class P, int Which>
__global__ __launch_bounds__(NT) void k_hscan1(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  if (Which == 0 && d.in_mode == IN_ALL) return;  // fixed list, scanned on the host
  const StepCounts &sc = d.ctl->ring[d.seq & 3];
  const int n = Which == 0 ? sc.n_ghub : sc.n_shub;
  const int32_t *list = Which == 0 ? d.ghub : d.shub;
  const int32_t *gofs = Which == 0 ? d.csc_off : d.csr_off;
  int b, e;
  scan_chunk(n, blockIdx.x, b, e);
  int sum = 0;
  for (int i = b + threadIdx.x; i < e; i += NT) {
    const int v = slot_vertex(d, p, list[i]);
    sum += gofs[v + 1] - gofs[v];
  }
  sum = block_sum(sum, lds);
  if (threadIdx.x == 0) d.bsum[blockIdx.x] = sum;
}

// Pass 2: offsets, tile partition and total of the hub list.
template <class P, int Which>
__global__ __launch_bounds__(NT) void k_hscan2(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  if (Which == 0 && d.in_mode == IN_ALL) return;
  StepCounts &sc = d.ctl->ring[d.seq & 3];
  const int n = Which == 0 ? sc.n_ghub : sc.n_shub;
  if (n == 0) return;
  const int32_t *list = Which == 0 ? d.ghub : d.shub;
  int32_t *off = Which == 0 ? d.ghoff : d.shoff;
  int32_t *part = Which == 0 ? d.ghpart : d.shpart;
  const int32_t *gofs = Which == 0 ? d.csc_off : d.csr_off;
  int b, e, pre, total;
  scan_chunk(n, blockIdx.x, b, e);
  chunk_prefix(d.bsum, blockIdx.x, pre, total, lds);
  int carry = pre;
  for (int base = b; base < e; base += NT) {
    const int i = base + threadIdx.x;
    int deg = 0;
    if (i < e) {
      const int v = slot_vertex(d, p, list[i]);
      deg = gofs[v + 1] - gofs[v];
    }
    int tot;
    const int ex = block_excl_scan(deg, tot, lds);
    if (i < e) {
      off[i] = carry + ex;
      mark_tiles(part, i, carry + ex, deg);
    }
    carry += tot;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    off[n] = total;
    part[(total + n + MP_NV - 1) / MP_NV] = n;
    if (Which == 0)
      sc.ghub_items = total;
    else
      sc.shub_items = total;
  }
}

// Hub gather: merge-path tiles over the hubs' CSC in-edges with the
// segmented reduce fused in (replaces kGatherMap + ReduceByKey,
// gpugas_kernels.cuh:53-134, gpugas.h:466-500).  A hub finished inside a
// tile is written to gtmp straight away; for a hub crossing tiles each tile
// leaves the partial before its first end (part_first) and after its last
// end (part_last), and k_hfix folds them.
template <class P>
__device__ __forceinline__ void hgather_tiles(const Dev<P> &d, int p) {
  typedef typename P::GatherResult GR;
  __shared__ int32_t s_end[HUB_SEG_MAX];
  __shared__ int32_t s_base[HUB_SEG_MAX];
  __shared__ int32_t s_vtx[HUB_SEG_MAX];
  __shared__ int32_t s_slot[HUB_SEG_MAX];
  __shared__ SegItem<P> s_wtot[NT / WAVE];
  constexpr int PADDED = MP_NV + MP_NV / 32;
  __shared__ int32_t s_e[PADDED];
  __shared__ int32_t s_k[is_source_only<P>::value ? 1 : PADDED];
  __shared__ GR s_val[PADDED];
  // source-only programs: the hottest source values staged as in the span
  // form (R-MAT-21 symmetrised CC: 104 -> 90 us per launch)
  constexpr int HOT = is_source_only<P>::value ? SPAN_HOT : 1;
  __shared__ GR s_hot[HOT];

  if (frontier_size(d) == 0) return;
  const bool all = d.in_mode == IN_ALL;
  const HubList<P> h = gather_hubs(d);
  if (h.n == 0) return;
  const int N = h.items + h.n;
  const int ntiles = (N + MP_NV - 1) / MP_NV;
  const int tid = threadIdx.x;
  if ((int)blockIdx.x >= ntiles) return;
  int nhot = 0;
  if constexpr (is_source_only<P>::value) {
    nhot = min(HOT, d.nV);
    for (int k = tid; k < nhot; k += NT) s_hot[k] = d.srcval[k];
    __syncthreads();
  }

  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int d0 = t * MP_NV;
    const int d1 = min(d0 + MP_NV, N);
    const int s0 = h.part[t];
    const int s1 = h.part[t + 1];
    const int nseg = min(s1, h.n - 1) - s0 + 1;
    for (int k = tid; k < nseg; k += NT) {
      const int s = s0 + k;
      const int slot = h.list[s];
      const int v = all ? slot : slot_vertex(d, p, slot);
      s_vtx[k] = v;
      s_slot[k] = slot;
      s_base[k] = d.csc_off[v] - h.off[s];
      s_end[k] = h.off[s + 1] + s;
    }
    __syncthreads();
    const bool first_fresh = (h.off[s0] + s0 == d0);

    // decode this thread's consecutive items; edge index (or -1 for a
    // segment end) into LDS, padded one word per 32 against bank conflicts
    const int p0 = d0 + tid * MP_VT;
    int k = (p0 < d1) ? lds_lower_bound(s_end, nseg, p0) : 0;
    int kind[MP_VT], kk[MP_VT];
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      const int pos = p0 + it;
      kind[it] = 0;
      kk[it] = 0;
      int e = -1;
      if (pos < d1) {
        kk[it] = k;
        if (s_end[k] == pos) {
          kind[it] = 2;
          ++k;
        } else {
          kind[it] = 1;
          e = s_base[k] + (pos - s0 - k);
        }
      }
      const int j = tid * MP_VT + it;
      s_e[j + (j >> 5)] = e;
      if constexpr (!is_source_only<P>::value) s_k[j + (j >> 5)] = kk[it];
    }
    __syncthreads();
    // striped: lanes of a wave take consecutive items, so the source ids
    // load coalesced and neighbouring edges' gathers share cache lines
    int ee[MP_VT], src[MP_VT];
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      const int j = it * NT + tid;
      ee[it] = s_e[j + (j >> 5)];
      src[it] = ee[it] >= 0 ? d.csc_src[ee[it]] : 0;
    }
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      const int j = it * NT + tid;
      if (ee[it] >= 0) {
        int dst = 0;
        if constexpr (!is_source_only<P>::value) dst = s_vtx[s_k[j + (j >> 5)]];
        s_val[j + (j >> 5)] = src[it] < nhot ? s_hot[src[it]]
                                             : gather_value(d, dst, src[it], ee[it]);
      }
    }
    __syncthreads();
    GR val[MP_VT];
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      const int j = tid * MP_VT + it;
      val[it] = kind[it] == 1 ? s_val[j + (j >> 5)] : GR();
    }
    // thread-serial segmented reduce
    Acc<P> acc = Acc<P>::empty(), head = Acc<P>::empty();
    int nend = 0, first_k = -1;
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      // a compensated sum adds the 0 of a non-edge item exactly: no branch
      if (is_compensated<P>::value || kind[it] == 1) acc.add(val[it]);
      if (kind[it] == 2) {
        if (nend == 0) {
          head = acc;
          first_k = kk[it];
        } else {
          d.gtmp[s_slot[kk[it]]] = acc.finish();
        }
        ++nend;
        acc = Acc<P>::empty();
      }
    }
    // workgroup-wide segmented scan of (end seen, partial)
    SegItem<P> me;
    me.a = acc;
    me.flag = nend > 0;
    SegItem<P> total;
    const SegItem<P> carry = block_seg_scan(me, total, s_wtot);
    if (nend > 0) {
      Acc<P> first = carry.a;
      first.merge(head);
      if (first_k > 0 || first_fresh) d.gtmp[s_slot[first_k]] = first.finish();
      if (!carry.flag) d.part_first[t] = first;
    }
    if (tid == NT - 1) {
      if (total.flag)
        d.part_last[t] = total.a;
      else
        d.part_first[t] = total.a;
    }
    __syncthreads();
  }
}

// Folds the tile partials of hubs whose edges cross tiles.
constexpr int HFIX_LOADS = 16;
template <class P>
__device__ __forceinline__ void hfix_tiles(const Dev<P> &d, int p) {
  if (frontier_size(d) == 0) return;
  const HubList<P> h = gather_hubs(d);
  if (h.n == 0) return;
  const int ntiles = (h.items + h.n + MP_NV - 1) / MP_NV;
  // one thread per tile boundary t: the segment holding position t*MP_NV
  // crosses it when it starts before; it is folded at its first crossing
  // (it starts in tile t-1), so each crossing segment exactly once
  for (int t = 1 + blockIdx.x * NT + threadIdx.x; t < ntiles;
       t += gridDim.x * NT) {
    const int s = h.part[t];
    const int st = h.off[s] + s;
    const int t0 = t - 1;
    if (st >= t * MP_NV || st < t0 * MP_NV) continue;
    const int en = h.off[s + 1] + s;
    const int t1 = en / MP_NV;
    Acc<P> a = (st > t0 * MP_NV) ? d.part_last[t0] : d.part_first[t0];
    // a hub of many tiles: loads issued HFIX_LOADS at a time, merged in
    // tile order
    for (int u = t0 + 1; u <= t1; u += HFIX_LOADS) {
      Acc<P> b[HFIX_LOADS];
#pragma unroll
      for (int j = 0; j < HFIX_LOADS; ++j)
        b[j] = u + j <= t1 ? d.part_first[u + j] : Acc<P>::empty();
#pragma unroll
      for (int j = 0; j < HFIX_LOADS; ++j)
        if (u + j <= t1) a.merge(b[j]);
    }
    d.gtmp[h.list[s]] = a.finish();
  }
}

// Hub gather over one contiguous span of hubs (all-active supersteps of a
// source-only, compensated-sum program whose hubs are the vertices
// [span_lo, span_hi), as in a graph numbered by descending in-degree):
// their in-edges are the CSC range [span_e0, span_e1), cut into tiles of
// SPAN_TE edges.  A tile's source ids load at once (striped: lanes take
// consecutive edges) beside the offsets of the segments it overlaps --
// span_first[t] names the vertex holding its first edge -- so its only
// dependent load is the gather itself.  Values go through LDS to thread
// order; each thread sums its SPAN_VT consecutive edges and closes a
// segment after its last edge; a segmented workgroup scan carries partial
// sums between threads; segments crossing tiles leave per-tile partials
// (part_first / part_last) that hfix_span folds in tile order.
template <class P>
__device__ __forceinline__ void hgather_span(const Dev<P> &d) {
  typedef typename P::GatherResult GR;
  static_assert(is_source_only<P>::value && is_compensated<P>::value,
                "span gather: source-only programs with a compensated sum");
  constexpr int PADDED = SPAN_TE + SPAN_TE / 32;
  __shared__ GR s_val[PADDED];
  __shared__ int32_t s_off[SPAN_SEG];
  __shared__ SegItem<P> s_wtot[NT / WAVE];
  __shared__ GR s_hot[SPAN_HOT];
  const int hi = d.span_hi, e0 = d.span_e0, e1 = d.span_e1;
  const int ntiles = (e1 - e0 + SPAN_TE - 1) / SPAN_TE;
  const int tid = threadIdx.x;
  if ((int)blockIdx.x >= ntiles) return;
  const int nhot = min(SPAN_HOT, d.nV);
  for (int k = tid; k < nhot; k += NT) s_hot[k] = d.srcval[k];
  __syncthreads();
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int E0 = e0 + t * SPAN_TE;
    const int vb = d.span_first[t];
    const int nsl = min(SPAN_SEG, hi - vb + 1);  // offsets of vb .. vb + nsl - 1
    int src[SPAN_VT];
#pragma unroll
    for (int it = 0; it < SPAN_VT; ++it) {
      const int e = E0 + it * NT + tid;
      src[it] = e < e1 ? d.csc_src[e] : 0;
    }
    for (int k = tid; k < nsl; k += NT) s_off[k] = d.csc_off[vb + k];
#pragma unroll
    for (int it = 0; it < SPAN_VT; ++it) {
      const int e = E0 + it * NT + tid;
      const int j = it * NT + tid;
      s_val[j + (j >> 5)] = e >= e1           ? GR()
                            : src[it] < nhot ? s_hot[src[it]]
                                             : d.srcval[src[it]];
    }
    __syncthreads();
    // this thread's edges q0 .. q0 + SPAN_VT - 1; k: the segment holding q0
    // (past the span's end: no edges, nothing closes)
    const int q0 = E0 + tid * SPAN_VT;
    int lo_k = 0, hi_k = nsl;  // last k with s_off[k] <= q0
    while (hi_k - lo_k > 1) {
      const int mid = (lo_k + hi_k) >> 1;
      if (s_off[mid] <= q0) lo_k = mid; else hi_k = mid;
    }
    int k = lo_k;
    int next = k + 1 < nsl ? s_off[k + 1] : INT_MAX;
    Acc<P> acc = Acc<P>::empty(), head = Acc<P>::empty();
    int nend = 0, first_k = 0;
#pragma unroll
    for (int i = 0; i < SPAN_VT; ++i) {
      const int j = tid * SPAN_VT + i;
      acc.add(s_val[j + (j >> 5)]);  // 0 past the end: exact
      if (q0 + i + 1 == next) {      // the last edge of segment vb + k
        if (nend == 0) {
          head = acc;
          first_k = k;
        } else {
          d.gtmp[vb + k] = acc.finish();
        }
        ++nend;
        acc = Acc<P>::empty();
        ++k;
        next = k + 1 < nsl ? s_off[k + 1] : INT_MAX;
      }
    }
    SegItem<P> me;
    me.a = acc;
    me.flag = nend > 0;
    SegItem<P> total;
    const SegItem<P> carry = block_seg_scan(me, total, s_wtot);
    if (nend > 0) {
      Acc<P> first = carry.a;
      first.merge(head);
      if (s_off[first_k] >= E0)
        d.gtmp[vb + first_k] = first.finish();  // started in this tile
      else
        d.part_first[t] = first;  // the tile's first segment, from earlier tiles
    }
    if (tid == NT - 1) {
      if (total.flag)
        d.part_last[t] = total.a;
      else
        d.part_first[t] = total.a;
    }
    __syncthreads();
  }
}

// Folds the partials of span segments that cross tiles: one thread per
// tile boundary t; the segment holding the tile's first edge crosses it
// when it starts before, and is folded at its first crossing (it starts in
// tile t - 1): its head in tile t - 1, then the first partials of the tiles
// up to the one holding its last edge, in tile order.
template <class P>
__device__ __forceinline__ void hfix_span(const Dev<P> &d) {
  const int e0 = d.span_e0, e1 = d.span_e1;
  const int ntiles = (e1 - e0 + SPAN_TE - 1) / SPAN_TE;
  for (int t = 1 + blockIdx.x * NT + threadIdx.x; t < ntiles;
       t += gridDim.x * NT) {
    const int s = d.span_first[t];
    const int st = d.csc_off[s];
    const int E0 = e0 + t * SPAN_TE;
    if (st >= E0 || st < E0 - SPAN_TE) continue;
    const int t1 = (d.csc_off[s + 1] - 1 - e0) / SPAN_TE;
    Acc<P> a = st > E0 - SPAN_TE ? d.part_last[t - 1] : d.part_first[t - 1];
    for (int u = t; u <= t1; u += HFIX_LOADS) {
      Acc<P> b[HFIX_LOADS];
#pragma unroll
      for (int j = 0; j < HFIX_LOADS; ++j)
        b[j] = u + j <= t1 ? d.part_first[u + j] : Acc<P>::empty();
#pragma unroll
      for (int j = 0; j < HFIX_LOADS; ++j)
        if (u + j <= t1) a.merge(b[j]);
    }
    d.gtmp[s] = a.finish();
  }
}

// Hub gather of a superstep: merge-path tiles over the hub list, or (Span)
// the contiguous-span form above.
template <class P, bool Span = false>
__global__ __launch_bounds__(NT) void k_hgather(Dev<P> d, int p) {
  if constexpr (Span)
    hgather_span(d);
  else
    hgather_tiles(d, p);
}

// Tile-crossing partials of the hub gather (Span: of the span form).
template <class P, bool Span = false>
__global__ __launch_bounds__(NT) void k_hfix(Dev<P> d, int p) {
  if constexpr (Span)
    hfix_span(d);
  else
    hfix_tiles(d, p);
}

// Optional program hook: `static double residual(const VertexData &before,
// const VertexData &after)`, summed in FP64 over the active set.
template <class P, class = void>
struct has_residual : std::false_type {};
template <class P>
struct has_residual<P, std::void_t<decltype(P::residual(
                           std::declval<const typename P::VertexData &>(),
                           std::declval<const typename P::VertexData &>()))>>
    : std::true_type {};

// Thread and wave tiers of the scatter for the slots whose apply returned
// true (`ret`); so / sdeg are the vertex's CSR offset and out-degree, loaded
// by the caller next to its other first loads.  Hubs go to the scatter hub
// list.  With d.out_list the
// destinations are claimed (claim_vertex) and the claimed ones appended to
// the next list q[p^1] -- one atomicAdd per workgroup for the thread tier
// and per wave for the wave tier; otherwise the activation byte is a plain
// store (every writer stores the same value).  Reached by every thread.
template <class P>
__device__ __forceinline__ int scatter_tiers(const Dev<P> &d, int p, bool ret, int slot,
                                             int so, int sdeg, int32_t *s_buf,
                                             int *s_cnt, int *s_base, int *lds) {
  uint8_t *fl = d.fl[p ^ 1];
  int32_t *qn = d.q[p ^ 1];
  int32_t *nn = &d.ctl->ring[(d.seq + 1) & 3].n_active;
  const int o = ret ? so : 0;
  const int deg = ret ? sdeg : 0;
  const bool small = ret && deg <= d.st1;
  if (d.out_list) {
    int dst[T1_MAX];
#pragma unroll
    for (int k = 0; k < T1_MAX; ++k)
      dst[k] = (small && k < deg) ? d.csr_dst[o + k] : -1;
    int nc = 0;
#pragma unroll
    for (int k = 0; k < T1_MAX; ++k) {
      if (dst[k] >= 0 && claim_vertex(fl, dst[k]))
        ++nc;
      else
        dst[k] = -1;
    }
    int tot;
    int pos = block_excl_scan(nc, tot, lds);
    if (threadIdx.x == 0 && tot) *s_base = atomicAdd(nn, tot);
    __syncthreads();
    if (nc) {
      pos += *s_base;
#pragma unroll
      for (int k = 0; k < T1_MAX; ++k)
        if (dst[k] >= 0) qn[pos++] = dst[k];
    }
    __syncthreads();
  } else if (small) {
#pragma unroll 4
    for (int k = 0; k < deg; ++k) fl[d.csr_dst[o + k]] = 1;
  }
  const bool mid = ret && deg > d.st1 && deg <= d.st2;
  uint64_t m = __ballot(mid);
  const int lane = lane_id();
  while (m) {
    const int j = __ffsll((unsigned long long)m) - 1;
    m &= m - 1;
    const int oj = __shfl(o, j, WAVE);
    const int dj = __shfl(deg, j, WAVE);
    for (int base = 0; base < dj; base += WAVE) {
      const int k = base + lane;
      const int u = k < dj ? d.csr_dst[oj + k] : 0;
      if (d.out_list) {
        const bool got = k < dj && claim_vertex(fl, u);
        const uint64_t gm = __ballot(got);
        if (gm) {
          int at = 0;
          if (lane == 0) at = atomicAdd(nn, __popcll(gm));
          at = __shfl(at, 0, WAVE);
          if (got) qn[at + lanemask_lt_count(gm)] = u;
        }
      } else if (k < dj) {
        fl[u] = 1;
      }
    }
  }
  block_append(ret && deg > d.st2, slot, d.shub,
               &d.ctl->ring[d.seq & 3].n_shub, s_buf, s_cnt, s_base);
  return ret ? deg : 0;
}

// Apply (kApply, gpugas_kernels.cuh:139-151).  Fuse = true also runs the
// scatter tiers (the whole superstep is being executed, so scatterActivate
// follows); otherwise the return flag is kept for k_scatter_vp.  Clears the
// activation bytes of its own input (so they are zero when the superstep
// after next writes them) and, for a bitmap input, counts its active slots
// per workgroup for the next superstep's report.
template <class P, bool HaveGather, bool Fuse>
__global__ __launch_bounds__(NT) void k_apply(Dev<P> d, int p) {
  typedef typename P::GatherResult GR;
  typedef typename P::VertexData VD;
  __shared__ int32_t s_buf[NT];
  __shared__ int32_t s_ids[BM_WORDS * 4];
  __shared__ int s_cnt, s_base;
  __shared__ int lds[NT / WAVE];
  __shared__ double ldsd[NT / WAVE];
  superstep_prologue(d, p, lds);
  if (threadIdx.x == 0) s_superstep = d.iter;
  __syncthreads();
  int scattered = 0, active = 0, returned = 0;
  double res = 0.0;
  // fused activation from the out-neighbour table: the ids load with the
  // state, no CSR offsets first (bitmap output only)
  const bool ell = Fuse && d.ell_dst && !d.out_list;
  for_each_active(d, p, Fuse, s_ids, lds, [&](bool valid, int i, int v) __attribute__((always_inline)) {
    bool ret = false;
    int so = 0, sdeg = 0;
    int e[ELL_W];
    if (valid && Fuse) {
      if (ell) {
        load_ell_row(d.ell_dst, v, e);
      } else {
        so = d.csr_off[v];
        sdeg = d.csr_off[v + 1] - so;
      }
    }
    if (valid) {
      ++active;
      if (d.in_mode == IN_LIST) d.fl[p][v] = 0;
      GR g;
      if (HaveGather)
        g = d.gtmp[i];
      else
        g = P::gatherZero;
      if constexpr (has_residual<P>::value) {
        VD before = d.vdata[v];
        ret = P::apply(d.vdata + v, g);
        res += P::residual(before, d.vdata[v]);
      } else {
        ret = P::apply(d.vdata + v, g);
      }
      if constexpr (is_source_only<P>::value)
        d.srcval[v] = P::gatherMap(d.vdata + v, d.vdata + v, nullptr);
      returned += ret ? 1 : 0;
    }
    if (Fuse && ell) {
      if (ret) {
        uint8_t *fl = d.fl[p ^ 1];
#pragma unroll
        for (int k = 0; k < ELL_W; ++k)
          if (e[k] >= 0) {
            fl[e[k]] = 1;
            ++scattered;
          }
      }
    } else if (Fuse) {
      scattered += scatter_tiers(d, p, ret, i, so, sdeg, s_buf, &s_cnt, &s_base, lds);
    } else if (valid) {
      d.rflag[i] = ret ? 1 : 0;
    }
  });
  {
    __shared__ int lds3[3 * (NT / WAVE)];
    block_sum3(active, scattered, returned, lds3);
  }
  if (d.in_mode == IN_BITMAP && threadIdx.x == 0)
    d.bcnt[p * SCAN_BLOCKS + blockIdx.x] = active;
  add_stat(d, 2, active);
  add_stat(d, 1, scattered);
  add_stat(d, 3, returned);
  if constexpr (has_residual<P>::value) {
    res = block_sum(res, ldsd);
    if (threadIdx.x == 0) d.rpart[blockIdx.x] = res;
  }
}

// A whole superstep in one launch for graphs without gather hubs, on an
// all or bitmap frontier: gather (thread / wave tiers) from the current
// state buffer, apply into the other buffer, scatter tiers.  Reading one
// buffer and writing the other keeps the reference's bulk-synchronous
// semantics (every gather of a superstep sees the state before any apply,
// refgas.h:143-168) without a kernel boundary between gather and apply.
// The other buffer is complete except for vertices changed in the previous
// superstep that are not active now: the workgroup owning their activation
// words copies them across (chg[p^1]) while staging its frontier.
template <class P>
__global__ __launch_bounds__(NT) void k_fused(Dev<P> d, int p) {
  typedef typename P::GatherResult GR;
  typedef typename P::VertexData VD;
  __shared__ int32_t s_buf[NT];
  __shared__ int32_t s_ids[BM_WORDS * 4];
  __shared__ int s_cnt, s_base;
  __shared__ int lds[NT / WAVE];
  __shared__ double ldsd[NT / WAVE];
  superstep_prologue(d, p, lds);
  if (threadIdx.x == 0) s_superstep = d.iter;
  __syncthreads();
  uint8_t *chg = d.chg[p];
  const int lane = lane_id();
  int gathered = 0, scattered = 0, active = 0, returned = 0;
  double res = 0.0;
  // vertices changed last superstep and inactive now: copy their state into
  // the buffer this superstep writes (bitmap input; after an all input or a
  // frontier initialisation nothing changed)
  uint32_t *cw = reinterpret_cast<uint32_t *>(d.chg[p ^ 1]);
  auto fixup = [&](int wi, uint32_t a) __attribute__((always_inline)) {
    const uint32_t c = cw[wi];
    if (!c) return;
    cw[wi] = 0u;
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (((c >> (8 * b)) & 0xffu) && !((a >> (8 * b)) & 0xffu))
        d.vnext[wi * 4 + b] = d.vdata[wi * 4 + b];
  };
  for_each_active(d, p, true, s_ids, lds, [&](bool valid, int i, int v) __attribute__((always_inline)) {
    int o = 0, deg = 0, so = 0, sdeg = 0;
    VD own{};
    if (valid) {
      o = d.csc_off[v];
      deg = d.csc_off[v + 1] - o;
      so = d.csr_off[v];
      sdeg = d.csr_off[v + 1] - so;
      own = d.vdata[v];
      gathered += deg;
      ++active;
    }
    Acc<P> acc = Acc<P>::empty();
    if (valid && deg <= d.gt1) {
      int src[T1_MAX];
#pragma unroll
      for (int k = 0; k < T1_MAX; ++k)
        src[k] = (k < deg) ? d.csc_src[o + k] : 0;
      GR val[T1_MAX];
#pragma unroll
      for (int k = 0; k < T1_MAX; ++k)
        if (k < deg)
          val[k] = gather_value(d, v, src[k], o + k);
#pragma unroll
      for (int k = 0; k < T1_MAX; ++k)
        if (k < deg) acc.add(val[k]);
    }
    uint64_t m = __ballot(valid && deg > d.gt1);
    while (m) {
      const int j = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const int vj = __shfl(v, j, WAVE);
      const int oj = __shfl(o, j, WAVE);
      const int dj = __shfl(deg, j, WAVE);
      Acc<P> w = Acc<P>::empty();
      for (int base = 0; base < dj; base += WAVE * MID_UNROLL) {
        int src[MID_UNROLL];
#pragma unroll
        for (int u = 0; u < MID_UNROLL; ++u) {
          const int k = base + u * WAVE + lane;
          src[u] = (k < dj) ? d.csc_src[oj + k] : 0;
        }
        GR val[MID_UNROLL];
#pragma unroll
        for (int u = 0; u < MID_UNROLL; ++u) {
          const int k = base + u * WAVE + lane;
          if (k < dj)
            val[u] = gather_value(d, vj, src[u], oj + k);
        }
#pragma unroll
        for (int u = 0; u < MID_UNROLL; ++u)
          if (base + u * WAVE + lane < dj) w.add(val[u]);
      }
#pragma unroll
      for (int x = WAVE / 2; x > 0; x >>= 1) w.merge(w.xor_peer(x));
      if (lane == j) acc = w;
    }
    bool ret = false;
    if (valid) {
      VD x = own;
      if constexpr (has_residual<P>::value) {
        const VD before = x;
        ret = P::apply(&x, acc.finish());
        res += P::residual(before, x);
      } else {
        ret = P::apply(&x, acc.finish());
      }
      d.vnext[v] = x;
      if (ret) chg[v] = 1;
      returned += ret ? 1 : 0;
    }
    scattered += scatter_tiers(d, p, ret, i, so, sdeg, s_buf, &s_cnt, &s_base, lds);
  }, fixup);
  {
    __shared__ int lds4[4 * (NT / WAVE)];
    block_sum4(active, gathered, scattered, returned, lds4);
  }
  if (d.in_mode == IN_BITMAP && threadIdx.x == 0)
    d.bcnt[p * SCAN_BLOCKS + blockIdx.x] = active;
  add_stat(d, 2, active);
  add_stat(d, 0, gathered);
  add_stat(d, 1, scattered);
  add_stat(d, 3, returned);
  if constexpr (has_residual<P>::value) {
    res = block_sum(res, ldsd);
    if (threadIdx.x == 0) d.rpart[blockIdx.x] = res;
  }
}

// A whole superstep of a pushMin program (see is_push_min) on an all or
// bitmap frontier of the vertices whose value changed in the previous
// superstep.  State is double-buffered like k_fused: vdata holds this
// superstep's values (read only), vnext the previous superstep's values
// everywhere except where they changed, which each frontier vertex first
// lowers to its own value with atomicMin (a push may already have lowered
// it further).  Each frontier vertex then pushes gatherMap(own value,
// edge) along its out-edges: where the pushed value is below the
// destination's value in vdata, an atomicMin into vnext and a mark (plain
// byte store) that the destination changes in this superstep -- exactly
// the destinations whose final vnext value ends below vdata.  After the
// launch vnext holds min(old, pushes) everywhere: the pull's values.  The
// frontier size reported for the next superstep counts the changed
// vertices that push (they have out-edges; a changed value is never one
// pushSkip names): the pull's next frontier is their out-neighbours, so
// the two stop after the same superstep.  Out-degrees are at most WAVE (one
// thread up to st1, one wave above), or at most ELL_W with the
// out-neighbour table.
template <class P>
__global__ __launch_bounds__(NT) void k_push(Dev<P> d, int p) {
  typedef typename P::VertexData VD;
  typedef typename P::EdgeData ED;
  __shared__ int32_t s_ids[BM_WORDS * 4];
  __shared__ int lds[NT / WAVE];
  if constexpr (is_push_min<P>::value) {
    superstep_prologue(d, p, lds);
    uint8_t *fn = d.fl[p ^ 1];
    const int lane = lane_id();
    const bool all = d.in_mode == IN_ALL;
    // per-thread counters in LDS, not in locals the lambdas capture (those
    // the compiler keeps in scratch here)
    __shared__ int s_cnt3[3][NT];
    int *const c_active = &s_cnt3[0][threadIdx.x];
    int *const c_pushed = &s_cnt3[1][threadIdx.x];
    int *const c_front = &s_cnt3[2][threadIdx.x];
    *c_active = 0;
    *c_pushed = 0;
    *c_front = 0;
    // a push can only matter below the destination's value in vdata (vnext
    // ends at most there: its own frontier entry lowers it to that), and a
    // push below it makes the destination change whatever else lands, so
    // the test on vdata alone decides both the atomic and the mark; the
    // atomic's old value is not needed (no returning round trip)
    auto push_one = [&](VD x, int u) __attribute__((always_inline)) {
      if (x < d.vdata[u]) {
        atomicMin(d.vnext + u, x);
        fn[u] = 1;
      }
    };
    if (d.ell_dst) {
      // out-edges from the out-neighbour table: ids and values load with
      // the vertex's own value (after it in the first, all-vertex
      // superstep, whose vertices mostly skip)
      for_each_active(d, p, true, s_ids, lds, [&](bool valid, int i, int v) __attribute__((always_inline)) {
        VD a{};
        int e[ELL_W];
        ED w[ELL_W];
        bool push = false;
        if (valid) {
          a = d.vdata[v];
          bool skip = false;
          if (all) {
            if constexpr (has_push_skip<P>::value) skip = P::pushSkip(a);
          } else {
            atomicMin(d.vnext + v, a);
          }
          if (!skip) {
            load_ell_row(d.ell_dst, v, e);
            if constexpr (!std::is_empty<ED>::value) {
              static_assert(sizeof(ED) == 4, "table values are 4 bytes");
              int wi[ELL_W];
              load_ell_row(reinterpret_cast<const int32_t *>(d.ell_w), v, wi);
#pragma unroll
              for (int k = 0; k < ELL_W; ++k) w[k] = __builtin_bit_cast(ED, wi[k]);
            }
            push = e[0] >= 0;
          }
          ++*c_front;
          if (push) ++*c_active;
        }
        if (push) {
          // every destination's value loaded before the first store (a byte
          // store may alias them, which would serialise the loads)
          VD cur[ELL_W];
#pragma unroll
          for (int k = 0; k < ELL_W; ++k)
            if (e[k] >= 0) cur[k] = d.vdata[e[k]];
#pragma unroll
          for (int k = 0; k < ELL_W; ++k)
            if (e[k] >= 0) {
              const VD x = P::gatherMap(nullptr, &a, w + k);
              if (x < cur[k]) {
                atomicMin(d.vnext + e[k], x);
                fn[e[k]] = 1;
              }
              ++*c_pushed;
            }
        }
      });
    } else
    for_each_active(d, p, true, s_ids, lds, [&](bool valid, int i, int v) __attribute__((always_inline)) {
      VD a{};
      int so = 0, deg = 0;
      if (valid) {
        a = d.vdata[v];
        so = d.csr_off[v];
        deg = d.csr_off[v + 1] - so;
        ++*c_front;
        if (!all) atomicMin(d.vnext + v, a);  // all: the buffers are equal
      }
      bool push = valid && deg > 0;
      if constexpr (has_push_skip<P>::value)
        push = push && !P::pushSkip(a);
      if (push) ++*c_active;
      if (push && deg <= d.st1) {
        int dst[T1_MAX];
        ED w[T1_MAX];
#pragma unroll
        for (int k = 0; k < T1_MAX; ++k) {
          dst[k] = k < deg ? d.csr_dst[so + k] : 0;
          if constexpr (!std::is_empty<ED>::value)
            w[k] = k < deg ? d.edata_csr[so + k] : ED();
        }
        VD cur[T1_MAX];
#pragma unroll
        for (int k = 0; k < T1_MAX; ++k)
          if (k < deg) cur[k] = d.vdata[dst[k]];
#pragma unroll
        for (int k = 0; k < T1_MAX; ++k)
          if (k < deg) {
            const VD x = P::gatherMap(nullptr, &a, w + k);
            if (x < cur[k]) {
              atomicMin(d.vnext + dst[k], x);
              fn[dst[k]] = 1;
            }
          }
        *c_pushed += deg;
      }
      uint64_t m = __ballot(push && deg > d.st1);
      while (m) {
        const int j = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        const VD aj = __shfl(a, j, WAVE);
        const int oj = __shfl(so, j, WAVE);
        const int dj = __shfl(deg, j, WAVE);
        if (lane < dj) {
          ED w{};
          if constexpr (!std::is_empty<ED>::value) w = d.edata_csr[oj + lane];
          push_one(P::gatherMap(nullptr, &aj, &w), d.csr_dst[oj + lane]);
        }
        if (lane == j) *c_pushed += dj;
      }
    });
    int active = *c_active, pushed = *c_pushed, frontier = *c_front;
    {
      __shared__ int lds3[3 * (NT / WAVE)];
      block_sum3(active, pushed, frontier, lds3);
    }
    if (d.in_mode == IN_BITMAP && threadIdx.x == 0)
      d.bcnt[p * SCAN_BLOCKS + blockIdx.x] = active;
    add_stat(d, 2, active);
    add_stat(d, 1, pushed);
    add_stat(d, 3, frontier);
  }
}

// Scatter tiers after an unfused apply (scatterActivate, gpugas.h:669-751).
template <class P>
__global__ __launch_bounds__(NT) void k_scatter_vp(Dev<P> d, int p) {
  __shared__ int32_t s_buf[NT];
  __shared__ int32_t s_ids[BM_WORDS * 4];
  __shared__ int s_cnt, s_base;
  __shared__ int lds[NT / WAVE];
  superstep_prologue(d, p, lds);
  int scattered = 0;
  for_each_active(d, p, true, s_ids, lds, [&](bool valid, int i, int v) __attribute__((always_inline)) {
    const bool ret = valid && d.rflag[i] != 0;
    int so = 0, sdeg = 0;
    if (ret) {
      so = d.csr_off[v];
      sdeg = d.csr_off[v + 1] - so;
    }
    scattered += scatter_tiers(d, p, ret, i, so, sdeg, s_buf, &s_cnt, &s_base, lds);
  });
  scattered = block_sum(scattered, lds);
  add_stat(d, 1, scattered);
}

// Program::scatter(src, dst, edge) over the out-edges of the slots whose
// apply returned true (refgas.h:178-192), for scatterEdges programs.  An
// edge is written only by its source's call, so no atomics; vertex data
// is read after every apply of the superstep.  Out-degree <= WAVE: the
// vertex's thread, edge by edge in CSR order; larger: the whole wave,
// lanes striding over the edges.  Reads the input frontier without
// clearing it (k_scatter_vp follows).
template <class P>
__global__ __launch_bounds__(NT) void k_edge_scatter(Dev<P> d, int p) {
  __shared__ int32_t s_ids[BM_WORDS * 4];
  __shared__ int lds[NT / WAVE];
  if constexpr (scatters_edges<P>::value) {
    const int lane = lane_id();
    for_each_active(d, p, false, s_ids, lds, [&](bool valid, int i, int v) __attribute__((always_inline)) {
      const bool ret = valid && d.rflag[i] != 0;
      int so = 0, deg = 0;
      if (ret) {
        so = d.csr_off[v];
        deg = d.csr_off[v + 1] - so;
      }
      if (ret && deg <= WAVE)
        for (int k = so; k < so + deg; ++k)
          P::scatter(d.vdata + v, d.vdata + d.csr_dst[k], d.edata_rw + d.csr_csc[k]);
      uint64_t m = __ballot(ret && deg > WAVE);
      while (m) {
        const int j = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        const int vj = __shfl(v, j, WAVE);
        const int oj = __shfl(so, j, WAVE);
        const int dj = __shfl(deg, j, WAVE);
        for (int k = lane; k < dj; k += WAVE)
          P::scatter(d.vdata + vj, d.vdata + d.csr_dst[oj + k],
                     d.edata_rw + d.csr_csc[oj + k]);
      }
    });
  }
}

// Hub scatter: merge-path tiles over the scatter hubs' CSR out-edges.
template <class P>
__global__ __launch_bounds__(NT) void k_hscatter(Dev<P> d, int p) {
  __shared__ int32_t s_end[HUB_SEG_MAX];
  __shared__ int32_t s_base[HUB_SEG_MAX];
  const StepCounts &sc = d.ctl->ring[d.seq & 3];
  const int n = sc.n_shub;
  if (n == 0) return;
  uint8_t *fl = d.fl[p ^ 1];
  int32_t *qn = d.q[p ^ 1];
  int32_t *nn = &d.ctl->ring[(d.seq + 1) & 3].n_active;
  const int N = sc.shub_items + n;
  const int ntiles = (N + MP_NV - 1) / MP_NV;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int d0 = t * MP_NV;
    const int d1 = min(d0 + MP_NV, N);
    const int s0 = d.shpart[t];
    const int s1 = d.shpart[t + 1];
    const int nseg = min(s1, n - 1) - s0 + 1;
    for (int k = tid; k < nseg; k += NT) {
      const int s = s0 + k;
      const int v = slot_vertex(d, p, d.shub[s]);
      s_base[k] = d.csr_off[v] - d.shoff[s];
      s_end[k] = d.shoff[s + 1] + s;
    }
    __syncthreads();
    const int p0 = d0 + tid * MP_VT;
    int k = (p0 < d1) ? lds_lower_bound(s_end, nseg, p0) : 0;
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      const int pos = p0 + it;
      int u = -1;
      if (pos < d1) {
        if (s_end[k] == pos)
          ++k;
        else
          u = d.csr_dst[s_base[k] + (pos - s0 - k)];
      }
      if (d.out_list) {
        const bool got = u >= 0 && claim_vertex(fl, u);
        const uint64_t gm = __ballot(got);
        if (gm) {
          int at = 0;
          if (lane == 0) at = atomicAdd(nn, __popcll(gm));
          at = __shfl(at, 0, WAVE);
          if (got) qn[at + lanemask_lt_count(gm)] = u;
        }
      } else if (u >= 0) {
        fl[u] = 1;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Small-frontier mode for graphs whose degrees are all at most SMALL_DEG_MAX:
// one workgroup of SMALL_NT threads runs consecutive supersteps from a list
// frontier, with workgroup barriers between gather, apply and scatter (all
// of one superstep's gathers finish before any of its applies, as
// refgas.h:143-168 requires), while the frontier has at most small_max
// vertices.  It stops when the frontier empties, outgrows small_max, or
// after max_steps supersteps, leaving a list frontier (q, activation bytes,
// counter ring) exactly as a superstep of the grid kernels would, and
// reports what it did in the control block.  One workgroup means one CU and
// one L2: the barriers are cheap and no cross-XCD fence is needed.
constexpr int SMALL_NT = 1024;
constexpr int SMALL_DEG_MAX = 64;
template <class P, bool HaveGather>
__global__ __launch_bounds__(SMALL_NT) void k_small(Dev<P> d, int p,
                                                    int max_steps,
                                                    int small_max) {
  typedef typename P::GatherResult GR;
  typedef typename P::VertexData VD;
  __shared__ int s_n, s_next;
  Ctl *c = d.ctl;
  if (threadIdx.x == 0) {
    s_n = c->ring[d.seq & 3].n_active;
    s_next = 0;
  }
  __syncthreads();
  int cur = p, steps = 0, n = s_n;
  int64_t gathered = 0, scattered = 0, active = 0, returned = 0;
  while (n > 0 && n <= small_max && steps < max_steps) {
    const int32_t *q = d.q[cur];
    int32_t *qn = d.q[cur ^ 1];
    uint8_t *fn = d.fl[cur ^ 1];
    if (HaveGather) {
      for (int i = threadIdx.x; i < n; i += SMALL_NT) {
        const int v = q[i];
        const int o = d.csc_off[v], o1 = d.csc_off[v + 1];
        gathered += o1 - o;
        Acc<P> acc = Acc<P>::empty();
        for (int e = o; e < o1; ++e) acc.add(gather_value(d, v, d.csc_src[e], e));
        d.gtmp[i] = acc.finish();
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) s_superstep = d.iter + steps;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += SMALL_NT) {
      const int v = q[i];
      d.fl[cur][v] = 0;
      ++active;
      const GR g = HaveGather ? d.gtmp[i] : P::gatherZero;
      VD x = d.vdata[v];
      const bool ret = P::apply(&x, g);
      d.vdata[v] = x;
      if constexpr (is_source_only<P>::value)
        d.srcval[v] = P::gatherMap(&x, &x, nullptr);
      if (ret) {
        const int o = d.csr_off[v], o1 = d.csr_off[v + 1];
        ++returned;
        scattered += o1 - o;
        for (int e = o; e < o1; ++e) {
          const int u = d.csr_dst[e];
          if (claim_vertex(fn, u)) qn[atomicAdd(&s_next, 1)] = u;
        }
      }
    }
    __syncthreads();
    n = s_next;
    __syncthreads();
    if (threadIdx.x == 0) s_next = 0;
    cur ^= 1;
    ++steps;
  }
  // statistics of this workgroup's threads into slot 0
  __shared__ unsigned long long s_stat[4];
  if (threadIdx.x < 4) s_stat[threadIdx.x] = 0;
  __syncthreads();
  if (gathered) atomicAdd(&s_stat[0], (unsigned long long)gathered);
  if (scattered) atomicAdd(&s_stat[1], (unsigned long long)scattered);
  if (active) atomicAdd(&s_stat[2], (unsigned long long)active);
  if (returned) atomicAdd(&s_stat[3], (unsigned long long)returned);
  __syncthreads();
  if (threadIdx.x < 4 && s_stat[threadIdx.x])
    atomicAdd(reinterpret_cast<unsigned long long *>(d.stat_acc +
                                                     threadIdx.x * SCAN_BLOCKS),
              s_stat[threadIdx.x]);
  // the frontier left for the grid kernels: superstep seq + steps reads
  // ring entry (seq + steps) & 3; the other entries start cleared
  if (threadIdx.x < 4 * STEP_COUNT_WORDS)
    reinterpret_cast<int32_t *>(c->ring)[threadIdx.x] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    c->ring[(d.seq + steps) & 3].n_active = n;
    c->small_steps = steps;
    c->small_n = n;
    c->small_empty = n == 0;
  }
}

// k_small for programs without a gather on graphs small enough that the
// out-edges, the vertex state, both lists and both activation byte arrays
// sit in LDS (SMALL_LDS_V vertices, SMALL_LDS_E edges, 4-byte state): the
// CSR and state are staged once, every superstep then touches LDS only
// (apply, claim, append, two barriers), and the state, the list left and its
// activation bytes go back to global memory at the end -- the same
// supersteps, values, list and report as k_small.
constexpr int SMALL_LDS_V = 4096;
constexpr int SMALL_LDS_E = 16384;
constexpr int SMALL_LDS_NT = 256;  // four waves: cheap barriers
template <class P>
__global__ __launch_bounds__(SMALL_LDS_NT) void k_small_lds(Dev<P> d, int p,
                                                        int max_steps,
                                                        int small_max, int g) {
  typedef typename P::VertexData VD;
  static_assert(sizeof(VD) <= 4, "LDS-resident small mode: 4-byte state");
  __shared__ int32_t s_off[SMALL_LDS_V + 1];
  __shared__ int32_t s_dst[SMALL_LDS_E];
  __shared__ VD s_vd[SMALL_LDS_V];
  __shared__ int32_t s_q[2][SMALL_LDS_V];
  __shared__ __align__(4) uint8_t s_fl[2][SMALL_LDS_V];
  // next-list counts by superstep mod 3: superstep L appends into s_next[L %
  // 3], which every thread reads after the barrier ending L; the slot of
  // L + 2 is cleared then (its last readers read it before that barrier)
  __shared__ int s_n, s_next[3];
  Ctl *c = d.ctl;
  const int V = d.nV, E = d.nE, tid = threadIdx.x;
  // staging: eight loads per thread in flight before their stores
  auto stage = [&](auto *dst, const auto *src, int count) __attribute__((always_inline)) {
    for (int b = 0; b < count; b += 8 * SMALL_LDS_NT) {
      typename std::remove_const<typename std::remove_pointer<decltype(src)>::type>::type x[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int i = b + k * SMALL_LDS_NT + tid;
        if (i < count) x[k] = src[i];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int i = b + k * SMALL_LDS_NT + tid;
        if (i < count) dst[i] = x[k];
      }
    }
  };
  stage(s_off, d.csr_off, V + 1);
  stage(s_dst, d.csr_dst, E);
  stage(s_vd, d.vdata, V);
  for (int i = tid; i < SMALL_LDS_V / 4; i += SMALL_LDS_NT) {
    reinterpret_cast<uint32_t *>(s_fl[0])[i] = 0u;
    reinterpret_cast<uint32_t *>(s_fl[1])[i] = 0u;
  }
  if (tid == 0) {
    s_n = c->ring[d.seq & 3].n_active;
    s_next[0] = s_next[1] = s_next[2] = 0;
    s_superstep = d.iter;
  }
  __syncthreads();
  int n = s_n;
  // the input list, its activation bytes cleared in global memory (the
  // supersteps below keep theirs in LDS)
  for (int i = tid; i < n; i += SMALL_LDS_NT) {
    const int v = d.q[p][i];
    s_q[0][i] = v;
    d.fl[p][v] = 0;
  }
  __syncthreads();
  int cur = 0, steps = 0;
  int64_t scattered = 0, active = 0, returned = 0;
  while (n > 0 && n <= small_max && steps < max_steps) {
    int *nx = &s_next[steps % 3];
    // g lanes per frontier vertex (g >= its out-degree, a power of two up to
    // the wave): the group's first lane applies, every lane takes one
    // out-edge, and the wave reserves its appends with one atomic
    for (int b = 0; b < n; b += SMALL_LDS_NT / g) {
      const int i = b + tid / g, lig = tid & (g - 1);
      const bool valid = i < n;
      int v = 0;
      bool ret = false;
      if (valid) {
        v = s_q[cur][i];
        if (lig == 0) {
          s_fl[cur][v] = 0;
          ++active;
          VD x = s_vd[v];
          ret = P::apply(&x, P::gatherZero);
          s_vd[v] = x;
        }
      }
      ret = __shfl(ret ? 1 : 0, lane_id() & ~(g - 1), WAVE) != 0;
      int u = -1;
      if (ret) {
        const int o = s_off[v], o1 = s_off[v + 1];
        if (lig == 0) {
          ++returned;
          scattered += o1 - o;
        }
        if (o + lig < o1) u = s_dst[o + lig];
      }
      const bool cl = u >= 0 && claim_vertex(s_fl[cur ^ 1], u);
      const uint64_t m = __ballot(cl);
      if (m) {
        const int lead = __ffsll((unsigned long long)m) - 1;
        int base = 0;
        if (lane_id() == lead) base = atomicAdd(nx, __popcll(m));
        base = __shfl(base, lead, WAVE);
        if (cl) s_q[cur ^ 1][base + __popcll(m & ((1ull << lane_id()) - 1))] = u;
      }
    }
    __syncthreads();
    n = *nx;
    if (tid == 0) {
      s_next[(steps + 2) % 3] = 0;
      s_superstep = d.iter + steps + 1;
    }
    __syncthreads();
    cur ^= 1;
    ++steps;
  }
  // state, the list left and its activation bytes back to global memory,
  // in the parity the grid kernels continue with
  const int pf = p ^ (steps & 1);
  for (int i = tid; i < V; i += SMALL_LDS_NT) d.vdata[i] = s_vd[i];
  for (int i = tid; i < n; i += SMALL_LDS_NT) {
    const int v = s_q[cur][i];
    d.q[pf][i] = v;
    d.fl[pf][v] = 1;
  }
  __shared__ unsigned long long s_stat[4];
  if (tid < 4) s_stat[tid] = 0;
  __syncthreads();
  if (scattered) atomicAdd(&s_stat[1], (unsigned long long)scattered);
  if (active) atomicAdd(&s_stat[2], (unsigned long long)active);
  if (returned) atomicAdd(&s_stat[3], (unsigned long long)returned);
  __syncthreads();
  if (tid < 4 && s_stat[tid])
    atomicAdd(reinterpret_cast<unsigned long long *>(d.stat_acc + tid * SCAN_BLOCKS),
              s_stat[tid]);
  if (tid < 4 * STEP_COUNT_WORDS) reinterpret_cast<int32_t *>(c->ring)[tid] = 0;
  __syncthreads();
  if (tid == 0) {
    c->ring[(d.seq + steps) & 3].n_active = n;
    c->small_steps = steps;
    c->small_n = n;
    c->small_empty = n == 0;
  }
}

// Ordered compaction of the activation bytes of the next parity into the
// next list (replaces scatter_if_inputloc_twophase, scatter_if_mgpu.h:130-157),
// used by the step-by-step API so its frontier is in ascending vertex order
// as in refgas.h:201-210.  Pass 1 counts set bytes per chunk of 32-bit words;
// pass 2 writes the ids of every chunk in ascending order at its prefix and
// clears the words.
template <class P>
__global__ __launch_bounds__(NT) void k_compact1(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  const uint32_t *fw = reinterpret_cast<const uint32_t *>(d.fl[p ^ 1]);
  int b, e;
  scan_chunk(d.nfwords, blockIdx.x, b, e);
  int cnt = 0;
  for (int i = b + threadIdx.x; i < e; i += NT) cnt += __popc(fw[i]);
  cnt = block_sum(cnt, lds);
  if (threadIdx.x == 0) d.bsum[blockIdx.x] = cnt;
}

template <class P>
__global__ __launch_bounds__(NT) void k_compact2(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  uint32_t *fw = reinterpret_cast<uint32_t *>(d.fl[p ^ 1]);
  int32_t *qn = d.q[p ^ 1];
  int b, e, pre, total;
  scan_chunk(d.nfwords, blockIdx.x, b, e);
  chunk_prefix(d.bsum, blockIdx.x, pre, total, lds);
  int carry = pre;
  for (int base = b; base < e; base += NT) {
    const int wi = base + threadIdx.x;
    uint32_t word = (wi < e) ? fw[wi] : 0u;
    int tot;
    int pos = carry + block_excl_scan(__popc(word), tot, lds);
    if (word) {
      fw[wi] = 0u;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if ((word >> (8 * k)) & 0xffu) qn[pos++] = wi * 4 + k;
    }
    carry += tot;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    d.ctl->ring[(d.seq + 1) & 3].n_active = total;
}

// gatherMap(v) of every vertex, after vertex state changed outside apply.
template <class P>
__global__ void k_srcval_all(Dev<P> d) {
  if constexpr (is_source_only<P>::value) {
    for (int64_t v = blockIdx.x * (int64_t)NT + threadIdx.x; v < d.nV;
         v += (int64_t)gridDim.x * NT)
      d.srcval[v] = P::gatherMap(d.vdata + v, d.vdata + v, nullptr);
  }
}

// Zeroes `words` 32-bit words (the engine's clears when it must not call
// the runtime's memset).
template <class W>
__global__ void k_zero_words(W *p, int64_t words) {
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < words;
       i += (int64_t)gridDim.x * NT)
    p[i] = W(0);
}

// Vertex-data initialisation on the device: `value` everywhere, `one` at
// vertex `at` (none if at < 0).  Replaces an upload of the whole array when
// the initial state is uniform, as it is for BFS and SSSP.
// `also` (if not null) gets the same values: the other state buffer of a
// double-buffered run, which then needs no copy before its first superstep.
template <class P>
__global__ void k_fill_vdata(Dev<P> d, typename P::VertexData value,
                             int32_t at, typename P::VertexData one,
                             typename P::VertexData *also) {
  for (int64_t v = blockIdx.x * (int64_t)NT + threadIdx.x; v < d.nV;
       v += (int64_t)gridDim.x * NT) {
    const typename P::VertexData x = v == at ? one : value;
    d.vdata[v] = x;
    if (also) also[v] = x;
  }
}

// Run totals: slot k of the per-workgroup counters summed over the
// workgroups into out[k] (host-mapped memory, system-scope stores), one
// workgroup in a fixed order; every load is issued before the first sum.
template <class P>
__global__ __launch_bounds__(NT) void k_stat_totals(Dev<P> d, int64_t *out) {
  static_assert(SCAN_BLOCKS % NT == 0, "whole rounds of the workgroup");
  constexpr int PER = SCAN_BLOCKS / NT;
  __shared__ int64_t lds[STAT_SLOTS][NT / WAVE];
  int64_t x[STAT_SLOTS][PER];
#pragma unroll
  for (int k = 0; k < STAT_SLOTS; ++k)
#pragma unroll
    for (int i = 0; i < PER; ++i) x[k][i] = d.stat_acc[k * SCAN_BLOCKS + i * NT + threadIdx.x];
#pragma unroll
  for (int k = 0; k < STAT_SLOTS; ++k) {
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) s += x[k][i];
    for (int o = WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, WAVE);
    if (lane_id() == 0) lds[k][threadIdx.x / WAVE] = s;
  }
  __syncthreads();
  if (threadIdx.x < STAT_SLOTS) {
    int64_t t = 0;
    for (int w = 0; w < NT / WAVE; ++w) t += lds[threadIdx.x][w];
    __hip_atomic_store(out + threadIdx.x, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

}  // namespace gas
