// gfx950 kernels of the GAS superstep, templated on the vertex program.
//
// One superstep of the reference engine (gpugas.h:421-751, semantics of
// refgas.h:143-210) runs here as a fixed sequence of kernels whose sizes
// all come from a device-resident control block, so the host never waits
// for a frontier size.  Supersteps alternate between two parities p = 0, 1:
// the active list of a superstep with parity p lives in queue p and its
// size in n_active[p]; the superstep builds the next list in queue p^1.  The
// host passes p as a kernel argument.
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
//   k_apply              Program::apply; fused with the
//                        thread/wave tiers of the scatter inside run(), or
//                        recording the return flag for scatterActivate()
//   k_scatter_vp         scatter tiers when apply was not fused
//   k_sscan1/2, k_hscatter          hub scatter (skipped without hubs)
//   k_compact1/2         ordered compaction of the byte activation flags
//                        into the next active list
//
// Activation writes one byte per destination with a plain
// store: every writer stores the same value, so no atomics are needed, and
// the compaction yields the next list in ascending vertex order, as
// refgas.h:201-210 does.  Wave size is 64.  No MFMA: there is no dense
// contraction on this path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace gas {

constexpr int NT = 256;            // threads per workgroup: 4 waves of 64
constexpr int WAVE = 64;
constexpr int SCAN_BLOCKS = 1024;  // chunks of per-vertex kernels and scans
constexpr int MP_VT = 8;           // merge-path items per thread
constexpr int MP_NV = NT * MP_VT;  // merge-path items per tile
constexpr int MP_BLOCKS = 2048;    // grid of the merge-path kernels
constexpr int HUB_MIN_T2 = 4;      // hubs have degree > t2 >= 4 ...
// ... so a segment spans >= 6 merge-path items and a tile touches at most
constexpr int HUB_SEG_MAX = MP_NV / (HUB_MIN_T2 + 2) + 2;  // segments
constexpr int T1_MAX = 16;         // largest thread tier (unrolled loads)
constexpr int MID_UNROLL = 8;      // wave tier: edges per lane per trip
constexpr int DEFAULT_T1 = 8;      // thread tier: degree <= t1
constexpr int DEFAULT_T2 = 8;      // wave tier: degree <= t2; above: hubs

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
// Control block, one per engine, in device memory.
struct Ctl {
  int32_t n_active[2];    // |A| of the superstep of each parity
  int32_t all_active[2];  // 1: A is implicitly [0, V)
  int32_t iteration;      // supersteps completed (BFS level)
  int32_t n_ghub;         // gather hubs appended this superstep
  int32_t ghub_items;     // their summed in-degree
  int32_t n_shub;         // scatter hubs appended this superstep
  int32_t shub_items;     // their summed out-degree
  int32_t pad[3];
  int64_t stat_gathered;  // totals over the run
  int64_t stat_scattered;
  int64_t stat_active;
  int64_t stat_supersteps;
};

// Optional program trait: `static constexpr bool compensatedSum = true`
// selects TwoSum-compensated accumulation of a float sum.
template <class P, class = void>
struct is_compensated : std::false_type {};
template <class P>
struct is_compensated<P, std::void_t<decltype(P::compensatedSum)>>
    : std::integral_constant<bool, P::compensatedSum> {};

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
// 4.2.2; Neumaier's variant for merging two partial sums).
template <class P>
struct Acc<P, true> {
  typedef typename P::GatherResult T;
  static_assert(std::is_same<T, float>::value, "compensated sum needs float");
  float s, c;
  int ok;
  __device__ static Acc empty() {
    Acc a;
    a.s = 0.f;
    a.c = 0.f;
    a.ok = 0;
    return a;
  }
  __device__ static float two_sum(float a, float b, float &err) {
    float t = a + b;
    float bp = t - a;
    err = (a - (t - bp)) + (b - bp);
    return t;
  }
  __device__ void add(float x) {
    if (!ok) {
      s = x;
      c = 0.f;
      ok = 1;
      return;
    }
    float err;
    s = two_sum(s, x, err);
    c += err;
  }
  __device__ void merge(const Acc &b) {
    if (!b.ok) return;
    if (!ok) {
      *this = b;
      return;
    }
    float err;
    s = two_sum(s, b.s, err);
    c = c + b.c + err;
  }
  __device__ Acc xor_peer(int m) const {
    Acc r;
    r.s = __shfl_xor(s, m, WAVE);
    r.c = __shfl_xor(c, m, WAVE);
    r.ok = __shfl_xor(ok, m, WAVE);
    return r;
  }
  __device__ Acc up_peer(int o) const {
    Acc r;
    r.s = __shfl_up(s, o, WAVE);
    r.c = __shfl_up(c, o, WAVE);
    r.ok = __shfl_up(ok, o, WAVE);
    return r;
  }
  __device__ T finish() const {
    T z = P::gatherZero;
    return ok ? P::gatherReduce(z, s + c) : z;
  }
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
  VD *vdata;
  Ctl *ctl;
  int32_t *q[2];     // active lists by parity
  uint8_t *flags;    // activation byte per vertex (V rounded up to 4)
  int32_t nfwords;   // flags as 32-bit words
  uint8_t *rflag;    // apply's return per active slot (unfused path)
  GR *gtmp;          // gather result per active slot
  int32_t gt1, gt2;  // gather degree tiers (thread <= gt1 < wave <= gt2 < hub)
  int32_t st1, st2;  // scatter degree tiers
  // gather hubs of the current superstep (queue mode): active slots
  int32_t *ghub, *ghoff, *ghpart;
  // gather hubs when every vertex is active (fixed per graph): vertex ids
  const int32_t *ahub, *ahoff, *ahpart;
  int32_t n_ahub, ahub_items;
  // scatter hubs of the current superstep: active slots
  int32_t *shub, *shoff, *shpart;
  int32_t *bsum;     // per-chunk sums of the two-pass scans
  int32_t *stat_blk; // per-workgroup counters: gathered | scattered
  double *rpart;     // per-chunk FP64 residual sums
  Acc<P> *part_first, *part_last;  // per-tile partials of crossing segments
};

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

// ---------------------------------------------------------------------------
// Frontier initialisation: setActive(start, end) (gpugas.h:342-349); clears
// the activation flags and the hub counters.
template <class P>
__global__ __launch_bounds__(NT) void k_set_active(Dev<P> d, int p, int start,
                                                   int end) {
  Ctl *c = d.ctl;
  const bool all = (start == 0 && end == d.nV);
  const int n = end - start;
  int32_t *q = d.q[p];
  const int stride = gridDim.x * NT;
  uint32_t *fw = reinterpret_cast<uint32_t *>(d.flags);
  for (int i = blockIdx.x * NT + threadIdx.x; i < d.nfwords; i += stride)
    fw[i] = 0u;
  if (!all)
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += stride)
      q[i] = start + i;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c->n_active[p] = n;
    c->all_active[p] = all ? 1 : 0;
    c->n_ghub = 0;
    c->n_shub = 0;
  }
}

// Thread and wave tiers of the gather; hub slots are appended to ghub
// (queue mode) or skipped (all active: the fixed hub list covers them).
template <class P>
__global__ __launch_bounds__(NT) void k_gather_vp(Dev<P> d, int p) {
  typedef typename P::GatherResult GR;
  __shared__ int32_t s_buf[NT];
  __shared__ int s_cnt, s_base;
  __shared__ int lds[NT / WAVE];
  Ctl *c = d.ctl;
  const int A = c->n_active[p];
  if (A == 0) return;
  const bool all = c->all_active[p] != 0;
  const int32_t *q = d.q[p];
  const int stride = gridDim.x * NT;
  const int rounds = (A + stride - 1) / stride;
  const int lane = lane_id();
  int gathered = 0;
  for (int r = 0; r < rounds; ++r) {
    const int i = r * stride + blockIdx.x * NT + threadIdx.x;
    const bool valid = i < A;
    int v = 0, o = 0, deg = 0;
    if (valid) {
      v = all ? i : q[i];
      o = d.csc_off[v];
      deg = d.csc_off[v + 1] - o;
      gathered += deg;
    }
    const bool small = valid && deg <= d.gt1;
    const bool mid = valid && deg > d.gt1 && deg <= d.gt2;
    const bool hub = valid && deg > d.gt2;
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
          val[k] = P::gatherMap(d.vdata + v, d.vdata + src[k], d.edata + o + k);
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
            val[u] = P::gatherMap(d.vdata + vj, d.vdata + src[u],
                                  d.edata + oj + k);
        }
#pragma unroll
        for (int u = 0; u < MID_UNROLL; ++u)
          if (base + u * WAVE + lane < dj) acc.add(val[u]);
      }
#pragma unroll
      for (int x = WAVE / 2; x > 0; x >>= 1) acc.merge(acc.xor_peer(x));
      if (lane == j) d.gtmp[i] = acc.finish();
    }
    if (!all) block_append(hub, i, d.ghub, &c->n_ghub, s_buf, &s_cnt, &s_base);
  }
  gathered = block_sum(gathered, lds);
  if (threadIdx.x == 0) d.stat_blk[blockIdx.x] = gathered;
}

// Hub list of a superstep: all-active supersteps use the fixed list.
template <class P>
struct HubList {
  const int32_t *list, *off, *part;
  int n, items;
};

template <class P>
__device__ inline HubList<P> gather_hubs(const Dev<P> &d, const Ctl *c,
                                         bool all) {
  HubList<P> h;
  if (all) {
    h.list = d.ahub;
    h.off = d.ahoff;
    h.part = d.ahpart;
    h.n = d.n_ahub;
    h.items = d.ahub_items;
  } else {
    h.list = d.ghub;
    h.off = d.ghoff;
    h.part = d.ghpart;
    h.n = c->n_ghub;
    h.items = c->ghub_items;
  }
  return h;
}

// Degree scan over a hub list (pass 1: chunk sums).  Which = 0: gather hubs
// (in-degrees), 1: scatter hubs (out-degrees).
template <class P, int Which>
__global__ __launch_bounds__(NT) void k_hscan1(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  const Ctl *c = d.ctl;
  if (c->n_active[p] == 0) return;
  const bool all = c->all_active[p] != 0;
  if (Which == 0 && all) return;  // fixed hub list, scanned on the host
  const int n = Which == 0 ? c->n_ghub : c->n_shub;
  const int32_t *list = Which == 0 ? d.ghub : d.shub;
  const int32_t *gofs = Which == 0 ? d.csc_off : d.csr_off;
  const int32_t *q = d.q[p];
  int b, e;
  scan_chunk(n, blockIdx.x, b, e);
  int sum = 0;
  for (int i = b + threadIdx.x; i < e; i += NT) {
    const int slot = list[i];
    const int v = all ? slot : q[slot];
    sum += gofs[v + 1] - gofs[v];
  }
  sum = block_sum(sum, lds);
  if (threadIdx.x == 0) d.bsum[blockIdx.x] = sum;
}

// Pass 2: offsets, tile partition and total of the hub list.
template <class P, int Which>
__global__ __launch_bounds__(NT) void k_hscan2(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  Ctl *c = d.ctl;
  if (c->n_active[p] == 0) return;
  const bool all = c->all_active[p] != 0;
  if (Which == 0 && all) return;
  const int n = Which == 0 ? c->n_ghub : c->n_shub;
  if (n == 0) return;
  const int32_t *list = Which == 0 ? d.ghub : d.shub;
  int32_t *off = Which == 0 ? d.ghoff : d.shoff;
  int32_t *part = Which == 0 ? d.ghpart : d.shpart;
  const int32_t *gofs = Which == 0 ? d.csc_off : d.csr_off;
  const int32_t *q = d.q[p];
  int b, e, pre, total;
  scan_chunk(n, blockIdx.x, b, e);
  chunk_prefix(d.bsum, blockIdx.x, pre, total, lds);
  int carry = pre;
  for (int base = b; base < e; base += NT) {
    const int i = base + threadIdx.x;
    int deg = 0;
    if (i < e) {
      const int slot = list[i];
      const int v = all ? slot : q[slot];
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
      c->ghub_items = total;
    else
      c->shub_items = total;
  }
}

// Hub gather: merge-path tiles over the hubs' CSC in-edges with the
// segmented reduce fused in (replaces kGatherMap + ReduceByKey,
// gpugas_kernels.cuh:53-134, gpugas.h:466-500).  A hub finished inside a
// tile is written to gtmp straight away; for a hub crossing tiles each tile
// leaves the partial before its first end (part_first) and after its last
// end (part_last), and k_hfix folds them.
template <class P>
__global__ __launch_bounds__(NT) void k_hgather(Dev<P> d, int p) {
  typedef typename P::GatherResult GR;
  __shared__ int32_t s_end[HUB_SEG_MAX];
  __shared__ int32_t s_base[HUB_SEG_MAX];
  __shared__ int32_t s_vtx[HUB_SEG_MAX];
  __shared__ int32_t s_slot[HUB_SEG_MAX];
  __shared__ SegItem<P> s_wtot[NT / WAVE];

  const Ctl *c = d.ctl;
  const int A = c->n_active[p];
  if (A == 0) return;
  const bool all = c->all_active[p] != 0;
  const HubList<P> h = gather_hubs(d, c, all);
  if (h.n == 0) return;
  const int32_t *q = d.q[p];
  const int N = h.items + h.n;
  const int ntiles = (N + MP_NV - 1) / MP_NV;
  const int tid = threadIdx.x;

  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int d0 = t * MP_NV;
    const int d1 = min(d0 + MP_NV, N);
    const int s0 = h.part[t];
    const int s1 = h.part[t + 1];
    const int nseg = min(s1, h.n - 1) - s0 + 1;
    for (int k = tid; k < nseg; k += NT) {
      const int s = s0 + k;
      const int slot = h.list[s];
      const int v = all ? slot : q[slot];
      s_vtx[k] = v;
      s_slot[k] = slot;
      s_base[k] = d.csc_off[v] - h.off[s];
      s_end[k] = h.off[s + 1] + s;
    }
    __syncthreads();
    const bool first_fresh = (h.off[s0] + s0 == d0);

    // decode this thread's items
    const int p0 = d0 + tid * MP_VT;
    int k = (p0 < d1) ? lds_lower_bound(s_end, nseg, p0) : 0;
    int kind[MP_VT], kk[MP_VT], eidx[MP_VT];
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      const int pos = p0 + it;
      kind[it] = 0;
      kk[it] = 0;
      eidx[it] = 0;
      if (pos < d1) {
        kk[it] = k;
        if (s_end[k] == pos) {
          kind[it] = 2;
          ++k;
        } else {
          kind[it] = 1;
          eidx[it] = s_base[k] + (pos - s0 - k);
        }
      }
    }
    int src[MP_VT];
#pragma unroll
    for (int it = 0; it < MP_VT; ++it)
      src[it] = (kind[it] == 1) ? d.csc_src[eidx[it]] : 0;
    GR val[MP_VT];
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      if (kind[it] == 1)
        val[it] = P::gatherMap(d.vdata + s_vtx[kk[it]], d.vdata + src[it],
                               d.edata + eidx[it]);
      else
        val[it] = GR();
    }
    // thread-serial segmented reduce
    Acc<P> acc = Acc<P>::empty(), head = Acc<P>::empty();
    int nend = 0, first_k = -1;
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      if (kind[it] == 1) {
        acc.add(val[it]);
      } else if (kind[it] == 2) {
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
template <class P>
__global__ __launch_bounds__(NT) void k_hfix(Dev<P> d, int p) {
  const Ctl *c = d.ctl;
  if (c->n_active[p] == 0) return;
  const bool all = c->all_active[p] != 0;
  const HubList<P> h = gather_hubs(d, c, all);
  for (int s = blockIdx.x * NT + threadIdx.x; s < h.n; s += gridDim.x * NT) {
    const int st = h.off[s] + s;
    const int en = h.off[s + 1] + s;
    const int t0 = st / MP_NV, t1 = en / MP_NV;
    if (t0 == t1) continue;
    Acc<P> a = (st > t0 * MP_NV) ? d.part_last[t0] : d.part_first[t0];
    for (int t = t0 + 1; t <= t1; ++t) a.merge(d.part_first[t]);
    d.gtmp[h.list[s]] = a.finish();
  }
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
// true (`ret`); hubs go to the scatter hub list.  Reached by every thread.
template <class P>
__device__ inline int scatter_tiers(const Dev<P> &d, Ctl *c, bool ret,
                                    int slot, int v, int32_t *s_buf,
                                    int *s_cnt, int *s_base) {
  int o = 0, deg = 0;
  if (ret) {
    o = d.csr_off[v];
    deg = d.csr_off[v + 1] - o;
  }
  if (ret && deg <= d.st1) {
#pragma unroll 4
    for (int k = 0; k < deg; ++k) d.flags[d.csr_dst[o + k]] = 1;
  }
  const bool mid = ret && deg > d.st1 && deg <= d.st2;
  uint64_t m = __ballot(mid);
  const int lane = lane_id();
  while (m) {
    const int j = __ffsll((unsigned long long)m) - 1;
    m &= m - 1;
    const int oj = __shfl(o, j, WAVE);
    const int dj = __shfl(deg, j, WAVE);
    for (int k = lane; k < dj; k += WAVE) d.flags[d.csr_dst[oj + k]] = 1;
  }
  block_append(ret && deg > d.st2, slot, d.shub, &c->n_shub, s_buf, s_cnt,
               s_base);
  return ret ? deg : 0;
}

// Apply (kApply, gpugas_kernels.cuh:139-151).  Fuse = true also runs the
// scatter tiers (the whole superstep is being executed, so scatterActivate
// follows); otherwise the return flag is kept for k_scatter_vp.  Resets the
// next parity's list size, which k_compact2 fills.
template <class P, bool HaveGather, bool Fuse>
__global__ __launch_bounds__(NT) void k_apply(Dev<P> d, int p) {
  typedef typename P::GatherResult GR;
  typedef typename P::VertexData VD;
  __shared__ int32_t s_buf[NT];
  __shared__ int s_cnt, s_base;
  __shared__ int lds[NT / WAVE];
  __shared__ double ldsd[NT / WAVE];
  Ctl *c = d.ctl;
  const int A = c->n_active[p];
  // The next parity starts empty even when this superstep is: otherwise a
  // superstep enqueued after the frontier emptied would re-run the list
  // left in the other queue two supersteps ago.
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c->n_active[p ^ 1] = 0;
    c->all_active[p ^ 1] = 0;
  }
  if (A == 0) return;
  if (threadIdx.x == 0) s_superstep = c->iteration;
  __syncthreads();
  const bool all = c->all_active[p] != 0;
  const int32_t *q = d.q[p];
  const int stride = gridDim.x * NT;
  const int rounds = (A + stride - 1) / stride;
  int scattered = 0;
  double res = 0.0;
  for (int r = 0; r < rounds; ++r) {
    const int i = r * stride + blockIdx.x * NT + threadIdx.x;
    const bool valid = i < A;
    int v = 0;
    bool ret = false;
    if (valid) {
      v = all ? i : q[i];
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
    }
    if (Fuse) {
      scattered += scatter_tiers(d, c, ret, i, v, s_buf, &s_cnt, &s_base);
    } else if (valid) {
      d.rflag[i] = ret ? 1 : 0;
    }
  }
  if (!HaveGather && threadIdx.x == 0) d.stat_blk[blockIdx.x] = 0;
  scattered = block_sum(scattered, lds);
  if (threadIdx.x == 0) d.stat_blk[SCAN_BLOCKS + blockIdx.x] = scattered;
  if constexpr (has_residual<P>::value) {
    res = block_sum(res, ldsd);
    if (threadIdx.x == 0) d.rpart[blockIdx.x] = res;
  }
}

// Scatter tiers after an unfused apply (scatterActivate, gpugas.h:669-751).
template <class P>
__global__ __launch_bounds__(NT) void k_scatter_vp(Dev<P> d, int p) {
  __shared__ int32_t s_buf[NT];
  __shared__ int s_cnt, s_base;
  __shared__ int lds[NT / WAVE];
  Ctl *c = d.ctl;
  const int A = c->n_active[p];
  if (A == 0) return;
  const bool all = c->all_active[p] != 0;
  const int32_t *q = d.q[p];
  const int stride = gridDim.x * NT;
  const int rounds = (A + stride - 1) / stride;
  int scattered = 0;
  for (int r = 0; r < rounds; ++r) {
    const int i = r * stride + blockIdx.x * NT + threadIdx.x;
    const bool valid = i < A;
    const bool ret = valid && d.rflag[i] != 0;
    const int v = valid ? (all ? i : q[i]) : 0;
    scattered += scatter_tiers(d, c, ret, i, v, s_buf, &s_cnt, &s_base);
  }
  scattered = block_sum(scattered, lds);
  if (threadIdx.x == 0) d.stat_blk[SCAN_BLOCKS + blockIdx.x] = scattered;
}

// Hub scatter: merge-path tiles over the scatter hubs' CSR out-edges.
template <class P>
__global__ __launch_bounds__(NT) void k_hscatter(Dev<P> d, int p) {
  __shared__ int32_t s_end[HUB_SEG_MAX];
  __shared__ int32_t s_base[HUB_SEG_MAX];
  const Ctl *c = d.ctl;
  const int A = c->n_active[p];
  if (A == 0) return;
  const int n = c->n_shub;
  if (n == 0) return;
  const bool all = c->all_active[p] != 0;
  const int32_t *q = d.q[p];
  const int N = c->shub_items + n;
  const int ntiles = (N + MP_NV - 1) / MP_NV;
  const int tid = threadIdx.x;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int d0 = t * MP_NV;
    const int d1 = min(d0 + MP_NV, N);
    const int s0 = d.shpart[t];
    const int s1 = d.shpart[t + 1];
    const int nseg = min(s1, n - 1) - s0 + 1;
    for (int k = tid; k < nseg; k += NT) {
      const int s = s0 + k;
      const int slot = d.shub[s];
      const int v = all ? slot : q[slot];
      s_base[k] = d.csr_off[v] - d.shoff[s];
      s_end[k] = d.shoff[s + 1] + s;
    }
    __syncthreads();
    const int p0 = d0 + tid * MP_VT;
    int k = (p0 < d1) ? lds_lower_bound(s_end, nseg, p0) : 0;
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      const int pos = p0 + it;
      if (pos < d1) {
        if (s_end[k] == pos)
          ++k;
        else
          d.flags[d.csr_dst[s_base[k] + (pos - s0 - k)]] = 1;
      }
    }
    __syncthreads();
  }
}

// Ordered compaction of the activation flags (replaces
// scatter_if_inputloc_twophase, scatter_if_mgpu.h:130-157): pass 1 counts
// set flags per chunk of 32-bit flag words and ends the superstep's
// bookkeeping; pass 2 writes the ids of every chunk in ascending order at
// its prefix and clears the words.
template <class P>
__global__ __launch_bounds__(NT) void k_compact1(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  Ctl *c = d.ctl;
  const int A = c->n_active[p];
  if (A == 0) return;
  const uint32_t *fw = reinterpret_cast<const uint32_t *>(d.flags);
  int b, e;
  scan_chunk(d.nfwords, blockIdx.x, b, e);
  int cnt = 0;
  for (int i = b + threadIdx.x; i < e; i += NT) cnt += __popc(fw[i]);
  cnt = block_sum(cnt, lds);
  if (threadIdx.x == 0) d.bsum[blockIdx.x] = cnt;
  if (blockIdx.x == 0) {
    // superstep totals (all per-chunk counters were written by earlier
    // kernels of this superstep)
    int g = 0, s = 0;
    for (int i = threadIdx.x; i < SCAN_BLOCKS; i += NT) {
      g += d.stat_blk[i];
      s += d.stat_blk[SCAN_BLOCKS + i];
    }
    const int gi = block_sum(g, lds);
    const int si = block_sum(s, lds);
    if (threadIdx.x == 0) {
      c->stat_gathered += gi;
      c->stat_scattered += si;
      c->stat_active += A;
      c->stat_supersteps += 1;
      c->iteration += 1;
      c->n_ghub = 0;
      c->n_shub = 0;
    }
  }
}

template <class P>
__global__ __launch_bounds__(NT) void k_compact2(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  Ctl *c = d.ctl;
  if (c->n_active[p] == 0) return;
  uint32_t *fw = reinterpret_cast<uint32_t *>(d.flags);
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
  if (blockIdx.x == 0 && threadIdx.x == 0) c->n_active[p ^ 1] = total;
}

// Vertex-data initialisation on the device: `value` everywhere, `one` at
// vertex `at` (none if at < 0).  Replaces an upload of the whole array when
// the initial state is uniform, as it is for BFS and SSSP.
template <class P>
__global__ void k_fill_vdata(Dev<P> d, typename P::VertexData value,
                             int32_t at, typename P::VertexData one) {
  for (int64_t v = blockIdx.x * (int64_t)NT + threadIdx.x; v < d.nV;
       v += (int64_t)gridDim.x * NT)
    d.vdata[v] = v == at ? one : value;
}

// Superstep counter write (host-side setIteration).
template <class P>
__global__ void k_set_iteration(Dev<P> d, int it) {
  if (threadIdx.x == 0) d.ctl->iteration = it;
}

}  // namespace gas
