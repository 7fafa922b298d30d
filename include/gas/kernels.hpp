// gfx950 kernels of the GAS superstep, templated on the vertex program.
//
// One superstep of the reference engine (gpugas.h:421-751, semantics of
// refgas.h:143-210) runs here as a fixed sequence of kernels whose sizes
// all come from a device-resident control block, so a superstep can be
// launched (or captured in a hipGraph) without the host knowing the
// frontier size.  Supersteps alternate between two parities p = 0, 1: the
// active list of a superstep with parity p lives in queue p and its size in
// n_active[p]; the superstep builds the next list in queue p^1.  The host
// passes p as a kernel argument, so no kernel has to publish a superstep's
// end to the others.
//
//   k_gscan1 / k_gscan2  exclusive scan of the active vertices' in-degrees
//                        plus the merge-path tile partition (skipped when
//                        every vertex is active: the CSC offsets already
//                        are that scan and the partition is precomputed)
//   k_gather             merge-path gather over the active vertices' CSC
//                        in-edges with the segmented reduce fused in; a
//                        segment that crosses a tile boundary leaves its
//                        per-tile partials behind
//   k_apply              finishes crossing segments from the tile partials,
//                        runs Program::apply, records the predicated
//                        out-degree of each returning vertex (the
//                        PredicatedEdgeCountIterator of gpugas.h:387-408)
//   k_sscan2             exclusive scan of those out-degrees + partition
//   k_scatter            merge-path walk of the returning vertices' CSR
//                        out-edges setting frontier bits; in queue mode the
//                        first setter of a bit also appends the vertex
//   k_compact1 / 2       bitmap mode only: ordered compaction of the
//                        frontier bitmap into the next active list
//                        (popcount + LDS block scan, no atomics)
//
// Wave size is 64 (CDNA).  No MFMA: there is no dense contraction here.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace gas {

constexpr int NT = 256;             // threads per workgroup: 4 waves of 64
constexpr int WAVE = 64;
constexpr int SCAN_BLOCKS = 1024;   // chunks of the two-pass scans
constexpr int MP_VT = 8;            // merge-path items per thread
constexpr int MP_NV = NT * MP_VT;   // merge-path items per tile
constexpr int MP_BLOCKS = 1024;     // grid of the merge-path kernels

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
  int32_t gather_items;   // sum of in-degrees over A
  int32_t scatter_items;  // sum of predicated out-degrees over A
  int32_t next_dense;     // 1: next frontier via bitmap + compaction
  int32_t dense_min;      // scatter_items at or above which bitmap mode runs
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

// Partial gather result with an explicit "holds a value" flag, so partial
// sums never need an identity element: the seed gatherZero is folded in
// exactly once when a segment is finished (refgas.h:148-158).
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
  int32_t *q[2];    // active lists by parity
  uint32_t *bm;     // frontier bitmap, (V+31)/32 words
  int32_t nwords;
  int32_t *goff;    // scan of in-degrees over A (V+1)
  int32_t *soff;    // scan of predicated out-degrees over A (V+1)
  int32_t *pdeg;    // predicated out-degree per active slot
  int32_t *bsum;    // per-chunk sums of the two-pass scans
  int32_t *gpart;   // first segment of every gather tile (queue mode)
  int32_t *apart;   // same for the all-active gather (fixed per graph)
  int32_t *spart;   // same for the scatter
  double *rpart;    // per-chunk FP64 residual sums
  GR *gtmp;         // gather result per active slot
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

// Chunk [b, e) of n items handled by workgroup `blk` in the two-pass scans.
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

// Merge-path layout: segment i (0 <= i < A) occupies the positions
// off[i]+i .. off[i+1]+i -- its edges, then its end item.  Tile t covers
// positions [t*MP_NV, (t+1)*MP_NV); part[t] is the segment holding the
// tile's first position and part[ntiles] = A.  The writer of segment i
// records itself for every tile that starts inside it.
__device__ inline void mark_tiles(int32_t *part, int i, int off_i, int deg) {
  const int start = off_i + i;
  const int end = off_i + deg + i;
  for (int t = (start + MP_NV - 1) / MP_NV; t * MP_NV <= end; ++t)
    part[t] = i;
}

// ---------------------------------------------------------------------------
// Frontier initialisation: setActive(start, end) (gpugas.h:342-349) plus
// clearing the frontier bitmap.
template <class P>
__global__ __launch_bounds__(NT) void k_set_active(Dev<P> d, int p, int start,
                                                   int end) {
  Ctl *c = d.ctl;
  const bool all = (start == 0 && end == d.nV);
  const int n = end - start;
  int32_t *q = d.q[p];
  const int stride = gridDim.x * NT;
  for (int i = blockIdx.x * NT + threadIdx.x; i < d.nwords; i += stride)
    d.bm[i] = 0u;
  if (!all)
    for (int i = blockIdx.x * NT + threadIdx.x; i < n; i += stride)
      q[i] = start + i;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c->n_active[p] = n;
    c->all_active[p] = all ? 1 : 0;
  }
}

// Tile partition of the all-active gather (depends on the graph only).
template <class P>
__global__ __launch_bounds__(NT) void k_static_partition(Dev<P> d) {
  const int V = d.nV;
  for (int i = blockIdx.x * NT + threadIdx.x; i < V; i += gridDim.x * NT) {
    int o = d.csc_off[i];
    mark_tiles(d.apart, i, o, d.csc_off[i + 1] - o);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int N = d.nE + V;
    d.apart[(N + MP_NV - 1) / MP_NV] = V;
  }
}

// Pass 1 of the in-degree scan over the active list.
template <class P>
__global__ __launch_bounds__(NT) void k_gscan1(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  const Ctl *c = d.ctl;
  const int A = c->n_active[p];
  if (A == 0 || c->all_active[p]) return;
  const int32_t *q = d.q[p];
  int b, e;
  scan_chunk(A, blockIdx.x, b, e);
  int sum = 0;
  for (int i = b + threadIdx.x; i < e; i += NT) {
    int v = q[i];
    sum += d.csc_off[v + 1] - d.csc_off[v];
  }
  sum = block_sum(sum, lds);
  if (threadIdx.x == 0) d.bsum[blockIdx.x] = sum;
}

// Pass 2 of the in-degree scan: goff[i] = sum of in-degrees of A[0..i),
// and the gather tile partition.
template <class P>
__global__ __launch_bounds__(NT) void k_gscan2(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  Ctl *c = d.ctl;
  const int A = c->n_active[p];
  if (A == 0) return;
  if (c->all_active[p]) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      c->gather_items = d.nE;
      c->stat_gathered += d.nE;
    }
    return;
  }
  const int32_t *q = d.q[p];
  int b, e, pre, total;
  scan_chunk(A, blockIdx.x, b, e);
  chunk_prefix(d.bsum, blockIdx.x, pre, total, lds);
  int carry = pre;
  for (int base = b; base < e; base += NT) {
    int i = base + threadIdx.x;
    int deg = 0;
    if (i < e) {
      int v = q[i];
      deg = d.csc_off[v + 1] - d.csc_off[v];
    }
    int tot;
    int ex = block_excl_scan(deg, tot, lds);
    if (i < e) {
      d.goff[i] = carry + ex;
      mark_tiles(d.gpart, i, carry + ex, deg);
    }
    carry += tot;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    d.goff[A] = total;
    d.gpart[(total + A + MP_NV - 1) / MP_NV] = A;
    c->gather_items = total;
    c->stat_gathered += total;
  }
}

// Merge-path gather + segmented reduce (north-star (a)+(b); replaces
// kGatherMap + ReduceByKey, gpugas_kernels.cuh:53-134, gpugas.h:466-500).
// Each tile takes MP_NV consecutive items of the path formed by every
// active vertex's in-edges followed by its end marker, so a hub's edge list
// is split across as many tiles as it needs while low-degree vertices share
// one.  A segment finished inside a tile is written straight to gtmp; for
// a segment crossing tiles each tile leaves the partial before its first
// end (part_first) and after its last end (part_last), and k_apply folds
// them.
template <class P>
__global__ __launch_bounds__(NT) void k_gather(Dev<P> d, int p) {
  typedef typename P::GatherResult GR;
  __shared__ int32_t s_end[MP_NV + 1];
  __shared__ int32_t s_base[MP_NV + 1];
  __shared__ int32_t s_vtx[MP_NV + 1];
  __shared__ SegItem<P> s_seg[NT];

  const Ctl *c = d.ctl;
  const int A = c->n_active[p];
  if (A == 0) return;
  const bool all = c->all_active[p] != 0;
  const int32_t *q = d.q[p];
  const int32_t *off = all ? d.csc_off : d.goff;
  const int32_t *part = all ? d.apart : d.gpart;
  const int N = c->gather_items + A;
  const int ntiles = (N + MP_NV - 1) / MP_NV;
  const int tid = threadIdx.x;

  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int d0 = t * MP_NV;
    const int d1 = min(d0 + MP_NV, N);
    const int s0 = part[t];
    const int s1 = part[t + 1];
    const int nseg = min(s1, A - 1) - s0 + 1;
    for (int k = tid; k < nseg; k += NT) {
      int s = s0 + k;
      int v = all ? s : q[s];
      int o = off[s];
      s_vtx[k] = v;
      s_base[k] = d.csc_off[v] - o;
      s_end[k] = off[s + 1] + s;
    }
    __syncthreads();
    const bool first_fresh = (off[s0] + s0 == d0);

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
          d.gtmp[s0 + kk[it]] = acc.finish();
        }
        ++nend;
        acc = Acc<P>::empty();
      }
    }
    // block-wide segmented scan of (end seen, partial)
    SegItem<P> me;
    me.a = acc;
    me.flag = nend > 0;
    for (int o = 1; o < NT; o <<= 1) {
      s_seg[tid] = me;
      __syncthreads();
      if (tid >= o) me = seg_combine(s_seg[tid - o], me);
      __syncthreads();
    }
    s_seg[tid] = me;  // inclusive
    __syncthreads();
    SegItem<P> carry;
    if (tid > 0) {
      carry = s_seg[tid - 1];
    } else {
      carry.a = Acc<P>::empty();
      carry.flag = 0;
    }
    if (nend > 0) {
      Acc<P> first = carry.a;
      first.merge(head);
      if (first_k > 0 || first_fresh) d.gtmp[s0 + first_k] = first.finish();
      if (!carry.flag) d.part_first[t] = first;
    }
    if (tid == NT - 1) {
      if (me.flag)
        d.part_last[t] = me.a;
      else
        d.part_first[t] = me.a;
    }
    __syncthreads();
  }
}

// Gather result of active slot i: direct, or folded from tile partials
// when its segment crosses tiles.
template <class P>
__device__ inline typename P::GatherResult gather_result(const Dev<P> &d,
                                                         const int32_t *off,
                                                         int i) {
  const int st = off[i] + i;
  const int en = off[i + 1] + i;
  const int t0 = st / MP_NV, t1 = en / MP_NV;
  if (t0 == t1) return d.gtmp[i];
  Acc<P> a = (st > t0 * MP_NV) ? d.part_last[t0] : d.part_first[t0];
  for (int t = t0 + 1; t <= t1; ++t) a.merge(d.part_first[t]);
  return a.finish();
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

// Apply (kApply, gpugas_kernels.cuh:139-151) fused with the fold of
// crossing segments and pass 1 of the scatter scan.  Also resets the next
// parity's list size, which k_scatter / k_compact2 fill.
template <class P, bool HaveGather>
__global__ __launch_bounds__(NT) void k_apply(Dev<P> d, int p) {
  typedef typename P::GatherResult GR;
  typedef typename P::VertexData VD;
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
  const int32_t *off = all ? d.csc_off : d.goff;
  int b, e;
  scan_chunk(A, blockIdx.x, b, e);
  int sum = 0;
  double res = 0.0;
  for (int i = b + threadIdx.x; i < e; i += NT) {
    const int v = all ? i : q[i];
    GR g;
    if (HaveGather)
      g = gather_result(d, off, i);
    else
      g = P::gatherZero;
    bool ret;
    if constexpr (has_residual<P>::value) {
      VD before = d.vdata[v];
      ret = P::apply(d.vdata + v, g);
      res += P::residual(before, d.vdata[v]);
    } else {
      ret = P::apply(d.vdata + v, g);
    }
    if (!all) d.bm[v >> 5] = 0u;  // every set bit belongs to A
    int pd = ret ? d.csr_off[v + 1] - d.csr_off[v] : 0;
    d.pdeg[i] = pd;
    sum += pd;
  }
  sum = block_sum(sum, lds);
  if (threadIdx.x == 0) d.bsum[blockIdx.x] = sum;
  if constexpr (has_residual<P>::value) {
    res = block_sum(res, ldsd);
    if (threadIdx.x == 0) d.rpart[blockIdx.x] = res;
  }
}

// Pass 2 of the scatter scan and the scatter tile partition; picks bitmap
// or queue mode for the next frontier from the number of edges to scatter,
// and counts the superstep (all applies of this superstep are done).
template <class P>
__global__ __launch_bounds__(NT) void k_sscan2(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  Ctl *c = d.ctl;
  const int A = c->n_active[p];
  if (A == 0) return;
  int b, e, pre, total;
  scan_chunk(A, blockIdx.x, b, e);
  chunk_prefix(d.bsum, blockIdx.x, pre, total, lds);
  int carry = pre;
  for (int base = b; base < e; base += NT) {
    int i = base + threadIdx.x;
    int x = (i < e) ? d.pdeg[i] : 0;
    int tot;
    int ex = block_excl_scan(x, tot, lds);
    if (i < e) {
      d.soff[i] = carry + ex;
      mark_tiles(d.spart, i, carry + ex, x);
    }
    carry += tot;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    d.soff[A] = total;
    d.spart[(total + A + MP_NV - 1) / MP_NV] = A;
    c->scatter_items = total;
    c->next_dense = total >= c->dense_min ? 1 : 0;
    c->stat_scattered += total;
    c->stat_active += A;
    c->stat_supersteps += 1;
    c->iteration += 1;
  }
}

// Scatter/activate (north-star (d); replaces IntervalGather into
// ActivateOutputIterator / ActivateOutputIteratorSmallSize,
// gpugas.h:545-597, 693-730).  Bitmap mode: fire-and-forget atomicOr.
// Queue mode: the lane whose atomicOr flips a bit appends the vertex; a
// tile collects its new vertices in LDS (ballot + mbcnt per wave) and
// reserves space in the next list with a single atomicAdd.
template <class P>
__global__ __launch_bounds__(NT) void k_scatter(Dev<P> d, int p) {
  __shared__ int32_t s_end[MP_NV + 1];
  __shared__ int32_t s_base[MP_NV + 1];
  __shared__ int32_t s_fresh[MP_NV];
  __shared__ int s_cnt, s_gbase;
  Ctl *c = d.ctl;
  const int A = c->n_active[p];
  if (A == 0) return;
  const int ES = c->scatter_items;
  if (ES == 0) return;
  const bool all = c->all_active[p] != 0;
  const bool dense = c->next_dense != 0;
  const int32_t *q = d.q[p];
  int32_t *qn = d.q[p ^ 1];
  int32_t *nn = &c->n_active[p ^ 1];
  const int32_t *off = d.soff;
  const int N = ES + A;
  const int ntiles = (N + MP_NV - 1) / MP_NV;
  const int tid = threadIdx.x;
  const int lane = lane_id();

  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int d0 = t * MP_NV;
    const int d1 = min(d0 + MP_NV, N);
    const int s0 = d.spart[t];
    const int s1 = d.spart[t + 1];
    const int nseg = min(s1, A - 1) - s0 + 1;
    for (int k = tid; k < nseg; k += NT) {
      int s = s0 + k;
      int v = all ? s : q[s];
      s_base[k] = d.csr_off[v] - off[s];
      s_end[k] = off[s + 1] + s;
    }
    __syncthreads();
    const int p0 = d0 + tid * MP_VT;
    int k = (p0 < d1) ? lds_lower_bound(s_end, nseg, p0) : 0;
    int eidx[MP_VT];
    bool isedge[MP_VT];
#pragma unroll
    for (int it = 0; it < MP_VT; ++it) {
      const int pos = p0 + it;
      isedge[it] = false;
      eidx[it] = 0;
      if (pos < d1) {
        if (s_end[k] == pos) {
          ++k;
        } else {
          isedge[it] = true;
          eidx[it] = s_base[k] + (pos - s0 - k);
        }
      }
    }
    int w[MP_VT];
#pragma unroll
    for (int it = 0; it < MP_VT; ++it)
      w[it] = isedge[it] ? d.csr_dst[eidx[it]] : 0;
    if (dense) {
#pragma unroll
      for (int it = 0; it < MP_VT; ++it)
        if (isedge[it]) atomicOr(&d.bm[w[it] >> 5], 1u << (w[it] & 31));
    } else {
      // stage this tile's newly activated vertices in LDS, then reserve
      // their slots in the next list with one atomicAdd per tile
      if (tid == 0) s_cnt = 0;
      __syncthreads();
#pragma unroll
      for (int it = 0; it < MP_VT; ++it) {
        bool fresh = false;
        if (isedge[it]) {
          uint32_t bit = 1u << (w[it] & 31);
          uint32_t old = atomicOr(&d.bm[w[it] >> 5], bit);
          fresh = (old & bit) == 0u;
        }
        uint64_t m = __ballot(fresh);
        if (m) {
          int leader = __ffsll((unsigned long long)m) - 1;
          int base = 0;
          if (lane == leader) base = atomicAdd(&s_cnt, __popcll(m));
          base = __shfl(base, leader, WAVE);
          if (fresh) {
            int rank = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
            s_fresh[base + rank] = w[it];
          }
        }
      }
      __syncthreads();
      const int cnt = s_cnt;
      if (tid == 0 && cnt) s_gbase = atomicAdd(nn, cnt);
      __syncthreads();
      for (int i = tid; i < cnt; i += NT) qn[s_gbase + i] = s_fresh[i];
    }
    __syncthreads();
  }
}

// Frontier compaction, bitmap mode (north-star (e); replaces
// scatter_if_inputloc_twophase, scatter_if_mgpu.h:130-157): pass 1 counts
// set bits per chunk of words, pass 2 writes the ids of every chunk in
// ascending order at its prefix and clears the words.
__device__ inline bool compact_dense(const Ctl *c, int p) {
  return c->n_active[p] > 0 && c->next_dense != 0 && c->scatter_items > 0;
}

template <class P>
__global__ __launch_bounds__(NT) void k_compact1(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  if (!compact_dense(d.ctl, p)) return;
  int b, e;
  scan_chunk(d.nwords, blockIdx.x, b, e);
  int cnt = 0;
  for (int i = b + threadIdx.x; i < e; i += NT) cnt += __popc(d.bm[i]);
  cnt = block_sum(cnt, lds);
  if (threadIdx.x == 0) d.bsum[blockIdx.x] = cnt;
}

template <class P>
__global__ __launch_bounds__(NT) void k_compact2(Dev<P> d, int p) {
  __shared__ int lds[NT / WAVE];
  Ctl *c = d.ctl;
  if (!compact_dense(c, p)) return;
  int32_t *qn = d.q[p ^ 1];
  int b, e, pre, total;
  scan_chunk(d.nwords, blockIdx.x, b, e);
  chunk_prefix(d.bsum, blockIdx.x, pre, total, lds);
  int carry = pre;
  for (int base = b; base < e; base += NT) {
    int wi = base + threadIdx.x;
    uint32_t word = (wi < e) ? d.bm[wi] : 0u;
    int tot;
    int pos = carry + block_excl_scan(__popc(word), tot, lds);
    if (word) {
      d.bm[wi] = 0u;
      while (word) {
        int bit = __ffs(word) - 1;
        qn[pos++] = wi * 32 + bit;
        word &= word - 1;
      }
    }
    carry += tot;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) c->n_active[p ^ 1] = total;
}

// Superstep counter write (host-side setIteration).
template <class P>
__global__ void k_set_iteration(Dev<P> d, int it) {
  if (threadIdx.x == 0) d.ctl->iteration = it;
}

}  // namespace gas
