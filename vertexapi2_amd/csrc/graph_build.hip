// Device-side CSR / CSC construction for large graphs (replaces the host
// sort of util.h:48-121 / gpugas.h:212-322): a stable LSD radix sort of
// (key vertex, edge index) pairs per layout, the neighbour ids and weights
// gathered through the sorted edge indices, and the offsets read off the
// boundaries between equal keys.  A stable sort keeps equal keys in
// edge-list order, so the layout is the one the host counting sort builds.
//
// The sort: 8-bit digits, ceil(bits(V - 1) / 8) passes (three for up to
// 2^24 vertices).  Each pass cuts the items into tiles of SORT_TILE
// consecutive items (one workgroup each) and runs
//   k_digit_hist   per-tile digit counts, stored digit-major
//                  (hist[d * ntiles + t]);
//   k_digit_scan   one workgroup per digit: exclusive scan of its counts
//                  over the tiles, and the digit's total;
//   k_digit_base   exclusive scan of the 256 totals;
//   k_digit_scatter  each tile again, in rounds of SORT_NT items in order:
//                  a wave finds the lanes holding its item's digit with
//                  eight ballots, ranks the item among them, adds the counts
//                  of earlier waves of the round and earlier rounds of the
//                  tile (LDS) and the digit's base and tile offset, and
//                  writes key and value there.
// Items keep their order within a digit in every pass, which is what makes
// the LSD sort stable.  Every launch goes through gasrt_launch on the
// caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <tuple>
#include <utility>

#include "gas.h"
#include "gas_internal.hpp"

namespace gas {
namespace {

constexpr int kThreads = 256;
constexpr int SORT_NT = 256;            // threads per sort workgroup (4 waves)
constexpr int SORT_IPT = 16;            // items per thread per tile
constexpr int SORT_TILE = SORT_NT * SORT_IPT;
constexpr int RADIX = 256;              // 8-bit digits
constexpr int SORT_WAVES = SORT_NT / 64;
static_assert(SORT_NT == RADIX, "one thread per digit in the tile kernels");

template <class... KP, class... A>
int launch(void (*k)(KP...), unsigned grid, unsigned block, gasrt_stream_t s, A &&...a) {
  std::tuple<std::decay_t<KP>...> t(std::forward<A>(a)...);
  void *argv[sizeof...(KP) + 1];
  std::apply([&](auto &...x) {
    int i = 0;
    ((argv[i++] = (void *)&x), ...);
  }, t);
  return gasrt_launch((const void *)k, grid, block, argv, 0, s);
}

__global__ __launch_bounds__(kThreads) void k_gather_by_index(
    const int32_t *__restrict__ in, const int32_t *__restrict__ idx,
    int32_t *__restrict__ out, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * kThreads)
    out[k] = in[idx[k]];
}

// off[v] = first slot whose key is >= v, for v in [0, nV]: slot k fills the
// vertices after the previous slot's key up to its own (slot n stands for
// key nV), so every vertex, empty ones included, is written once.
__global__ __launch_bounds__(kThreads) void k_offsets_from_keys(
    const uint32_t *__restrict__ keys, int64_t n, int32_t nV,
    int32_t *__restrict__ off) {
  for (int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x; k <= n;
       k += (int64_t)gridDim.x * kThreads) {
    const int64_t prev = k == 0 ? -1 : (int64_t)keys[k - 1];
    const int64_t cur = k == n ? (int64_t)nV : (int64_t)keys[k];
    for (int64_t v = prev + 1; v <= cur; ++v) off[v] = (int32_t)k;
  }
}

// Digit counts of tile blockIdx.x (items [t * SORT_TILE, ...) of n).
__global__ __launch_bounds__(SORT_NT) void k_digit_hist(
    const uint32_t *__restrict__ keys, int64_t n, int shift, int32_t ntiles,
    int32_t *__restrict__ hist) {
  __shared__ int32_t cnt[RADIX];
  cnt[threadIdx.x] = 0;  // SORT_NT == RADIX
  __syncthreads();
  const int64_t t0 = (int64_t)blockIdx.x * SORT_TILE;
#pragma unroll 4
  for (int r = 0; r < SORT_IPT; ++r) {
    const int64_t i = t0 + (int64_t)r * SORT_NT + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (RADIX - 1)], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of digit blockIdx.x's counts over the tiles, in place;
// its total into totals[digit].
__global__ __launch_bounds__(SORT_NT) void k_digit_scan(int32_t *__restrict__ hist,
                                                        int32_t ntiles,
                                                        int32_t *__restrict__ totals) {
  __shared__ int32_t wsum[SORT_WAVES];
  int32_t *h = hist + (int64_t)blockIdx.x * ntiles;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t carry = 0;
  for (int32_t base = 0; base < ntiles; base += SORT_NT) {
    const int32_t i = base + threadIdx.x;
    const int32_t x = i < ntiles ? h[i] : 0;
    int32_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t y = __shfl_up(inc, o, 64);
      if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int32_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < SORT_WAVES; ++k) {
      pre += k < w ? wsum[k] : 0;
      tot += wsum[k];
    }
    if (i < ntiles) h[i] = carry + pre + inc - x;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// base[d] = sum of totals[d'] for d' < d (one workgroup of RADIX threads).
__global__ __launch_bounds__(SORT_NT) void k_digit_base(const int32_t *__restrict__ totals,
                                                        int32_t *__restrict__ base) {
  __shared__ int32_t wsum[SORT_WAVES];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int32_t x = totals[threadIdx.x];
  int32_t inc = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int32_t pre = 0;
#pragma unroll
  for (int k = 0; k < SORT_WAVES; ++k) pre += k < w ? wsum[k] : 0;
  base[threadIdx.x] = pre + inc - x;
}

// Stable scatter of tile blockIdx.x.  vals_in == nullptr: the value of item
// i is i (the first pass sorts edge indices).
__global__ __launch_bounds__(SORT_NT) void k_digit_scatter(
    const uint32_t *__restrict__ keys_in, const int32_t *__restrict__ vals_in, int64_t n,
    int shift, int32_t ntiles, const int32_t *__restrict__ hist,
    const int32_t *__restrict__ base, uint32_t *__restrict__ keys_out,
    int32_t *__restrict__ vals_out) {
  __shared__ int32_t run[RADIX];                // items of each digit in earlier rounds
  __shared__ int32_t wcnt[SORT_WAVES][RADIX];   // this round's items per wave and digit
  __shared__ int32_t dst0[RADIX];               // where the tile's items of a digit start
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  run[tid] = 0;
#pragma unroll
  for (int k = 0; k < SORT_WAVES; ++k) wcnt[k][tid] = 0;
  dst0[tid] = base[tid] + hist[(int64_t)tid * ntiles + blockIdx.x];
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t t0 = (int64_t)blockIdx.x * SORT_TILE;
  for (int r = 0; r < SORT_IPT; ++r) {
    const int64_t i = t0 + (int64_t)r * SORT_NT + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? keys_in[i] : 0u;
    const int32_t val = valid ? (vals_in ? vals_in[i] : (int32_t)i) : 0;
    const int digit = (int)((key >> shift) & (RADIX - 1));
    // lanes of this wave holding a valid item of the same digit
    uint64_t same = __ballot(valid);
    if (!valid) same = ~same;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bal = __ballot((digit >> b) & 1);
      same &= ((digit >> b) & 1) ? bal : ~bal;
    }
    const int below = __popcll(same & lt);
    if (valid && below == 0) wcnt[w][digit] = __popcll(same);
    __syncthreads();
    if (valid) {
      int32_t pos = dst0[digit] + run[digit] + below;
      for (int k = 0; k < w; ++k) pos += wcnt[k][digit];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    int32_t add = 0;
#pragma unroll
    for (int k = 0; k < SORT_WAVES; ++k) {
      add += wcnt[k][tid];
      wcnt[k][tid] = 0;
    }
    run[tid] += add;
    __syncthreads();
  }
}

// Row v of the out-neighbour table: its CSR out-neighbours (weights) then
// -1 (0) up to kEllWidth.
__global__ __launch_bounds__(kThreads) void k_build_ell(
    int32_t nV, const int32_t *__restrict__ off, const int32_t *__restrict__ dst,
    const int32_t *__restrict__ w, int32_t *__restrict__ ell_dst,
    int32_t *__restrict__ ell_w) {
  for (int64_t v = (int64_t)blockIdx.x * kThreads + threadIdx.x; v < nV;
       v += (int64_t)gridDim.x * kThreads) {
    const int32_t o = off[v], deg = off[v + 1] - o;
#pragma unroll
    for (int k = 0; k < kEllWidth; ++k) {
      ell_dst[v * kEllWidth + k] = k < deg ? dst[o + k] : -1;
      if (w) ell_w[v * kEllWidth + k] = k < deg ? w[o + k] : 0;
    }
  }
}

unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(8192, (n + kThreads - 1) / kThreads));
}

int64_t tiles_for(int64_t n) { return (n + SORT_TILE - 1) / SORT_TILE; }

// Scratch of one layout build, carved from d_temp.
struct SortScratch {
  uint32_t *keys_alt;
  int32_t *vals_alt;
  int32_t *hist, *totals, *base;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

SortScratch carve(void *temp, int64_t nE) {
  SortScratch s;
  char *p = static_cast<char *>(temp);
  const size_t eb = align256((size_t)std::max<int64_t>(nE, 1) * 4);
  s.keys_alt = reinterpret_cast<uint32_t *>(p);
  p += eb;
  s.vals_alt = reinterpret_cast<int32_t *>(p);
  p += eb;
  s.hist = reinterpret_cast<int32_t *>(p);
  p += align256((size_t)RADIX * std::max<int64_t>(tiles_for(nE), 1) * 4);
  s.totals = reinterpret_cast<int32_t *>(p);
  p += align256(RADIX * 4);
  s.base = reinterpret_cast<int32_t *>(p);
  return s;
}

}  // namespace

int device_build_ell(int32_t nV, const int32_t *d_off, const int32_t *d_dst,
                     const int32_t *d_w, int32_t *ell_dst, int32_t *ell_w,
                     gasrt_stream_t stream) {
  return launch(k_build_ell, grid_for(nV), kThreads, stream, nV, d_off, d_dst, d_w,
                ell_dst, ell_w);
}

size_t device_build_temp_bytes(int64_t nE) {
  const size_t eb = align256((size_t)std::max<int64_t>(nE, 1) * 4);
  return 2 * eb + align256((size_t)RADIX * std::max<int64_t>(tiles_for(nE), 1) * 4) +
         2 * align256(RADIX * 4);
}

int device_build_layout(int32_t nV, int64_t nE, const int32_t *d_keys,
                        const int32_t *d_vals, const int32_t *d_w,
                        int32_t *d_off, int32_t *d_out_vals, int32_t *d_eid,
                        int32_t *d_out_w, uint32_t *d_sorted_keys, void *d_temp,
                        size_t temp_bytes, gasrt_stream_t stream) {
  if (temp_bytes < device_build_temp_bytes(nE))
    return set_error(GAS_ERR_ARG, "device graph build: scratch too small");
  if (nE > INT32_MAX)
    return set_error(GAS_ERR_ARG, "device graph build: more than 2^31 - 1 edges");
  int bits = 1;
  while (bits < 32 && ((int64_t)1 << bits) < (int64_t)nV) ++bits;
  const int passes = (bits + 7) / 8;
  if (nE > 0) {
    const SortScratch s = carve(d_temp, nE);
    const int32_t ntiles = (int32_t)tiles_for(nE);
    const uint32_t *kin = reinterpret_cast<const uint32_t *>(d_keys);
    const int32_t *vin = nullptr;
    for (int p = 0; p < passes; ++p) {
      // the last pass writes the caller's outputs
      const bool to_out = (p % 2) == ((passes - 1) % 2);
      uint32_t *kout = to_out ? d_sorted_keys : s.keys_alt;
      int32_t *vout = to_out ? d_eid : s.vals_alt;
      const int shift = 8 * p;
      GAS_TRY(launch(k_digit_hist, (unsigned)ntiles, SORT_NT, stream, kin, nE, shift, ntiles,
                     s.hist));
      GAS_TRY(launch(k_digit_scan, RADIX, SORT_NT, stream, s.hist, ntiles, s.totals));
      GAS_TRY(launch(k_digit_base, 1, SORT_NT, stream, s.totals, s.base));
      GAS_TRY(launch(k_digit_scatter, (unsigned)ntiles, SORT_NT, stream, kin, vin, nE, shift,
                     ntiles, s.hist, s.base, kout, vout));
      kin = kout;
      vin = vout;
    }
    GAS_TRY(launch(k_gather_by_index, grid_for(nE), kThreads, stream, d_vals,
                   (const int32_t *)d_eid, d_out_vals, nE));
    if (d_w)
      GAS_TRY(launch(k_gather_by_index, grid_for(nE), kThreads, stream, d_w,
                     (const int32_t *)d_eid, d_out_w, nE));
  }
  GAS_TRY(launch(k_offsets_from_keys, grid_for(nE + 1), kThreads, stream,
                 (const uint32_t *)d_sorted_keys, nE, nV, d_off));
  return GAS_OK;
}

}  // namespace gas
