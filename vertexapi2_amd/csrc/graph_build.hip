// Device-side CSR / CSC construction for large graphs (replaces the host
// sort of util.h:48-121 / gpugas.h:212-322): one stable LSD radix sort of
// (key vertex, edge index) pairs per layout (rocPRIM), the neighbour ids and
// weights gathered through the sorted edge indices, and the offsets read
// off the boundaries between equal keys.  A stable sort keeps equal keys in
// edge-list order, so the layout is the one the host counting sort builds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>  // rocprim's iterators use memset without including it

#include <rocprim/rocprim.hpp>

#include "gas.h"
#include "gas_internal.hpp"

namespace gas {
namespace {

constexpr int kThreads = 256;

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

unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(8192, (n + kThreads - 1) / kThreads));
}

int hip_fail(hipError_t e, const char *what) {
  return set_error(GAS_ERR_HIP, "device graph build: %s: %s", what, hipGetErrorString(e));
}

}  // namespace

int device_build_layout(int32_t nV, int64_t nE, const int32_t *d_keys,
                        const int32_t *d_vals, const int32_t *d_w,
                        int32_t *d_off, int32_t *d_out_vals, int32_t *d_eid,
                        int32_t *d_out_w, uint32_t *d_sorted_keys, void *d_temp,
                        size_t temp_bytes) {
  unsigned end_bit = 1;
  while (end_bit < 32 && ((int64_t)1 << end_bit) < (int64_t)nV) ++end_bit;
  hipError_t e = hipSuccess;
  if (nE > 0) {
    size_t need = temp_bytes;
    e = rocprim::radix_sort_pairs(d_temp, need,
                                  reinterpret_cast<const uint32_t *>(d_keys),
                                  d_sorted_keys, rocprim::counting_iterator<int32_t>(0),
                                  d_eid, (size_t)nE, 0, end_bit, hipStream_t(0));
    if (e != hipSuccess) return hip_fail(e, "radix sort");
    k_gather_by_index<<<grid_for(nE), kThreads>>>(d_vals, d_eid, d_out_vals, nE);
    if (d_w) k_gather_by_index<<<grid_for(nE), kThreads>>>(d_w, d_eid, d_out_w, nE);
  }
  k_offsets_from_keys<<<grid_for(nE + 1), kThreads>>>(d_sorted_keys, nE, nV, d_off);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "launch");
  return GAS_OK;
}

size_t device_build_temp_bytes(int64_t nE) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const uint32_t *>(nullptr),
                            static_cast<uint32_t *>(nullptr),
                            rocprim::counting_iterator<int32_t>(0),
                            static_cast<int32_t *>(nullptr), (size_t)std::max<int64_t>(nE, 1),
                            0, 32, hipStream_t(0));
  return bytes;
}

}  // namespace gas
