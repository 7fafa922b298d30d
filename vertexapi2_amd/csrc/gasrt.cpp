// Thin HIP runtime layer behind the gasrt_* C ABI.
//
// Stands in for the helpers the reference takes from its (absent)
// cuda_util.h: gpuAlloc / gpuFree / copyToGPU / copyToHost / CHECK /
// SYNC_CHECK (call sites gpugas.h:177-195, 227-318).  Every function returns
// a status and records the failure text per thread instead of aborting.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "gas.h"
#include "gas_internal.hpp"

namespace gas {

static thread_local std::string t_last_error;
static thread_local int t_engine_set = 0;

int &engine_set() { return t_engine_set; }

int set_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_last_error = buf;
  return code;
}

void clear_error() { t_last_error.clear(); }

int hip_status(int hip_err, const char *what) {
  hipError_t e = static_cast<hipError_t>(hip_err);
  if (e == hipSuccess) return GAS_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return set_error(GAS_ERR_NOMEM, "%s: %s", what, hipGetErrorString(e));
  return set_error(GAS_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace gas

using gas::hip_status;
using gas::set_error;

extern "C" {

const char *gas_last_error(void) { return gas::t_last_error.c_str(); }

int gasrt_device_count(int *count) {
  if (!count) return set_error(GAS_ERR_ARG, "gasrt_device_count: null");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return hip_status(e, "hipGetDeviceCount");
  }
  *count = n;
  return GAS_OK;
}

int gasrt_set_device(int device) {
  return hip_status(hipSetDevice(device), "hipSetDevice");
}

int gasrt_device_info(int device, int *cu_count, int *lds_bytes_per_block,
                      size_t *hbm_bytes, char *name, size_t name_len) {
  hipDeviceProp_t p;
  int rc = hip_status(hipGetDeviceProperties(&p, device),
                      "hipGetDeviceProperties");
  if (rc) return rc;
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (lds_bytes_per_block) *lds_bytes_per_block = (int)p.sharedMemPerBlock;
  if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  return GAS_OK;
}

int gasrt_malloc(void **ptr, size_t bytes) {
  if (!ptr) return set_error(GAS_ERR_ARG, "gasrt_malloc: null");
  *ptr = nullptr;
  if (bytes == 0) return GAS_OK;
  return hip_status(hipMalloc(ptr, bytes), "hipMalloc");
}

int gasrt_free(void *ptr) {
  if (!ptr) return GAS_OK;
  return hip_status(hipFree(ptr), "hipFree");
}

int gasrt_host_alloc(void **ptr, size_t bytes) {
  if (!ptr) return set_error(GAS_ERR_ARG, "gasrt_host_alloc: null");
  *ptr = nullptr;
  if (bytes == 0) return GAS_OK;
  return hip_status(hipHostMalloc(ptr, bytes, hipHostMallocDefault),
                    "hipHostMalloc");
}

int gasrt_host_alloc_mapped(void **host_ptr, void **dev_ptr, size_t bytes) {
  if (!host_ptr || !dev_ptr || bytes == 0)
    return set_error(GAS_ERR_ARG, "gasrt_host_alloc_mapped: bad arguments");
  *host_ptr = *dev_ptr = nullptr;
  int rc = hip_status(
      hipHostMalloc(host_ptr, bytes, hipHostMallocMapped | hipHostMallocCoherent),
      "hipHostMalloc(mapped, coherent)");
  if (rc != GAS_OK) return rc;
  rc = hip_status(hipHostGetDevicePointer(dev_ptr, *host_ptr, 0),
                  "hipHostGetDevicePointer");
  if (rc != GAS_OK) {
    hipHostFree(*host_ptr);
    *host_ptr = nullptr;
  }
  return rc;
}

int gasrt_host_free(void *ptr) {
  if (!ptr) return GAS_OK;
  return hip_status(hipHostFree(ptr), "hipHostFree");
}

static hipStream_t S(gasrt_stream_t s) { return static_cast<hipStream_t>(s); }

int gasrt_memcpy_h2d(void *dst, const void *src, size_t bytes,
                     gasrt_stream_t stream) {
  if (!bytes) return GAS_OK;
  int rc = hip_status(
      hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)),
      "hipMemcpyAsync(H2D)");
  if (rc) return rc;
  return hip_status(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
}

int gasrt_memcpy_d2h(void *dst, const void *src, size_t bytes,
                     gasrt_stream_t stream) {
  if (!bytes) return GAS_OK;
  int rc = hip_status(
      hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)),
      "hipMemcpyAsync(D2H)");
  if (rc) return rc;
  return hip_status(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
}

int gasrt_memcpy_d2d(void *dst, const void *src, size_t bytes,
                     gasrt_stream_t stream) {
  if (!bytes) return GAS_OK;
  return hip_status(
      hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)),
      "hipMemcpyAsync(D2D)");
}

int gasrt_memcpy_d2h_async(void *dst, const void *src, size_t bytes,
                           gasrt_stream_t stream) {
  if (!bytes) return GAS_OK;
  return hip_status(
      hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)),
      "hipMemcpyAsync(D2H)");
}

int gasrt_memset(void *dst, int value, size_t bytes, gasrt_stream_t stream) {
  if (!bytes) return GAS_OK;
  return hip_status(hipMemsetAsync(dst, value, bytes, S(stream)),
                    "hipMemsetAsync");
}

int gasrt_stream_create(gasrt_stream_t *stream) {
  if (!stream) return set_error(GAS_ERR_ARG, "gasrt_stream_create: null");
  hipStream_t s = nullptr;
  int rc = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking),
                      "hipStreamCreateWithFlags");
  *stream = s;
  return rc;
}

int gasrt_stream_destroy(gasrt_stream_t stream) {
  if (!stream) return GAS_OK;
  return hip_status(hipStreamDestroy(S(stream)), "hipStreamDestroy");
}

int gasrt_stream_synchronize(gasrt_stream_t stream) {
  return hip_status(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
}

int gasrt_stream_query(gasrt_stream_t stream, int *idle) {
  if (!idle) return set_error(GAS_ERR_ARG, "gasrt_stream_query: null");
  hipError_t e = hipStreamQuery(S(stream));
  if (e == hipErrorNotReady) {
    *idle = 0;
    return GAS_OK;
  }
  *idle = 1;
  return hip_status(e, "hipStreamQuery");
}

int gasrt_event_create(gasrt_event_t *event) {
  if (!event) return set_error(GAS_ERR_ARG, "gasrt_event_create: null");
  hipEvent_t e = nullptr;
  int rc = hip_status(hipEventCreate(&e), "hipEventCreate");
  *event = e;
  return rc;
}

int gasrt_event_destroy(gasrt_event_t event) {
  if (!event) return GAS_OK;
  return hip_status(hipEventDestroy(static_cast<hipEvent_t>(event)),
                    "hipEventDestroy");
}

int gasrt_event_record(gasrt_event_t event, gasrt_stream_t stream) {
  return hip_status(hipEventRecord(static_cast<hipEvent_t>(event), S(stream)),
                    "hipEventRecord");
}

int gasrt_event_synchronize(gasrt_event_t event) {
  return hip_status(hipEventSynchronize(static_cast<hipEvent_t>(event)),
                    "hipEventSynchronize");
}

int gasrt_event_query(gasrt_event_t event, int *done) {
  hipError_t e = hipEventQuery(static_cast<hipEvent_t>(event));
  if (e == hipErrorNotReady) {
    *done = 0;
    return GAS_OK;
  }
  *done = 1;
  return hip_status(e, "hipEventQuery");
}

int gasrt_event_elapsed_ms(gasrt_event_t start, gasrt_event_t stop,
                           float *ms) {
  return hip_status(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                        static_cast<hipEvent_t>(stop)),
                    "hipEventElapsedTime");
}

int gasrt_launch(const void *func, unsigned grid_x, unsigned block_x,
                 void **args, size_t shared_bytes, gasrt_stream_t stream) {
  if (grid_x == 0) return GAS_OK;
  return hip_status(hipLaunchKernel(func, dim3(grid_x), dim3(block_x), args,
                                    shared_bytes, S(stream)),
                    "hipLaunchKernel");
}

int gasrt_capture_begin(gasrt_stream_t stream) {
  return hip_status(
      hipStreamBeginCapture(S(stream), hipStreamCaptureModeThreadLocal),
      "hipStreamBeginCapture");
}

int gasrt_capture_end(gasrt_stream_t stream, gasrt_graph_exec_t *exec) {
  hipGraph_t g = nullptr;
  int rc = hip_status(hipStreamEndCapture(S(stream), &g),
                      "hipStreamEndCapture");
  if (rc) return rc;
  hipGraphExec_t ge = nullptr;
  rc = hip_status(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0),
                  "hipGraphInstantiate");
  hipGraphDestroy(g);
  if (rc) return rc;
  *exec = ge;
  return GAS_OK;
}

int gasrt_graph_launch(gasrt_graph_exec_t exec, gasrt_stream_t stream) {
  return hip_status(
      hipGraphLaunch(static_cast<hipGraphExec_t>(exec), S(stream)),
      "hipGraphLaunch");
}

int gasrt_graph_destroy(gasrt_graph_exec_t exec) {
  if (!exec) return GAS_OK;
  return hip_status(hipGraphExecDestroy(static_cast<hipGraphExec_t>(exec)),
                    "hipGraphExecDestroy");
}

int gasrt_mem_info(size_t *free_bytes, size_t *total_bytes) {
  return hip_status(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}

}  // extern "C"
