// Runtime part of the C-ABI: device selection, library stream, budgeted device
// memory, copies, fills and events.
#include <cstdarg>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "fap_internal.h"

namespace fap_hip {

namespace {
thread_local std::string g_last_error;

std::mutex g_mu;
bool g_init = false;
int g_device = -1;
hipStream_t g_stream = nullptr;
hipStream_t g_aux[2] = {nullptr, nullptr};
int g_cus = 0;  // compute units of the device
hipEvent_t g_ws_fence = nullptr;  // end of the last call that used the pipeline workspaces
bool g_ws_fence_set = false;
int64_t g_budget = 0;  // 0 = unlimited
int64_t g_in_use = 0;
int64_t g_high = 0;
struct Block {
  int64_t bytes;   // size of the device allocation
  int64_t booked;  // bytes the caller asked for (what in_use and the budget count)
};
std::unordered_map<void *, Block> g_sizes;
// Freed blocks kept for reuse: hipMalloc / hipFree of multi-GB buffers while kernels
// run on other streams took 0.6-2.3 s each in the request server (0.4 ms on an idle
// device), so blocks up to kCacheBlockMax go back to this cache (at most kCacheMax in
// all) and a later request takes the smallest cached block of at least its size and at
// most twice it.  Larger blocks (a placed n x n result) are freed at once.
constexpr int64_t kCacheBlockMax = 32LL << 30;
constexpr int64_t kCacheMax = 64LL << 30;
std::multimap<int64_t, void *> g_cache;
int64_t g_cached = 0;

// Frees every cached block (with g_mu held by the caller).
void drop_cache_locked(std::vector<void *> &out) {
  for (auto &kv : g_cache) out.push_back(kv.second);
  g_cache.clear();
  g_cached = 0;
}
}  // namespace

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}
void clear_error() { g_last_error.clear(); }

hipStream_t lib_stream() { return g_stream; }
hipStream_t aux_stream(int i) { return g_aux[i & 1]; }
int device_cus() { return g_cus; }
int lib_device() { return g_device; }
bool initialized() { return g_init; }

int ws_fence_wait(hipStream_t st) {
  if (g_ws_fence_set) FAP_CHECK_HIP(hipStreamWaitEvent(st, g_ws_fence, 0));
  return FAP_OK;
}

int ws_fence_record(hipStream_t st) {
  FAP_CHECK_HIP(hipEventRecord(g_ws_fence, st));
  g_ws_fence_set = true;
  return FAP_OK;
}

int dev_alloc(void **p, int64_t bytes) {
  if (!p || bytes < 0) {
    set_error("dev_alloc: invalid arguments");
    return FAP_ERR_INVALID;
  }
  if (!g_init) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  *p = nullptr;
  if (bytes == 0) return FAP_OK;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_budget > 0 && g_in_use + bytes > g_budget) {
      set_error("HBM budget exceeded: %lld in use + %lld requested > %lld budget",
                (long long)g_in_use, (long long)bytes, (long long)g_budget);
      return FAP_ERR_BUDGET;
    }
    auto it = g_cache.lower_bound(bytes);
    if (it != g_cache.end() && it->first <= 2 * bytes) {
      *p = it->second;
      g_sizes[*p] = Block{it->first, bytes};
      g_cached -= it->first;
      g_cache.erase(it);
      g_in_use += bytes;
      if (g_in_use > g_high) g_high = g_in_use;
      return FAP_OK;
    }
  }
  hipError_t e = hipMalloc(p, (size_t)bytes);
  if (e != hipSuccess) {  // out of memory: give the cached blocks back and try again
    (void)hipGetLastError();
    std::vector<void *> drop;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      drop_cache_locked(drop);
    }
    for (void *q : drop) (void)hipFree(q);
    e = drop.empty() ? e : hipMalloc(p, (size_t)bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    set_error("hipMalloc(%lld) failed: %s", (long long)bytes, hipGetErrorString(e));
    return FAP_ERR_NOMEM;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_sizes[*p] = Block{bytes, bytes};
  g_in_use += bytes;
  if (g_in_use > g_high) g_high = g_in_use;
  return FAP_OK;
}

int dev_free(void *p) {
  if (!p) return FAP_OK;
  std::vector<void *> drop;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_sizes.find(p);
    if (it == g_sizes.end()) {
      set_error("fap_hip_free: pointer %p was not allocated by fap_hip_malloc", p);
      return FAP_ERR_INVALID;
    }
    const Block b = it->second;
    g_in_use -= b.booked;
    g_sizes.erase(it);
    if (b.bytes > kCacheBlockMax) {
      drop.push_back(p);
    } else {
      g_cache.insert({b.bytes, p});
      g_cached += b.bytes;
      while (g_cached > kCacheMax) {  // the largest cached blocks go first
        auto last = std::prev(g_cache.end());
        g_cached -= last->first;
        drop.push_back(last->second);
        g_cache.erase(last);
      }
    }
  }
  for (void *q : drop) FAP_CHECK_HIP(hipFree(q));
  return FAP_OK;
}

int Workspace::ensure(int64_t need) {
  if (need <= bytes) return FAP_OK;
  if (ptr) {
    // The old buffer may still be read by queued work on the library stream.
    FAP_CHECK_HIP(hipDeviceSynchronize());
    FAP_TRY(dev_free(ptr));
    ptr = nullptr;
    bytes = 0;
  }
  int64_t grown = need + need / 4;
  int s = dev_alloc(&ptr, grown);
  if (s != FAP_OK) {
    s = dev_alloc(&ptr, need);
    if (s != FAP_OK) return s;
    grown = need;
  }
  bytes = grown;
  return FAP_OK;
}

__global__ void fill_i32_kernel(int32_t *__restrict__ p, int32_t v, int64_t count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

// 16-byte stores; count4 = number of int4 elements.
__global__ void fill_i32x4_kernel(int4 *__restrict__ p, int32_t v, int64_t count4) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int4 val = make_int4(v, v, v, v);
  for (; i < count4; i += stride) p[i] = val;
}

__global__ void copy_i32x4_kernel(int4 *__restrict__ dst, const int4 *__restrict__ src,
                                  int64_t count4) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count4; i += stride) dst[i] = src[i];
}

__global__ void copy_i32_kernel(int32_t *__restrict__ dst, const int32_t *__restrict__ src,
                                int64_t count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) dst[i] = src[i];
}

}  // namespace fap_hip

using namespace fap_hip;

extern "C" {

int fap_hip_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_init) {
    if (device == g_device) return FAP_OK;
    set_error("fap_hip_init: already initialised on device %d", g_device);
    return FAP_ERR_STATE;
  }
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= device || device < 0) {
    (void)hipGetLastError();
    set_error("fap_hip_init: no HIP device %d (count %d, %s)", device, count,
              hipGetErrorString(e));
    return FAP_ERR_NODEV;
  }
  FAP_CHECK_HIP(hipSetDevice(device));
  FAP_CHECK_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  // Two more streams (default priority) for overlapping the stages of different parts.
  FAP_CHECK_HIP(hipStreamCreateWithFlags(&g_aux[0], hipStreamNonBlocking));
  FAP_CHECK_HIP(hipStreamCreateWithFlags(&g_aux[1], hipStreamNonBlocking));
  FAP_CHECK_HIP(hipEventCreateWithFlags(&g_ws_fence, hipEventDisableTiming));
  FAP_CHECK_HIP(hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, device));
  FAP_TRY(floyd_scratch_setup(device));
  g_device = device;
  g_init = true;
  return FAP_OK;
}

const char *fap_hip_last_error(void) { return g_last_error.c_str(); }

int fap_hip_device_info(int *cu_count, int64_t *lds_bytes, int64_t *hbm_bytes, char *name,
                        int name_len) {
  if (!g_init) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  hipDeviceProp_t prop;
  FAP_CHECK_HIP(hipGetDeviceProperties(&prop, g_device));
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (lds_bytes) *lds_bytes = (int64_t)prop.sharedMemPerBlock;
  if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
  if (name && name_len > 0) snprintf(name, name_len, "%s", prop.gcnArchName);
  return FAP_OK;
}

int fap_hip_set_budget(int64_t bytes) {
  if (bytes < 0) {
    set_error("fap_hip_set_budget: negative budget");
    return FAP_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_budget = bytes;
  return FAP_OK;
}

int fap_hip_mem_stats(int64_t *in_use, int64_t *high_water, int64_t *budget) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (in_use) *in_use = g_in_use;
  if (high_water) *high_water = g_high;
  if (budget) *budget = g_budget;
  return FAP_OK;
}

int fap_hip_malloc(void **d_ptr, int64_t bytes) { return dev_alloc(d_ptr, bytes); }
int fap_hip_free(void *d_ptr) { return dev_free(d_ptr); }

int fap_hip_memcpy_h2d(void *d_dst, const void *h_src, int64_t bytes, void *stream) {
  if (bytes == 0) return FAP_OK;
  FAP_REQUIRE(d_dst && h_src && bytes > 0, "fap_hip_memcpy_h2d: invalid arguments");
  hipStream_t s = resolve(stream);
  FAP_CHECK_HIP(hipMemcpyAsync(d_dst, h_src, (size_t)bytes, hipMemcpyHostToDevice, s));
  FAP_CHECK_HIP(hipStreamSynchronize(s));
  return FAP_OK;
}

int fap_hip_memcpy_d2h(void *h_dst, const void *d_src, int64_t bytes, void *stream) {
  if (bytes == 0) return FAP_OK;
  FAP_REQUIRE(h_dst && d_src && bytes > 0, "fap_hip_memcpy_d2h: invalid arguments");
  hipStream_t s = resolve(stream);
  FAP_CHECK_HIP(hipMemcpyAsync(h_dst, d_src, (size_t)bytes, hipMemcpyDeviceToHost, s));
  FAP_CHECK_HIP(hipStreamSynchronize(s));
  return FAP_OK;
}

int fap_hip_memcpy_d2d(void *d_dst, const void *d_src, int64_t bytes, void *stream) {
  if (bytes == 0) return FAP_OK;
  FAP_REQUIRE(d_dst && d_src && bytes > 0, "fap_hip_memcpy_d2d: invalid arguments");
  FAP_CHECK_HIP(hipMemcpyAsync(d_dst, d_src, (size_t)bytes, hipMemcpyDeviceToDevice,
                               resolve(stream)));
  return FAP_OK;
}

int fap_hip_memcpy2d_d2h_i32(int32_t *h_dst, int64_t h_ld, const int32_t *d_src, int64_t d_ld,
                             int64_t rows, int64_t cols, void *stream) {
  if (rows == 0 || cols == 0) return FAP_OK;
  FAP_REQUIRE(h_dst && d_src && rows > 0 && cols > 0 && h_ld >= cols && d_ld >= cols,
              "fap_hip_memcpy2d_d2h_i32: invalid arguments");
  hipStream_t s = resolve(stream);
  FAP_CHECK_HIP(hipMemcpy2DAsync(h_dst, (size_t)h_ld * 4, d_src, (size_t)d_ld * 4,
                                 (size_t)cols * 4, (size_t)rows, hipMemcpyDeviceToHost, s));
  FAP_CHECK_HIP(hipStreamSynchronize(s));
  return FAP_OK;
}

int fap_hip_fill_i32(int32_t *d_ptr, int32_t value, int64_t count, void *stream) {
  if (count == 0) return FAP_OK;
  FAP_REQUIRE(d_ptr && count > 0, "fap_hip_fill_i32: invalid arguments");
  if (count % 4 == 0 && (uintptr_t)d_ptr % 16 == 0) {
    int64_t blocks = (count / 4 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_i32x4_kernel, dim3((unsigned)blocks), dim3(256), 0, resolve(stream),
                       reinterpret_cast<int4 *>(d_ptr), value, count / 4);
  } else {
    int64_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, resolve(stream),
                       d_ptr, value, count);
  }
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int fap_hip_copy_i32(int32_t *d_dst, const int32_t *d_src, int64_t count, void *stream) {
  if (count == 0) return FAP_OK;
  FAP_REQUIRE(d_dst && d_src && count > 0, "fap_hip_copy_i32: invalid arguments");
  if (count % 4 == 0 && (uintptr_t)d_dst % 16 == 0 && (uintptr_t)d_src % 16 == 0) {
    int64_t blocks = (count / 4 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(copy_i32x4_kernel, dim3((unsigned)blocks), dim3(256), 0, resolve(stream),
                       reinterpret_cast<int4 *>(d_dst), reinterpret_cast<const int4 *>(d_src),
                       count / 4);
  } else {
    int64_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(copy_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, resolve(stream),
                       d_dst, d_src, count);
  }
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int fap_hip_stream_sync(void *stream) {
  FAP_CHECK_HIP(hipStreamSynchronize(resolve(stream)));
  return FAP_OK;
}

int fap_hip_device_sync(void) {
  FAP_CHECK_HIP(hipDeviceSynchronize());
  return FAP_OK;
}

int fap_hip_event_create(void **ev) {
  FAP_REQUIRE(ev, "fap_hip_event_create: null");
  hipEvent_t e;
  FAP_CHECK_HIP(hipEventCreate(&e));
  *ev = e;
  return FAP_OK;
}
int fap_hip_event_destroy(void *ev) {
  if (ev) FAP_CHECK_HIP(hipEventDestroy(static_cast<hipEvent_t>(ev)));
  return FAP_OK;
}
int fap_hip_event_record(void *ev, void *stream) {
  FAP_REQUIRE(ev, "fap_hip_event_record: null");
  FAP_CHECK_HIP(hipEventRecord(static_cast<hipEvent_t>(ev), resolve(stream)));
  return FAP_OK;
}
int fap_hip_event_query(void *ev, int32_t *done) {
  FAP_REQUIRE(ev && done, "fap_hip_event_query: null");
  const hipError_t e = hipEventQuery(static_cast<hipEvent_t>(ev));
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    *done = 0;
    return FAP_OK;
  }
  FAP_CHECK_HIP(e);
  *done = 1;
  return FAP_OK;
}

int fap_hip_stream_create(void **stream) {
  FAP_REQUIRE(stream, "fap_hip_stream_create: null");
  hipStream_t s;
  FAP_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
  *stream = s;
  return FAP_OK;
}

int fap_hip_stream_destroy(void *stream) {
  if (stream) FAP_CHECK_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return FAP_OK;
}

int fap_hip_event_elapsed_ms(void *a, void *b, float *ms) {
  FAP_REQUIRE(a && b && ms, "fap_hip_event_elapsed_ms: null");
  FAP_CHECK_HIP(hipEventSynchronize(static_cast<hipEvent_t>(b)));
  FAP_CHECK_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)));
  return FAP_OK;
}

}  // extern "C"
