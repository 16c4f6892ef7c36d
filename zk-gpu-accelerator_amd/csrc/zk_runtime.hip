// Lifecycle, memory, streams and events of the C ABI (include/zkgpu.h).
#include "zk_runtime.hpp"

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/zkgpu.h"
#include "zk_msm.hpp"

namespace zk {

static thread_local std::string g_last_error;

int fail(const std::string& msg) {
  g_last_error = msg;
  return 1;
}

int fail_hip(hipError_t err, const char* what) {
  return fail(std::string(what) + ": " + hipGetErrorString(err));
}

// ---- caching allocator -------------------------------------------------
// Size classes: powers of two up to 1 MiB, then multiples of 2 MiB.
static size_t size_class(size_t bytes) {
  if (bytes <= 256) return 256;
  if (bytes <= (1u << 20)) {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
  }
  const size_t g = size_t(2) << 20;
  return (bytes + g - 1) / g * g;
}

struct Block {
  size_t bytes;
  hipStream_t stream;
};

static std::mutex g_alloc_mu;
static std::unordered_map<void*, Block> g_live;
static std::map<std::pair<hipStream_t, size_t>, std::vector<void*>> g_free;
static size_t g_cached_bytes = 0;

static void release_cache_locked() {
  for (auto& kv : g_free)
    for (void* p : kv.second) (void)hipFree(p);
  g_free.clear();
  g_cached_bytes = 0;
}

void* dev_alloc(size_t bytes, hipStream_t stream) {
  size_t c = size_class(bytes);
  {
    // the smallest cached block of this stream that holds the request and is at most twice its size
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    for (auto it = g_free.lower_bound({stream, c}); it != g_free.end() && it->first.first == stream &&
                                                    it->first.second <= 2 * c;
         ++it) {
      if (it->second.empty()) continue;
      void* p = it->second.back();
      it->second.pop_back();
      g_cached_bytes -= it->first.second;
      g_live[p] = Block{it->first.second, stream};
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, c);
  if (e != hipSuccess) {
    // Out of memory with blocks parked in the cache: give them back and retry once.
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    (void)hipGetLastError();
    (void)hipDeviceSynchronize();
    release_cache_locked();
    e = hipMalloc(&p, c);
    if (e != hipSuccess) {
      fail_hip(e, "hipMalloc");
      return nullptr;
    }
  }
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_live[p] = Block{c, stream};
  return p;
}

void dev_free(void* ptr) {
  if (!ptr) return;
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  auto it = g_live.find(ptr);
  if (it == g_live.end()) return;
  Block b = it->second;
  g_live.erase(it);
  g_free[{b.stream, b.bytes}].push_back(ptr);
  g_cached_bytes += b.bytes;
}

Scratch::~Scratch() {
  for (int i = 0; i < count_; ++i) dev_free(ptrs_[i]);
}

void* Scratch::raw(size_t bytes) {
  if (count_ == kMax) {
    fail("internal: too many scratch buffers in one call");
    return nullptr;
  }
  void* p = dev_alloc(bytes ? bytes : 1, stream_);
  if (p) ptrs_[count_++] = p;
  return p;
}

// ---- opt-in kernel timing ------------------------------------------------
struct ProfRecord {
  std::string name;
  hipEvent_t start, end;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRecord> g_prof;

bool prof_begin(hipStream_t stream, hipEvent_t* start, hipEvent_t* stop) {
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_on) return false;
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return false;
  if (hipEventCreate(start) != hipSuccess) return false;
  if (hipEventCreate(stop) != hipSuccess) {
    (void)hipEventDestroy(*start);
    return false;
  }
  return true;
}

void prof_commit(const char* name, hipEvent_t start, hipEvent_t stop) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.push_back(ProfRecord{name, start, stop});
}

bool launch_escapes(hipStream_t alloc_stream, hipStream_t stream) {
  if (stream != alloc_stream) return true;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone;
}

static int g_cus = 0;
int device_cus() { return g_cus > 0 ? g_cus : 256; }

// Plain grid-stride copy of 16-byte words: every byte read once and written once.
__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t words) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) dst[i] = src[i];
}

__global__ void probe_kernel(int64_t* x, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + 1;
}

// v_mad_u64_u32 throughput: eight independent 64-bit accumulators per thread,
// each step one 32x32+64 multiply-add (the instruction every field product here is built from).
__global__ void __launch_bounds__(256) k_mad_peak(uint32_t* out, int iters, uint32_t seed) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) acc[c] = (uint64_t)(seed ^ (t * 2654435761u + c)) << 7;
  const uint32_t b = seed * 40503u + t;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[c] = (uint64_t)(uint32_t)acc[c] * b + acc[c];  // multiplicand from the chain
    }
  }
  uint64_t s = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) s ^= acc[c];
  out[t] = (uint32_t)s ^ (uint32_t)(s >> 32);
}

}  // namespace zk

using namespace zk;

extern "C" {

const char* zk_last_error(void) { return g_last_error.c_str(); }

int zk_device_count(int* count) {
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    (void)hipGetLastError();
    return fail_hip(e, "hipGetDeviceCount");
  }
  return 0;
}

int zk_init(int device) {
  ZK_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  ZK_HIP(hipGetDeviceProperties(&prop, device));
  g_cus = prop.multiProcessorCount;
  return msm_scratch_init();
}

int zk_device_info(char* name, int name_len, int* compute_units, int* clock_khz) {
  int dev = 0;
  ZK_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  ZK_HIP(hipGetDeviceProperties(&prop, dev));
  if (name && name_len > 0) {
    snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (compute_units) *compute_units = prop.multiProcessorCount;
  if (clock_khz) {
    int clk = 0;
    ZK_HIP(hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, dev));
    *clock_khz = clk;
  }
  return 0;
}

int zk_mem_info(uint64_t* free_bytes, uint64_t* total_bytes) {
  size_t f = 0, t = 0;
  ZK_HIP(hipMemGetInfo(&f, &t));
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  *free_bytes = (uint64_t)f + g_cached_bytes;  // cached blocks are free to this library
  *total_bytes = (uint64_t)t;
  return 0;
}

int zk_alloc(void** ptr, uint64_t bytes, void* stream) {
  *ptr = dev_alloc((size_t)bytes, (hipStream_t)stream);
  return *ptr ? 0 : 1;
}

int zk_free(void* ptr) {
  dev_free(ptr);
  return 0;
}

int zk_memcpy_h2d(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (!bytes) return 0;
  ZK_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  ZK_HIP(hipStreamSynchronize((hipStream_t)stream));  // the host buffer may be released after return
  return 0;
}

// Pinned host blocks: freed blocks stay pinned in a cache of at most kHostCacheBytes (pinning
// costs milliseconds per hundred MB, so the public calls reuse their staging blocks).
namespace {
constexpr size_t kHostCacheBytes = (size_t)2 << 30;
std::mutex g_host_mu;
std::map<void*, size_t> g_host_live;        // block -> size
std::multimap<size_t, void*> g_host_cache;  // size -> block
size_t g_host_cached = 0;
}  // namespace

int zk_host_alloc(void** ptr, uint64_t bytes) {
  *ptr = nullptr;
  const size_t want = std::max<size_t>((size_t)bytes, 1);
  {
    std::lock_guard<std::mutex> lock(g_host_mu);
    auto it = g_host_cache.lower_bound(want);
    if (it != g_host_cache.end() && it->first <= 2 * want) {  // reuse a block not much larger
      *ptr = it->second;
      g_host_live[it->second] = it->first;
      g_host_cached -= it->first;
      g_host_cache.erase(it);
      return 0;
    }
  }
  const size_t size = (want + 0xfffff) & ~(size_t)0xfffff;  // whole MiB
  void* p = nullptr;
  ZK_HIP(hipHostMalloc(&p, size, hipHostMallocDefault));
  std::lock_guard<std::mutex> lock(g_host_mu);
  g_host_live[p] = size;
  *ptr = p;
  return 0;
}

int zk_host_free(void* ptr) {
  if (!ptr) return 0;
  std::vector<void*> drop;
  {
    std::lock_guard<std::mutex> lock(g_host_mu);
    auto it = g_host_live.find(ptr);
    if (it == g_host_live.end()) return fail("zk_host_free: not a block from zk_host_alloc");
    const size_t size = it->second;
    g_host_live.erase(it);
    g_host_cache.emplace(size, ptr);
    g_host_cached += size;
    while (g_host_cached > kHostCacheBytes) {  // the smallest blocks go first
      auto small = g_host_cache.begin();
      g_host_cached -= small->first;
      drop.push_back(small->second);
      g_host_cache.erase(small);
    }
  }
  for (void* p : drop) ZK_HIP(hipHostFree(p));
  return 0;
}

int zk_memcpy_h2d_async(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (!bytes) return 0;
  ZK_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return 0;
}

int zk_memcpy_d2h(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (!bytes) return 0;
  ZK_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  ZK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int zk_memcpy_d2d(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (!bytes) return 0;
  ZK_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return 0;
}

int zk_copy(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) return fail("zk_copy needs 16-byte aligned buffers and sizes");
  if (!bytes) return 0;
  const int64_t words = (int64_t)(bytes / 16);
  const int64_t blocks = std::min<int64_t>((words + 255) / 256, (int64_t)device_cus() * 16);
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, static_cast<const uint4*>(src),
                     static_cast<uint4*>(dst), words);
  ZK_LAUNCH_CHECK("copy");
  return 0;
}

int zk_memset(void* dst, int value, uint64_t bytes, void* stream) {
  if (!bytes) return 0;
  ZK_HIP(hipMemsetAsync(dst, value, (size_t)bytes, (hipStream_t)stream));
  return 0;
}

int zk_stream_create(void** stream) {
  hipStream_t s;
  ZK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return 0;
}

int zk_stream_destroy(void* stream) {
  if (!stream) return 0;
  // Its work finishes first: the cached blocks may still be read by it.
  ZK_HIP(hipStreamSynchronize((hipStream_t)stream));
  {
    // Blocks of this stream, cached or live, are ordered by the null stream from now on (a later
    // stream may get the same handle).
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    for (auto it = g_free.begin(); it != g_free.end();) {
      if (it->first.first == (hipStream_t)stream) {
        for (void* p : it->second) g_free[{nullptr, it->first.second}].push_back(p);
        it = g_free.erase(it);
      } else {
        ++it;
      }
    }
    for (auto& kv : g_live)
      if (kv.second.stream == (hipStream_t)stream) kv.second.stream = nullptr;
  }
  ZK_HIP(hipStreamDestroy((hipStream_t)stream));
  return 0;
}

int zk_stream_sync(void* stream) {
  ZK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int zk_device_sync(void) {
  ZK_HIP(hipDeviceSynchronize());
  return 0;
}

int zk_event_create(void** event) {
  hipEvent_t e;
  ZK_HIP(hipEventCreate(&e));
  *event = (void*)e;
  return 0;
}

int zk_event_destroy(void* event) {
  if (event) ZK_HIP(hipEventDestroy((hipEvent_t)event));
  return 0;
}

int zk_event_record(void* event, void* stream) {
  ZK_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return 0;
}

int zk_event_elapsed_ms(void* start, void* end, float* ms) {
  ZK_HIP(hipEventSynchronize((hipEvent_t)end));
  ZK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
  return 0;
}

int zk_event_query(void* event, int* done) {
  hipError_t e = hipEventQuery((hipEvent_t)event);
  if (e == hipSuccess) {
    *done = 1;
    return 0;
  }
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    *done = 0;
    return 0;
  }
  return fail_hip(e, "hipEventQuery");
}

int zk_event_sync(void* event) {
  ZK_HIP(hipEventSynchronize((hipEvent_t)event));
  return 0;
}

int zk_stream_wait_event(void* stream, void* event) {
  ZK_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
  return 0;
}

int zk_prof_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = on != 0;
  return 0;
}

int zk_prof_read(const char* name, double* total_ms, int64_t* count, int reset) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  double sum = 0;
  int64_t c = 0;
  for (auto& r : g_prof) {
    if (r.name != name) continue;
    float ms = 0;
    ZK_HIP(hipEventSynchronize(r.end));
    ZK_HIP(hipEventElapsedTime(&ms, r.start, r.end));
    sum += ms;
    ++c;
  }
  *total_ms = sum;
  *count = c;
  if (reset) {
    for (auto& r : g_prof) {
      (void)hipEventDestroy(r.start);
      (void)hipEventDestroy(r.end);
    }
    g_prof.clear();
  }
  return 0;
}

int zk_graph_begin(void* stream) {
  if (!stream) return fail("graph capture needs a created stream, not the null stream");
  ZK_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return 0;
}

int zk_graph_end(void* stream, void** graph_exec) {
  hipGraph_t g = nullptr;
  ZK_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail_hip(e, "hipGraphInstantiate");
  *graph_exec = (void*)ex;
  return 0;
}

int zk_graph_launch(void* graph_exec, void* stream) {
  ZK_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
  return 0;
}

int zk_graph_destroy(void* graph_exec) {
  if (graph_exec) ZK_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
  return 0;
}

int zk_valu_peak(double* mads_per_second_t) {
  const int blocks = device_cus() * 8, threads = 256, iters = 2000;
  uint32_t* out = nullptr;
  ZK_HIP(hipMalloc(&out, (size_t)blocks * threads * 4));
  hipEvent_t e0, e1;
  ZK_HIP(hipEventCreate(&e0));
  ZK_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_mad_peak, dim3(blocks), dim3(threads), 0, 0, out, 10, 1u);  // warm up
  ZK_HIP(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(k_mad_peak, dim3(blocks), dim3(threads), 0, 0, out, iters, 7u);
  ZK_HIP(hipEventRecord(e1, 0));
  ZK_HIP(hipEventSynchronize(e1));
  float ms = 0;
  ZK_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  const double mads = (double)blocks * threads * iters * 16.0 * 8.0;
  *mads_per_second_t = mads / (ms * 1e-3) / 1e12;
  return 0;
}

int zk_selftest(void) {
  const int n = 4;
  int64_t host[n] = {0, 1, 2, 3};
  int64_t* dev = nullptr;
  ZK_HIP(hipMalloc(&dev, sizeof(host)));
  hipError_t e = hipMemcpy(dev, host, sizeof(host), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(probe_kernel, dim3(1), dim3(64), 0, 0, dev, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(host, dev, sizeof(host), hipMemcpyDeviceToHost);
  (void)hipFree(dev);
  if (e != hipSuccess) return fail_hip(e, "probe kernel");
  for (int i = 0; i < n; ++i)
    if (host[i] != i + 1) return fail("probe kernel ran but returned wrong values");
  return 0;
}

}  // extern "C"
