// Host-side runtime shared by every translation unit of libzkgpu_hip.so:
// error reporting, the caching device allocator and scratch buffers.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace zk {

// Records msg as the calling thread's last error and returns a non-zero status.
int fail(const std::string& msg);
int fail_hip(hipError_t err, const char* what);

// Device memory from a per-size-class cache. Blocks are tagged with the
// stream they were allocated on and only reused on that stream, so reuse is
// ordered after every earlier use without a synchronisation.
void* dev_alloc(size_t bytes, hipStream_t stream);
void dev_free(void* ptr);

// Scratch memory owned by the library for one call. It is released to the
// cache (on the same stream) when the Scratch goes out of scope.
class Scratch {
 public:
  explicit Scratch(hipStream_t s) : stream_(s) {}
  ~Scratch();
  // Returns nullptr (after recording an error) if the allocation fails.
  template <class T>
  T* get(size_t count) { return static_cast<T*>(raw(count * sizeof(T))); }
  void* raw(size_t bytes);
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;

 private:
  static constexpr int kMax = 48;
  hipStream_t stream_;
  void* ptrs_[kMax] = {};
  int count_ = 0;
};

int device_cus();  // compute units of the current device

// Whether work on `stream` may outlive what the plan's allocating stream orders: the launch runs on
// another stream or is being recorded into a graph (which may be replayed on any stream). A plan
// that answered yes waits for the device before its memory returns to the allocator's cache.
bool launch_escapes(hipStream_t alloc_stream, hipStream_t stream);

// Opt-in kernel timing (zk_prof_enable): launch_timed passes a start and a stop event to
// hipExtLaunchKernelGGL, so the dispatch itself takes the timestamps (marker packets between
// kernels would add a gap of about 10 us each); zk_prof_read sums the durations per name.
// prof_begin creates the two events when timing is on and the stream is not being captured.
bool prof_begin(hipStream_t stream, hipEvent_t* start, hipEvent_t* stop);
void prof_commit(const char* name, hipEvent_t start, hipEvent_t stop);

template <class Kernel, class... Args>
void launch_timed(const char* name, Kernel kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t stream,
                  Args... args) {
  hipEvent_t start = nullptr, stop = nullptr;
  if (prof_begin(stream, &start, &stop)) {
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, start, stop, 0, args...);
    prof_commit(name, start, stop);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, args...);
  }
}

}  // namespace zk

#define ZK_HIP(expr)                                      \
  do {                                                    \
    hipError_t zk_err_ = (expr);                          \
    if (zk_err_ != hipSuccess) return ::zk::fail_hip(zk_err_, #expr); \
  } while (0)

#define ZK_LAUNCH_CHECK(what)                             \
  do {                                                    \
    hipError_t zk_err_ = hipGetLastError();               \
    if (zk_err_ != hipSuccess) return ::zk::fail_hip(zk_err_, what); \
  } while (0)
