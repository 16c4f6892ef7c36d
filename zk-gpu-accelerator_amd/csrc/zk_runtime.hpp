// Host-side runtime shared by every translation unit of libzkgpu_hip.so:
// error reporting, the caching device allocator and scratch buffers.
#pragma once

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

// Opt-in kernel timing (zk_prof_enable): a Prof brackets one launch with HIP
// events on its stream; zk_prof_read sums the durations per name.
class Prof {
 public:
  Prof(const char* name, hipStream_t stream);
  ~Prof();
  Prof(const Prof&) = delete;
  Prof& operator=(const Prof&) = delete;

 private:
  const char* name_;
  hipStream_t stream_;
  hipEvent_t start_ = nullptr;
};

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
