// MSM entry points shared by the plain and batched paths.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zk {

// Window (bits per signed digit) chosen from the device cost model for n points.
int default_window(int64_t n);

// out_jac (device, 12 limbs, 256-form) = sum s_i P_i; c <= 0 and k0 <= 0 pick defaults.
// entries (optional): receives the sorted-entry count the accumulation processed (synchronises).
int msm_g1(const uint64_t* points, const uint64_t* scalars, int64_t n, int c, int k0, uint64_t* out_jac, hipStream_t stream,
           uint64_t* entries = nullptr);

// Looks up the current device's address of the no-alloc MSM's scratch array (once per device).
int msm_scratch_init();

}  // namespace zk
