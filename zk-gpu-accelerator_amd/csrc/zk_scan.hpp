// Device-wide exclusive prefix sums.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zk_runtime.hpp"

namespace zk {

// out[i] = in[0] + ... + in[i-1]; *total (device, optional) = the sum of all n.
// Temporary buffers come from sc.
int scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, hipStream_t stream, Scratch& sc);

}  // namespace zk
