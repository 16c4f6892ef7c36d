// Device-wide exclusive prefix sums.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zk_runtime.hpp"

namespace zk {

// out[i] = in[0] + ... + in[i-1]; *total (device, optional) = the sum of all n.
// Temporary buffers come from sc.
int scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, hipStream_t stream, Scratch& sc);

// The same with a caller-provided workspace of scan_workspace_words(n) words
// (no allocation, so it can be captured into a graph).
size_t scan_workspace_words(int64_t n);
int scan_u32_ws(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, uint32_t* ws, hipStream_t stream);

}  // namespace zk
