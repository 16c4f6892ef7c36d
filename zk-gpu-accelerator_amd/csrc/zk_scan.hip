// Exclusive prefix sums of uint32 counts on the device (used by the MSM's bucket sort).
//
// Each block scans 1024 entries (256 threads x 4), writes its total, the block
// totals are scanned recursively, and a final pass adds each block's offset.
#include "zk_scan.hpp"

namespace zk {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanBlock = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* sh, uint32_t* total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    uint32_t add = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  uint32_t incl = sh[t];
  if (total) *total = sh[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_blocks(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                              int64_t n, uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t sh[kScanThreads];
  __shared__ uint32_t tot;
  const int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    sum += v[k];
  }
  uint32_t pre = block_exclusive(sum, sh, threadIdx.x == 0 ? &tot : nullptr);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = pre;
    pre += v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(uint32_t* __restrict__ out, int64_t n,
                                                           const uint32_t* __restrict__ block_offsets) {
  const int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanItems;
  const uint32_t off = block_offsets[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) out[base + k] += off;
}

__global__ void k_scan_total(const uint32_t* __restrict__ in, const uint32_t* __restrict__ out, int64_t n,
                             uint32_t* __restrict__ total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *total = n ? out[n - 1] + in[n - 1] : 0u;
}

size_t scan_workspace_words(int64_t n) {
  size_t words = 0;
  while (n > 1) {
    const int64_t blocks = (n + kScanBlock - 1) / kScanBlock;
    words += 2 * (size_t)blocks;
    if (blocks == 1) break;
    n = blocks;
  }
  return words + 2;
}

int scan_u32_ws(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, uint32_t* ws, hipStream_t stream) {
  if (n <= 0) {
    if (total) {
      hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, stream, in, out, (int64_t)0, total);
      ZK_LAUNCH_CHECK("scan total");
    }
    return 0;
  }
  const int64_t blocks = (n + kScanBlock - 1) / kScanBlock;
  uint32_t* sums = ws;
  uint32_t* offs = ws + blocks;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)blocks), dim3(kScanThreads), 0, stream, in, out, n, sums);
  ZK_LAUNCH_CHECK("scan blocks");
  if (blocks > 1) {
    if (int rc = scan_u32_ws(sums, offs, blocks, nullptr, ws + 2 * blocks, stream)) return rc;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)blocks), dim3(kScanThreads), 0, stream, out, n, offs);
    ZK_LAUNCH_CHECK("scan add");
  }
  if (total) {
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, stream, in, out, n, total);
    ZK_LAUNCH_CHECK("scan total");
  }
  return 0;
}

int scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, hipStream_t stream, Scratch& sc) {
  uint32_t* ws = sc.get<uint32_t>(scan_workspace_words(n));
  if (!ws) return 1;
  return scan_u32_ws(in, out, n, total, ws, stream);
}

}  // namespace zk
