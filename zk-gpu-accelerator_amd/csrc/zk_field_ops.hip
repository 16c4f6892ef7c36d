// Element-wise field kernels and the field registry (include/zkgpu.h, "field arithmetic").
//
// One thread per row, 256 threads per block: each row is 32 bytes read as
// two 16-byte loads, so a wavefront reads 2 KiB of consecutive rows. These
// kernels are HBM-bound (a multiplication is two 9-limb Montgomery products
// against 96 bytes of traffic).
#include <mutex>
#include <string>
#include <vector>

#include "../../include/zkgpu.h"
#include "zk_fields.hpp"

namespace zk {

// ---- registry ------------------------------------------------------------

static std::mutex g_field_mu;
static std::vector<FieldRT> g_fields;  // id = index + 2

bool lookup_field(int id, FieldRT* out) {
  std::lock_guard<std::mutex> lk(g_field_mu);
  int i = id - 2;
  if (i < 0 || i >= (int)g_fields.size()) return false;
  *out = g_fields[(size_t)i];
  return true;
}

namespace {
struct Big {  // 320-bit unsigned scratch for host-side constants
  uint64_t w[5] = {0, 0, 0, 0, 0};
};
bool geq(const Big& a, const Big& b) {
  for (int i = 4; i >= 0; --i)
    if (a.w[i] != b.w[i]) return a.w[i] > b.w[i];
  return true;
}
void sub_in(Big& a, const Big& b) {
  unsigned __int128 br = 0;
  for (int i = 0; i < 5; ++i) {
    unsigned __int128 d = (unsigned __int128)a.w[i] - b.w[i] - br;
    a.w[i] = (uint64_t)d;
    br = (d >> 64) ? 1 : 0;
  }
}
void dbl_mod(Big& a, const Big& m) {
  uint64_t carry = 0;
  for (int i = 0; i < 5; ++i) {
    uint64_t nc = a.w[i] >> 63;
    a.w[i] = (a.w[i] << 1) | carry;
    carry = nc;
  }
  if (geq(a, m)) sub_in(a, m);
}
void to_limbs29(const Big& a, uint32_t out[NL]) {
  for (int k = 0; k < NL; ++k) {
    int bit = LB * k, word = bit / 64, sh = bit % 64;
    uint64_t v = a.w[word] >> sh;
    if (sh > 64 - LB && word + 1 < 5) v |= a.w[word + 1] << (64 - sh);
    out[k] = (uint32_t)v & LMASK;
  }
}
Big pow2_mod(int k, const Big& m) {
  Big x;
  x.w[0] = 1;
  if (geq(x, m)) sub_in(x, m);
  for (int i = 0; i < k; ++i) dbl_mod(x, m);
  return x;
}
}  // namespace

bool make_field_rt(const uint64_t modulus[4], FieldRT* out, std::string* err) {
  Big m;
  for (int i = 0; i < 4; ++i) m.w[i] = modulus[i];
  bool odd = modulus[0] & 1;
  bool below = (modulus[3] >> 63) == 0;
  bool at_least_3 = modulus[1] | modulus[2] | modulus[3] || modulus[0] >= 3;
  if (!odd || !below || !at_least_3) {
    *err = "modulus must be odd and in [3, 2^255)";
    return false;
  }
  to_limbs29(m, out->p);
  Big m2;  // 2p < 2^256
  uint64_t carry = 0;
  for (int i = 0; i < 4; ++i) {
    m2.w[i] = (m.w[i] << 1) | carry;
    carry = m.w[i] >> 63;
  }
  to_limbs29(m2, out->p2);
  to_limbs29(pow2_mod(256, m), out->c256);
  to_limbs29(pow2_mod(266, m), out->c266);
  to_limbs29(pow2_mod(517, m), out->c517);
  to_limbs29(pow2_mod(522, m), out->c522);
  to_limbs29(pow2_mod(LB * NL, m), out->one);
  for (int i = 0; i < 8; ++i) out->p32[i] = (uint32_t)(modulus[i / 2] >> (32 * (i % 2)));
  uint32_t inv = 1;  // Newton: inv = p^-1 mod 2^32
  for (int i = 0; i < 5; ++i) inv *= 2u - (uint32_t)modulus[0] * inv;
  out->ninv = (0u - inv) & LMASK;
  return true;
}

// ---- kernels --------------------------------------------------------------

enum class Op { kMul, kAdd, kSub };

template <class F, Op op>
__global__ void __launch_bounds__(256) k_binary(F f, const uint64_t* __restrict__ a, int64_t a_step,
                                                const uint64_t* __restrict__ b, int64_t b_step,
                                                uint64_t* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fe x = load_row(a + i * a_step), y = load_row(b + i * b_step), r;
  if constexpr (op == Op::kMul) r = mul256(f, x, y);
  else if constexpr (op == Op::kAdd) r = fadd(f, x, y);
  else r = fsub(f, x, y);
  store_row(out + 4 * i, r);
}

template <class F, bool to>
__global__ void __launch_bounds__(256) k_convert(F f, const uint64_t* __restrict__ in, uint64_t* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fe x = load_row(in + 4 * i);
  store_row(out + 4 * i, to ? plain_to_256(f, x) : form256_to_plain(f, x));
}

template <Op op>
static int binary(int field_id, const uint64_t* a, int a_bcast, const uint64_t* b, int b_bcast, uint64_t* out,
                  int64_t n, void* stream) {
  if (n <= 0) return 0;
  return with_field(field_id, [&](auto f) {
    hipLaunchKernelGGL((k_binary<decltype(f), op>), dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, f, a,
                       a_bcast ? (int64_t)0 : (int64_t)4, b, b_bcast ? (int64_t)0 : (int64_t)4, out, n);
    ZK_LAUNCH_CHECK("field kernel launch");
    return 0;
  });
}

}  // namespace zk

using namespace zk;

// Product throughput: every thread runs kChains independent chains of 9-limb Montgomery products
// (the serial column schedule the throughput kernels use), values kept below 2p. No memory traffic
// beyond one store per thread: the rate is the ceiling the field arithmetic itself sets.
constexpr int kPeakChains = 4;

__global__ void __launch_bounds__(256) k_product_peak(uint32_t* out, int iters, uint32_t seed) {
  const FqConst f;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  Fe a[kPeakChains], b;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    b.l[i] = (seed * 2654435761u + t * 40503u + (uint32_t)i * 97u) & (LMASK >> 2);
#pragma unroll
    for (int c = 0; c < kPeakChains; ++c) a[c].l[i] = (t * 2246822519u + (uint32_t)(c * 9 + i) * 3266489917u) & (LMASK >> 2);
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int c = 0; c < kPeakChains; ++c) a[c] = mont_raw_serial(f, a[c], b);
  }
  uint32_t x = 0;
#pragma unroll
  for (int c = 0; c < kPeakChains; ++c)
#pragma unroll
    for (int i = 0; i < NL; ++i) x ^= a[c].l[i];
  out[t] = x;
}

extern "C" {

int zk_product_peak(double* gproducts_per_s) {
  const int blocks = device_cus() * 12, threads = 256, iters = 400;
  uint32_t* out = nullptr;
  ZK_HIP(hipMalloc(&out, (size_t)blocks * threads * 4));
  hipEvent_t e0, e1;
  ZK_HIP(hipEventCreate(&e0));
  ZK_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_product_peak, dim3(blocks), dim3(threads), 0, 0, out, 4, 1u);  // warm up
  ZK_HIP(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(k_product_peak, dim3(blocks), dim3(threads), 0, 0, out, iters, 7u);
  ZK_HIP(hipEventRecord(e1, 0));
  ZK_HIP(hipEventSynchronize(e1));
  float ms = 0;
  ZK_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  *gproducts_per_s = (double)blocks * threads * iters * kPeakChains / (ms * 1e-3) / 1e9;
  return 0;
}

int zk_field_register(const uint64_t modulus[4], int* field_id) {
  FieldRT f;
  std::string err;
  if (!make_field_rt(modulus, &f, &err)) return fail(err);
  std::lock_guard<std::mutex> lk(g_field_mu);
  for (size_t i = 0; i < g_fields.size(); ++i) {
    bool same = true;
    for (int k = 0; k < NL; ++k) same = same && g_fields[i].p[k] == f.p[k];
    if (same) {
      *field_id = (int)i + 2;
      return 0;
    }
  }
  g_fields.push_back(f);
  *field_id = (int)g_fields.size() + 1;
  return 0;
}

int zk_field_to_mont(int field_id, const uint64_t* plain, uint64_t* mont, int64_t n, void* stream) {
  if (n <= 0) return 0;
  return with_field(field_id, [&](auto f) {
    hipLaunchKernelGGL((k_convert<decltype(f), true>), dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, f,
                       plain, mont, n);
    ZK_LAUNCH_CHECK("to_mont launch");
    return 0;
  });
}

int zk_field_from_mont(int field_id, const uint64_t* mont, uint64_t* plain, int64_t n, void* stream) {
  if (n <= 0) return 0;
  return with_field(field_id, [&](auto f) {
    hipLaunchKernelGGL((k_convert<decltype(f), false>), dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, f,
                       mont, plain, n);
    ZK_LAUNCH_CHECK("from_mont launch");
    return 0;
  });
}

int zk_field_mul(int field_id, const uint64_t* a, int a_bcast, const uint64_t* b, int b_bcast, uint64_t* out, int64_t n,
                 void* stream) {
  return binary<Op::kMul>(field_id, a, a_bcast, b, b_bcast, out, n, stream);
}

int zk_field_add(int field_id, const uint64_t* a, int a_bcast, const uint64_t* b, int b_bcast, uint64_t* out, int64_t n,
                 void* stream) {
  return binary<Op::kAdd>(field_id, a, a_bcast, b, b_bcast, out, n, stream);
}

int zk_field_sub(int field_id, const uint64_t* a, int a_bcast, const uint64_t* b, int b_bcast, uint64_t* out, int64_t n,
                 void* stream) {
  return binary<Op::kSub>(field_id, a, a_bcast, b, b_bcast, out, n, stream);
}

}  // extern "C"
