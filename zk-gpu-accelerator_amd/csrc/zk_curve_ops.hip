// Element-wise G1 kernels, the on-curve check and fixed-base multiples
// (include/zkgpu.h, "G1 point kernels").
#include <mutex>

#include "../../include/zkgpu.h"
#include "zk_curve.hpp"
#include "zk_fields.hpp"

namespace zk {

using Fq = FqConst;

template <int kind>  // 0: add, 1: add_affine, 2: double
__global__ void __launch_bounds__(256) k_g1(const uint64_t* __restrict__ a, int64_t a_step, const uint64_t* __restrict__ b,
                                            int64_t b_step, uint64_t* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Fq f;
  Jac p = load_jac(f, a + i * a_step), r;
  if constexpr (kind == 0) r = jac_add(f, p, load_jac(f, b + i * b_step));
  else if constexpr (kind == 1) r = jac_madd(f, p, load_aff(f, b + i * b_step));
  else r = jac_dbl(f, p);
  store_jac(f, out + 12 * i, r);
}

// first_bad = min index whose affine point is off the curve (initialised to n by the host).
__global__ void __launch_bounds__(256) k_on_curve(const uint64_t* __restrict__ pts, int64_t n,
                                                  unsigned long long* first_bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Fq f;
  Aff q = load_aff(f, pts + 8 * i);
  Fe three = fadd(f, fdbl(f, fe_const<Fq>(f.one)), fe_const<Fq>(f.one));
  Fe lhs = mont_sqr(f, q.y);
  Fe rhs = fadd(f, mont(f, mont_sqr(f, q.x), q.x), three);
  if (!fe_eq(lhs, rhs)) atomicMin(first_bad, (unsigned long long)i);
}

// Table of d * 2^(8j) * G for j < 32, d < 256 (Jacobian, 261-form): thread j walks d upwards.
__global__ void k_fb_table_jac(Jac* table) {
  const int j = threadIdx.x;
  if (j >= 32) return;
  const Fq f;
  Jac g;
  g.x = fe_const<Fq>(f.one);               // x = 1
  g.y = fadd(f, g.x, g.x);                 // y = 2
  g.z = fe_const<Fq>(f.one);
  for (int k = 0; k < 8 * j; ++k) g = jac_dbl(f, g);
  Jac acc = jac_identity(f);
  for (int d = 0; d < 256; ++d) {
    table[j * 256 + d] = acc;
    acc = jac_add(f, acc, g);
  }
}

__global__ void __launch_bounds__(256) k_fb_table_aff(const Jac* __restrict__ jt, Aff* __restrict__ at) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 32 * 256) return;
  const Fq f;
  Jac p = jt[i];
  Aff a;
  if (fe_is_zero(p.z)) {
    a.x = fe_zero();
    a.y = fe_zero();
  } else {
    a = jac_to_aff(f, p);
  }
  at[i] = a;
}

// out[i] = k[i] * G (affine 256-form rows; zero rows for the identity).
__global__ void __launch_bounds__(256) k_fixed_base(const uint64_t* __restrict__ k, int64_t n, const Aff* __restrict__ table,
                                                    uint64_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Fq f;
  const uint4* kp = reinterpret_cast<const uint4*>(k + 4 * i);
  uint4 k0 = kp[0], k1 = kp[1];
  uint32_t w[8] = {k0.x, k0.y, k0.z, k0.w, k1.x, k1.y, k1.z, k1.w};
  Jac acc = jac_identity(f);
  for (int j = 0; j < 32; ++j) {
    uint32_t d = (w[j >> 2] >> (8 * (j & 3))) & 0xffu;
    if (d) acc = jac_madd(f, acc, table[j * 256 + d]);
  }
  if (fe_is_zero(acc.z)) {
    store_row(out + 8 * i, fe_zero());
    store_row(out + 8 * i + 4, fe_zero());
    return;
  }
  Aff a = jac_to_aff(f, acc);
  store_coord(f, out + 8 * i, a.x);
  store_coord(f, out + 8 * i + 4, a.y);
}

static std::mutex g_fb_mu;
static Aff* g_fb_table = nullptr;

static int fixed_base_table(hipStream_t stream, const Aff** out) {
  std::lock_guard<std::mutex> lk(g_fb_mu);
  if (!g_fb_table) {
    Jac* jt = nullptr;
    Aff* at = nullptr;
    ZK_HIP(hipMalloc(&jt, sizeof(Jac) * 32 * 256));
    ZK_HIP(hipMalloc(&at, sizeof(Aff) * 32 * 256));
    hipLaunchKernelGGL(k_fb_table_jac, dim3(1), dim3(64), 0, stream, jt);
    ZK_LAUNCH_CHECK("fixed-base table");
    hipLaunchKernelGGL(k_fb_table_aff, dim3(32), dim3(256), 0, stream, jt, at);
    ZK_LAUNCH_CHECK("fixed-base table (affine)");
    ZK_HIP(hipStreamSynchronize(stream));
    (void)hipFree(jt);
    g_fb_table = at;
  }
  *out = g_fb_table;
  return 0;
}

}  // namespace zk

using namespace zk;

extern "C" {

int zk_g1_add(const uint64_t* a, int a_bcast, const uint64_t* b, int b_bcast, uint64_t* out, int64_t n, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_g1<0>, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, a, a_bcast ? (int64_t)0 : (int64_t)12,
                     b, b_bcast ? (int64_t)0 : (int64_t)12, out, n);
  ZK_LAUNCH_CHECK("g1 add");
  return 0;
}

int zk_g1_add_affine(const uint64_t* a, int a_bcast, const uint64_t* b, int b_bcast, uint64_t* out, int64_t n,
                     void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_g1<1>, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, a, a_bcast ? (int64_t)0 : (int64_t)12,
                     b, b_bcast ? (int64_t)0 : (int64_t)8, out, n);
  ZK_LAUNCH_CHECK("g1 add_affine");
  return 0;
}

int zk_g1_double(const uint64_t* a, uint64_t* out, int64_t n, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_g1<2>, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, a, (int64_t)12,
                     (const uint64_t*)nullptr, (int64_t)0, out, n);
  ZK_LAUNCH_CHECK("g1 double");
  return 0;
}

int zk_g1_on_curve(const uint64_t* points_affine, int64_t n, int64_t* first_bad, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  *first_bad = -1;
  if (n <= 0) return 0;
  Scratch sc(stream);
  unsigned long long* d = sc.get<unsigned long long>(1);
  if (!d) return 1;
  unsigned long long init = (unsigned long long)n;
  ZK_HIP(hipMemcpyAsync(d, &init, 8, hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_on_curve, dim3(grid_for(n, 256)), dim3(256), 0, stream, points_affine, n, d);
  ZK_LAUNCH_CHECK("on-curve check");
  unsigned long long res = 0;
  ZK_HIP(hipMemcpyAsync(&res, d, 8, hipMemcpyDeviceToHost, stream));
  ZK_HIP(hipStreamSynchronize(stream));
  *first_bad = res >= (unsigned long long)n ? -1 : (int64_t)res;
  return 0;
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

int zk_g1_on_curve_async(const uint64_t* points_affine, int64_t n, uint64_t* first_bad_dev, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  unsigned long long* d = reinterpret_cast<unsigned long long*>(first_bad_dev);
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, stream, d, (unsigned long long)(n > 0 ? n : 0));
  ZK_LAUNCH_CHECK("on-curve check");
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_on_curve, dim3(grid_for(n, 256)), dim3(256), 0, stream, points_affine, n, d);
  ZK_LAUNCH_CHECK("on-curve check");
  return 0;
}

int zk_g1_fixed_base(const uint64_t* k, int64_t n, uint64_t* out_affine, void* stream) {
  if (n <= 0) return 0;
  const Aff* table = nullptr;
  if (int rc = fixed_base_table((hipStream_t)stream, &table)) return rc;
  hipLaunchKernelGGL(k_fixed_base, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, k, n, table, out_affine);
  ZK_LAUNCH_CHECK("fixed-base multiples");
  return 0;
}

}  // extern "C"
