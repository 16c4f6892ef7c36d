// The 64-bit prime-field path: elementwise products, sums and powers modulo any 2 <= m < 2^63
// over broadcast integer operands (widened to int64 and reduced into [0, m) on load), and the NTT
// of int64 vectors modulo a prime m < 2^63.
//
// Products use Barrett reduction (HAC 14.42 with base 2 and k = bit length of m) on the exact
// 128-bit product, so every modulus below 2^63 takes the same path. The NTT is a multi-pass
// Cooley-Tukey decomposition N = n_1 n_2 ... (at most 2^10 points per pass): each pass loads a
// tile of several sub-transforms into LDS, runs their radix-2 stages there with precomputed
// Shoup twiddles, applies the inter-pass twiddle while storing, and the last pass writes natural
// order (and n^-1 for an inverse). A 2^20 transform is two passes, one HBM round trip each.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/zkgpu.h"
#include "zk_fields.hpp"
#include "zk_ntt.hpp"
#include "zk_runtime.hpp"

namespace zk {
namespace {

typedef unsigned __int128 u128;

// Barrett constants for m: mu = floor(2^(2 len) / m), len = bit length of m. A power of two reduces
// with a mask instead (its mu can need 65 bits).
struct Mod64 {
  uint64_t m, mu;
  uint32_t len, pow2;
};

Mod64 make_mod(uint64_t m) {
  Mod64 k{m, 0, 0, 0};
  k.len = 64 - __builtin_clzll(m);
  k.pow2 = (m & (m - 1)) == 0;
  if (!k.pow2) k.mu = (uint64_t)(((u128)1 << (2 * k.len)) / m);
  return k;
}

// a * b mod m for a, b < m.
__host__ __device__ inline uint64_t mulmod(uint64_t a, uint64_t b, const Mod64& k) {
  const u128 x = (u128)a * b;  // < m^2 <= 2^(2 len)
  if (k.pow2) return (uint64_t)x & (k.m - 1);
  const uint64_t q1 = (uint64_t)(x >> (k.len - 1));                   // < 2^(len + 1)
  const uint64_t q3 = (uint64_t)(((u128)q1 * k.mu) >> (k.len + 1));  // floor(x / m) - {0, 1, 2}
  u128 r = x - (u128)q3 * k.m;
  if (r >= k.m) r -= k.m;
  if (r >= k.m) r -= k.m;
  return (uint64_t)r;
}

__host__ __device__ inline uint64_t addmod(uint64_t a, uint64_t b, uint64_t m) {  // a, b < m < 2^63
  const uint64_t s = a + b;
  return s >= m ? s - m : s;
}

__host__ __device__ inline uint64_t submod(uint64_t a, uint64_t b, uint64_t m) {
  return a >= b ? a - b : a + (m - b);
}

// a * w mod m with wq = floor(w 2^64 / m) precomputed (Shoup); any a < 2^64, w < m < 2^63.
__device__ inline uint64_t shoup(uint64_t a, uint64_t w, uint64_t wq, uint64_t m) {
  const uint64_t q = __umul64hi(a, wq);
  const uint64_t r = a * w - q * m;  // the exact value lies in [0, 2m)
  return r >= m ? r - m : r;
}

// The same product for m < 2^31 in 32-bit words: wq = floor(w 2^32 / m), any a < 2^32, w < m.
__device__ inline uint32_t shoup(uint32_t a, uint32_t w, uint32_t wq, uint32_t m) {
  const uint32_t q = __umulhi(a, wq);
  const uint32_t r = a * w - q * m;  // the exact value lies in [0, 2m)
  return r >= m ? r - m : r;
}

__device__ inline uint32_t addmod(uint32_t a, uint32_t b, uint32_t m) {  // a, b < m < 2^31
  const uint32_t s = a + b;
  return s >= m ? s - m : s;
}

__device__ inline uint32_t submod(uint32_t a, uint32_t b, uint32_t m) { return a >= b ? a - b : a + (m - b); }

// Floor-mod of an int64 into [0, m): in-range values skip the (software) 64-bit division.
__device__ inline uint64_t reduce_i64(int64_t x, uint64_t m) {
  if (x >= 0) {
    const uint64_t u = (uint64_t)x;
    return u < m ? u : u % m;
  }
  return m - 1 - (~(uint64_t)x) % m;  // ~x = -x - 1 >= 0
}

__device__ inline int64_t load_i64(const void* p, int dtype, uint64_t i) {
  switch (dtype) {
    case ZK_I8: return ((const int8_t*)p)[i];
    case ZK_I16: return ((const int16_t*)p)[i];
    case ZK_I32: return ((const int32_t*)p)[i];
    case ZK_U8: return ((const uint8_t*)p)[i];
    case ZK_U16: return ((const uint16_t*)p)[i];
    case ZK_U32: return ((const uint32_t*)p)[i];
    default: return ((const int64_t*)p)[i];  // ZK_I64, and ZK_U64 read as its int64 bit pattern
  }
}

constexpr int kMaxDims = 8;

struct Operand {
  const void* p;
  int64_t stride[kMaxDims];  // elements; 0 along broadcast dimensions
  int32_t dtype;
};

struct Shape {
  uint64_t dim[kMaxDims];
  int32_t ndim;
};

// Element offsets of the linear output index i in both operands. Idx is uint32_t when the output
// has fewer than 2^32 elements (the division is much cheaper).
template <class Idx>
__device__ inline void offsets(Idx i, const Shape& sh, const Operand& a, const Operand* b, int64_t* oa, int64_t* ob) {
  int64_t x = 0, y = 0;
  for (int d = sh.ndim - 1; d > 0; --d) {
    const Idx n = (Idx)sh.dim[d];
    const Idx q = i / n, c = i - q * n;
    x += (int64_t)c * a.stride[d];
    if (b) y += (int64_t)c * b->stride[d];
    i = q;
  }
  x += (int64_t)i * a.stride[0];
  if (b) y += (int64_t)i * b->stride[0];
  *oa = x;
  *ob = y;
}

template <int kOp, class Idx>
__global__ void __launch_bounds__(256) k_f64_binary(Operand a, Operand b, Shape sh, int64_t* __restrict__ out, uint64_t n,
                                                    Mod64 k) {
  const Idx step = (Idx)gridDim.x * 256;
  for (Idx i = (Idx)blockIdx.x * 256 + threadIdx.x; i < (Idx)n; i += step) {
    int64_t oa, ob;
    if (sh.ndim == 1) {
      oa = (int64_t)i * a.stride[0];
      ob = (int64_t)i * b.stride[0];
    } else {
      offsets<Idx>(i, sh, a, &b, &oa, &ob);
    }
    const uint64_t x = reduce_i64(load_i64(a.p, a.dtype, oa), k.m);
    const uint64_t y = reduce_i64(load_i64(b.p, b.dtype, ob), k.m);
    out[i] = (int64_t)(kOp == 0 ? mulmod(x, y, k) : addmod(x, y, k.m));
  }
}

// Up to 4096 exponent bits per launch; longer exponents continue from the previous launch's output.
struct ExpChunk {
  uint64_t w[64];
  int32_t bits;
};

// out = prev^(2^bits) * base^chunk (prev = 1 on the first chunk): left-to-right square-and-multiply.
template <class Idx>
__global__ void __launch_bounds__(256) k_f64_pow(Operand a, Shape sh, const int64_t* __restrict__ prev,
                                                 int64_t* __restrict__ out, uint64_t n, Mod64 k, ExpChunk e) {
  const Idx step = (Idx)gridDim.x * 256;
  for (Idx i = (Idx)blockIdx.x * 256 + threadIdx.x; i < (Idx)n; i += step) {
    int64_t oa, ob;
    if (sh.ndim == 1) oa = (int64_t)i * a.stride[0];
    else offsets<Idx>(i, sh, a, nullptr, &oa, &ob);
    const uint64_t b = reduce_i64(load_i64(a.p, a.dtype, oa), k.m);
    uint64_t r = prev ? (uint64_t)prev[i] : 1;
    for (int j = e.bits - 1; j >= 0; --j) {
      r = mulmod(r, r, k);
      if ((e.w[j >> 6] >> (j & 63)) & 1) r = mulmod(r, b, k);
    }
    out[i] = (int64_t)r;
  }
}

__global__ void k_clear_flag(int* flag) {
  if (threadIdx.x == 0) *flag = 0;
}

template <class Idx>
__global__ void __launch_bounds__(256) k_f64_any_zero(Operand a, Shape sh, uint64_t n, uint64_t m, int* flag) {
  const Idx step = (Idx)gridDim.x * 256;
  for (Idx i = (Idx)blockIdx.x * 256 + threadIdx.x; i < (Idx)n; i += step) {
    int64_t oa, ob;
    if (sh.ndim == 1) oa = (int64_t)i * a.stride[0];
    else offsets<Idx>(i, sh, a, nullptr, &oa, &ob);
    if (reduce_i64(load_i64(a.p, a.dtype, oa), m) == 0) *flag = 1;  // every writer stores the same 1
  }
}

// ---- NTT -------------------------------------------------------------------------------------

// One pass over digit p of N = n_1 ... n_q: R = n_p points per sub-transform, P = n_1 ... n_(p-1)
// rows already split off, S = N / (P R) columns left. For row r and column J the pass computes
//   Y_p[(r + P k) S + J] = w_N^(P J k) * sum_j w_R^(j k) Y_(p-1)[r R S + j S + J],   k < R,
// the input of the next pass's row r + P k; the last pass (S = 1) writes X[r + P k], natural order.
// A workgroup takes 2^log_cr consecutive rows by 2^log_cj consecutive columns, so loads and stores
// move runs of consecutive addresses.
struct Ntt64Pass {
  const void* in;
  int64_t* out;
  const uint64_t* stage;  // pairs (w, floor(w 2^64 / m)) for w = w_Rmax^i, i < Rmax / 2
  const uint64_t* lo;     // pairs for w_N^i, i < 2^log_t
  const uint64_t* hi;     // pairs for w_N^(i 2^log_t)
  uint64_t m;
  uint64_t scale, scale_q;  // n^-1 and its Shoup quotient (applied by the last pass of an inverse)
  int32_t in_dtype;         // the first pass widens and reduces its input; later passes read int64
  uint32_t log_r, log_rmax, log_p, log_s, log_t, log_cr, log_cj;
  uint32_t first, last, do_scale;
};

constexpr int kNttThreads = 512;
// Points per sub-transform: 2^10 keeps every tile at least 8 columns wide (64-byte runs) and its
// LDS at 74 KiB, two workgroups per CU.
constexpr int kNttMaxLog = 10;
constexpr int kNttTileLog = 13;    // points per workgroup tile (64 KiB of LDS)
constexpr int kNtt64MaxLogN = 30;  // three passes
// LDS of the largest tile: 2^13 points, one padding word per row, the stage twiddle pairs.
constexpr int kNttMaxLds = ((1 << kNttTileLog) + (1 << kNttTileLog) / 2 + (1 << kNttMaxLog)) * 8;

// T = uint64_t for any modulus below 2^63; T = uint32_t for moduli below 2^31 (the pass then holds
// 32-bit residues and twiddle quotients floor(w 2^32 / m), so a Shoup product is three 32-bit
// instructions instead of about sixteen).
template <class T>
__global__ void __launch_bounds__(kNttThreads) k_ntt64_pass(Ntt64Pass d) {
  extern __shared__ uint64_t lds_words[];
  T* lds = reinterpret_cast<T*>(lds_words);
  const uint32_t R = 1u << d.log_r, CJ = 1u << d.log_cj, CR = 1u << d.log_cr;
  const uint32_t slab = (R << d.log_cj) + 1;  // elements per row of the tile, one of them padding
  T* tw = lds + (size_t)CR * slab;           // this pass's stage twiddles, R / 2 pairs
  const uint64_t P = 1ull << d.log_p;
  const uint64_t blk = blockIdx.x;
  const uint32_t jt_bits = d.log_s - d.log_cj;
  const uint64_t J0 = (blk & ((1ull << jt_bits) - 1)) << d.log_cj;
  const uint64_t r0 = (blk >> jt_bits) << d.log_cr;
  const uint32_t E = R << (d.log_cj + d.log_cr);
  const T m = (T)d.m;

  for (uint32_t i = threadIdx.x; i < R / 2; i += kNttThreads) {
    const uint64_t src = (uint64_t)i << (d.log_rmax - d.log_r);
    tw[2 * i] = (T)d.stage[2 * src];
    tw[2 * i + 1] = (T)d.stage[2 * src + 1];
  }
  // Load: columns fastest, then points, then rows; point j lands at its bit-reversed slot.
  for (uint32_t e = threadIdx.x; e < E; e += kNttThreads) {
    const uint32_t cj = e & (CJ - 1), j = (e >> d.log_cj) & (R - 1), cr = e >> (d.log_cj + d.log_r);
    const uint64_t addr = ((r0 + cr) << (d.log_r + d.log_s)) + ((uint64_t)j << d.log_s) + J0 + cj;
    const uint64_t v = d.first ? reduce_i64(load_i64(d.in, d.in_dtype, addr), d.m) : (uint64_t)((const int64_t*)d.in)[addr];
    const uint32_t jr = d.log_r ? (__brev(j) >> (32 - d.log_r)) : 0;
    lds[cr * slab + (jr << d.log_cj) + cj] = (T)v;
  }
  __syncthreads();
  // Decimation-in-time stages on bit-reversed input give natural order. Stages go in pairs
  // (radix 4: a thread takes points i, i + h, i + 2h, i + 3h through stages lh and lh + 1 in
  // registers, one LDS round trip and one barrier for two stages); an odd count ends radix 2.
  uint32_t lh = 0;
  for (; lh + 1 < d.log_r; lh += 2) {
    const uint32_t h = 1u << lh;
    for (uint32_t q = threadIdx.x; q < E / 4; q += kNttThreads) {
      const uint32_t cj = q & (CJ - 1), t = (q >> d.log_cj) & (R / 4 - 1), cr = q >> (d.log_cj + d.log_r - 2);
      const uint32_t tt = t & (h - 1);
      const uint32_t i = ((t >> lh) << (lh + 2)) + tt;
      const uint32_t w1 = tt << (d.log_r - 1 - lh);        // w_(2h)^tt
      const uint32_t w2 = tt << (d.log_r - 2 - lh);        // w_(4h)^tt
      const uint32_t w3 = (tt + h) << (d.log_r - 2 - lh);  // w_(4h)^(tt + h)
      T* row = lds + cr * slab + cj;
      const T a0 = row[i << d.log_cj], a1 = row[(i + h) << d.log_cj], a2 = row[(i + 2 * h) << d.log_cj],
              a3 = row[(i + 3 * h) << d.log_cj];
      T x = shoup(a1, tw[2 * w1], tw[2 * w1 + 1], m);
      const T b0 = addmod(a0, x, m), b1 = submod(a0, x, m);
      x = shoup(a3, tw[2 * w1], tw[2 * w1 + 1], m);
      const T b2 = addmod(a2, x, m), b3 = submod(a2, x, m);
      x = shoup(b2, tw[2 * w2], tw[2 * w2 + 1], m);
      row[i << d.log_cj] = addmod(b0, x, m);
      row[(i + 2 * h) << d.log_cj] = submod(b0, x, m);
      x = shoup(b3, tw[2 * w3], tw[2 * w3 + 1], m);
      row[(i + h) << d.log_cj] = addmod(b1, x, m);
      row[(i + 3 * h) << d.log_cj] = submod(b1, x, m);
    }
    __syncthreads();
  }
  for (; lh < d.log_r; ++lh) {
    const uint32_t h = 1u << lh;
    for (uint32_t q = threadIdx.x; q < E / 2; q += kNttThreads) {
      const uint32_t cj = q & (CJ - 1), t = (q >> d.log_cj) & (R / 2 - 1), cr = q >> (d.log_cj + d.log_r - 1);
      const uint32_t tt = t & (h - 1);
      const uint32_t i = ((t >> lh) << (lh + 1)) + tt;
      const uint32_t wi = tt << (d.log_r - 1 - lh);  // w_(2h)^tt = w_R^(tt R / 2h)
      T* row = lds + cr * slab + cj;
      const T u = row[i << d.log_cj], v = row[(i + h) << d.log_cj];
      const T x = shoup(v, tw[2 * wi], tw[2 * wi + 1], m);
      row[i << d.log_cj] = addmod(u, x, m);
      row[(i + h) << d.log_cj] = submod(u, x, m);
    }
    __syncthreads();
  }
  // Store: columns fastest when the tile has several (their outputs are adjacent), else rows.
  for (uint32_t e = threadIdx.x; e < E; e += kNttThreads) {
    uint32_t cj, k, cr;
    if (d.log_cj) {
      cj = e & (CJ - 1);
      k = (e >> d.log_cj) & (R - 1);
      cr = e >> (d.log_cj + d.log_r);
    } else {
      cj = 0;
      cr = e & (CR - 1);
      k = (e >> d.log_cr) & (R - 1);
    }
    T v = lds[cr * slab + (k << d.log_cj) + cj];
    if (!d.last) {
      const uint64_t ex = P * (J0 + cj) * k;  // < N
      const uint64_t eh = ex >> d.log_t, el = ex & ((1ull << d.log_t) - 1);
      v = shoup(v, (T)d.hi[2 * eh], (T)d.hi[2 * eh + 1], m);
      v = shoup(v, (T)d.lo[2 * el], (T)d.lo[2 * el + 1], m);
    } else if (d.do_scale) {
      v = shoup(v, (T)d.scale, (T)d.scale_q, m);
    }
    d.out[((r0 + cr + P * k) << d.log_s) + J0 + cj] = (int64_t)v;
  }
}

uint64_t powmod_host(uint64_t b, uint64_t e, const Mod64& k) {
  uint64_t r = 1 % k.m;
  for (b %= k.m; e; e >>= 1, b = mulmod(b, b, k))
    if (e & 1) r = mulmod(r, b, k);
  return r;
}

uint64_t shoup_q(uint64_t w, uint64_t m) { return (uint64_t)(((u128)w << 64) / m); }

struct PassShape {
  uint32_t log_r, log_p, log_s, log_cr, log_cj;
};

std::vector<PassShape> ntt64_passes(int log_n) {
  const int np = log_n == 0 ? 1 : (log_n + kNttMaxLog - 1) / kNttMaxLog;
  std::vector<PassShape> out;
  int done = 0;
  for (int p = 0; p < np; ++p) {
    const int lr = log_n / np + (p < log_n % np ? 1 : 0);
    PassShape s;
    s.log_r = lr;
    s.log_p = done;
    s.log_s = log_n - done - lr;
    int lc = kNttTileLog - lr;
    if (lc < 0) lc = 0;
    if (lc > (int)(s.log_p + s.log_s)) lc = s.log_p + s.log_s;
    s.log_cj = lc < (int)s.log_s ? lc : s.log_s;
    s.log_cr = lc - s.log_cj;
    if (s.log_cr > s.log_p) s.log_cr = s.log_p;
    out.push_back(s);
    done += lr;
  }
  return out;
}

int check_operand_args(int dtype, int ndim, const int64_t* shape) {
  if (dtype < ZK_I8 || dtype > ZK_U64) return fail("unknown integer dtype code");
  if (ndim < 1 || ndim > kMaxDims) return fail("operands must have 1 to 8 dimensions after merging");
  for (int d = 0; d < ndim; ++d)
    if (shape[d] < 0) return fail("negative dimension");
  return 0;
}

int f64_launch_shape(int ndim, const int64_t* shape, Shape* sh, uint64_t* n) {
  sh->ndim = ndim;
  uint64_t total = 1;
  for (int d = 0; d < kMaxDims; ++d) sh->dim[d] = d < ndim ? (uint64_t)shape[d] : 1;
  for (int d = 0; d < ndim; ++d) total *= (uint64_t)shape[d];
  *n = total;
  return 0;
}

unsigned f64_grid(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (unsigned)(g < 65536 ? g : 65536);
}

Operand make_operand(const void* p, int dtype, const int64_t* strides, int ndim) {
  Operand o{p, {}, dtype};
  for (int d = 0; d < ndim; ++d) o.stride[d] = strides[d];
  return o;
}

int check_mod(uint64_t m) {
  if (m < 2 || m >= (1ull << 63)) return fail("the modulus must satisfy 2 <= mod < 2^63");
  return 0;
}

}  // namespace
}  // namespace zk

using namespace zk;

extern "C" {

int zk_f64_binary(int op, const void* a, int a_dtype, const int64_t* a_strides, const void* b, int b_dtype,
                  const int64_t* b_strides, int ndim, const int64_t* shape, int64_t* out, uint64_t mod, void* stream) {
  if (int rc = check_mod(mod)) return rc;
  if (op != 0 && op != 1) return fail("zk_f64_binary: op must be 0 (mul) or 1 (add)");
  if (int rc = check_operand_args(a_dtype, ndim, shape)) return rc;
  if (int rc = check_operand_args(b_dtype, ndim, shape)) return rc;
  Shape sh;
  uint64_t n;
  f64_launch_shape(ndim, shape, &sh, &n);
  if (n == 0) return 0;
  const Operand oa = make_operand(a, a_dtype, a_strides, ndim), ob = make_operand(b, b_dtype, b_strides, ndim);
  const Mod64 k = make_mod(mod);
  const hipStream_t st = (hipStream_t)stream;
  const bool small = n < (1ull << 32);
  if (op == 0) {
    if (small) hipLaunchKernelGGL((k_f64_binary<0, uint32_t>), dim3(f64_grid(n)), dim3(256), 0, st, oa, ob, sh, out, n, k);
    else hipLaunchKernelGGL((k_f64_binary<0, uint64_t>), dim3(f64_grid(n)), dim3(256), 0, st, oa, ob, sh, out, n, k);
  } else {
    if (small) hipLaunchKernelGGL((k_f64_binary<1, uint32_t>), dim3(f64_grid(n)), dim3(256), 0, st, oa, ob, sh, out, n, k);
    else hipLaunchKernelGGL((k_f64_binary<1, uint64_t>), dim3(f64_grid(n)), dim3(256), 0, st, oa, ob, sh, out, n, k);
  }
  ZK_LAUNCH_CHECK("64-bit field kernel launch");
  return 0;
}

int zk_f64_pow(const void* base, int dtype, const int64_t* strides, int ndim, const int64_t* shape, const uint64_t* exp_words,
               int exp_len, int64_t* out, uint64_t mod, void* stream) {
  if (int rc = check_mod(mod)) return rc;
  if (int rc = check_operand_args(dtype, ndim, shape)) return rc;
  if (exp_len < 0) return fail("zk_f64_pow: negative exponent length");
  Shape sh;
  uint64_t n;
  f64_launch_shape(ndim, shape, &sh, &n);
  if (n == 0) return 0;
  const Operand oa = make_operand(base, dtype, strides, ndim);
  const Mod64 k = make_mod(mod);
  const hipStream_t st = (hipStream_t)stream;
  int top = exp_len;  // words up to the highest set bit
  while (top > 0 && exp_words[top - 1] == 0) --top;
  const int bits = top == 0 ? 0 : 64 * (top - 1) + (64 - __builtin_clzll(exp_words[top - 1]));
  const bool small = n < (1ull << 32);
  // Chunks of at most 4096 bits, most significant first; the first launch also covers exp = 0.
  int hi = bits;
  const int64_t* prev = nullptr;
  do {
    const int lo = hi > 4096 ? hi - 4096 : 0;
    ExpChunk e{};
    e.bits = hi - lo;
    for (int j = 0; j < e.bits; ++j) {
      const int src = lo + j;
      if ((exp_words[src >> 6] >> (src & 63)) & 1) e.w[j >> 6] |= 1ull << (j & 63);
    }
    if (small) hipLaunchKernelGGL((k_f64_pow<uint32_t>), dim3(f64_grid(n)), dim3(256), 0, st, oa, sh, prev, out, n, k, e);
    else hipLaunchKernelGGL((k_f64_pow<uint64_t>), dim3(f64_grid(n)), dim3(256), 0, st, oa, sh, prev, out, n, k, e);
    ZK_LAUNCH_CHECK("64-bit power launch");
    prev = out;  // each thread reads and writes only its own element
    hi = lo;
  } while (hi > 0);
  return 0;
}

int zk_f64_any_zero(const void* base, int dtype, const int64_t* strides, int ndim, const int64_t* shape, uint64_t mod,
                    int* result, void* stream) {
  if (int rc = check_mod(mod)) return rc;
  if (int rc = check_operand_args(dtype, ndim, shape)) return rc;
  *result = 0;
  Shape sh;
  uint64_t n;
  f64_launch_shape(ndim, shape, &sh, &n);
  if (n == 0) return 0;
  const hipStream_t st = (hipStream_t)stream;
  Scratch sc(st);
  int* flag = sc.get<int>(1);
  if (!flag) return 1;
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, st, flag);
  const Operand oa = make_operand(base, dtype, strides, ndim);
  if (n < (1ull << 32)) hipLaunchKernelGGL((k_f64_any_zero<uint32_t>), dim3(f64_grid(n)), dim3(256), 0, st, oa, sh, n, mod, flag);
  else hipLaunchKernelGGL((k_f64_any_zero<uint64_t>), dim3(f64_grid(n)), dim3(256), 0, st, oa, sh, n, mod, flag);
  ZK_LAUNCH_CHECK("64-bit zero test launch");
  ZK_HIP(hipMemcpyAsync(result, flag, sizeof(int), hipMemcpyDeviceToHost, st));
  ZK_HIP(hipStreamSynchronize(st));
  return 0;
}

int zk_ntt64(const void* in, int dtype, int64_t* out, int log_n, uint64_t mod, uint64_t omega, int invert, void* stream) {
  if (mod < 3 || mod >= (1ull << 63) || (mod & 1) == 0) return fail("zk_ntt64: the modulus must be an odd prime below 2^63");
  if (log_n < 0 || log_n > kNtt64MaxLogN) return fail("zk_ntt64: log_n must be in [0, 30]");
  if (dtype < ZK_I8 || dtype > ZK_U64) return fail("unknown integer dtype code");
  if (in == (const void*)out) return fail("zk_ntt64: out must not be the input");
  if (omega == 0 || omega >= mod) return fail("zk_ntt64: omega must be a nonzero residue");
  const Mod64 k = make_mod(mod);
  const uint64_t n = 1ull << log_n;
  if (powmod_host(omega, n, k) != 1 || (log_n > 0 && powmod_host(omega, n / 2, k) == 1))
    return fail("zk_ntt64: omega is not a primitive 2^log_n-th root of unity");
  const hipStream_t st = (hipStream_t)stream;
  const std::vector<PassShape> passes = ntt64_passes(log_n);
  const uint32_t log_rmax = passes[0].log_r, log_t = (log_n + 1) / 2;
  const uint64_t rmax_half = log_rmax ? (1ull << (log_rmax - 1)) : 1;
  const uint64_t nlo = 1ull << log_t, nhi = 1ull << (log_n - log_t);
  const size_t words = 2 * (rmax_half + nlo + nhi);
  const bool small = mod < (1ull << 31);  // 32-bit residues and Shoup quotients
  const uint64_t tag[4] = {mod, omega, (uint64_t)(invert ? 1 : 0), small ? 0x32u : 0x64u};
  auto quotient = [&](uint64_t v) { return small ? (uint64_t)(((u128)v << 32) / mod) : shoup_q(v, mod); };
  TableHold hold;
  const uint64_t* tab = nullptr;
  int rc = host_table(log_n, tag, words, [&](uint64_t* h) {
    const uint64_t w = invert ? powmod_host(omega, mod - 2, k) : omega;
    auto put = [&](uint64_t* dst, uint64_t v) {
      dst[0] = v;
      dst[1] = quotient(v);
    };
    const uint64_t wr = powmod_host(w, 1ull << (log_n - log_rmax), k);  // w_Rmax
    uint64_t acc = 1;
    for (uint64_t i = 0; i < rmax_half; ++i, acc = mulmod(acc, wr, k)) put(h + 2 * i, acc);
    uint64_t* lo = h + 2 * rmax_half;
    acc = 1;
    for (uint64_t i = 0; i < nlo; ++i, acc = mulmod(acc, w, k)) put(lo + 2 * i, acc);
    const uint64_t wt = acc;  // w^(2^log_t)
    uint64_t* hi = lo + 2 * nlo;
    acc = 1;
    for (uint64_t i = 0; i < nhi; ++i, acc = mulmod(acc, wt, k)) put(hi + 2 * i, acc);
  }, st, &tab);
  if (rc) return rc;
  const uint64_t* lo = tab + 2 * rmax_half;
  const uint64_t* hi = lo + 2 * nlo;
  const uint64_t ninv = powmod_host(n % mod, mod - 2, k);
  Scratch sc(st);
  int64_t* tmp = nullptr;
  if (passes.size() > 1) {
    tmp = sc.get<int64_t>(n);
    if (!tmp) return 1;
  }
  const int np = (int)passes.size();
  static const hipError_t attr =
      hipFuncSetAttribute((const void*)k_ntt64_pass<uint64_t>, hipFuncAttributeMaxDynamicSharedMemorySize, kNttMaxLds);
  ZK_HIP(attr);
  static const hipError_t attr32 =
      hipFuncSetAttribute((const void*)k_ntt64_pass<uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, kNttMaxLds);
  ZK_HIP(attr32);
  const void* src = in;
  for (int p = 0; p < np; ++p) {
    const PassShape& s = passes[p];
    int64_t* dst = ((np - 1 - p) % 2 == 0) ? out : tmp;
    Ntt64Pass d{};
    d.in = src;
    d.out = dst;
    d.stage = tab;
    d.lo = lo;
    d.hi = hi;
    d.m = mod;
    d.scale = ninv;
    d.scale_q = quotient(ninv);
    d.in_dtype = p == 0 ? dtype : ZK_I64;
    d.log_r = s.log_r;
    d.log_rmax = log_rmax;
    d.log_p = s.log_p;
    d.log_s = s.log_s;
    d.log_t = log_t;
    d.log_cr = s.log_cr;
    d.log_cj = s.log_cj;
    d.first = p == 0;
    d.last = p == np - 1;
    d.do_scale = invert && p == np - 1;
    const uint64_t tile = (1ull << (s.log_r + s.log_cj + s.log_cr));
    const unsigned blocks = (unsigned)(n / tile);
    const size_t lds = ((size_t)(1u << s.log_cr) * (((size_t)1 << (s.log_r + s.log_cj)) + 1) + (1u << s.log_r)) *
                       (small ? 4 : 8);
    if (lds > (size_t)kNttMaxLds) return fail("zk_ntt64: pass tile exceeds its LDS budget");
    if (small) launch_timed("k_ntt64_pass", k_ntt64_pass<uint32_t>, dim3(blocks), dim3(kNttThreads), (uint32_t)lds, st, d);
    else launch_timed("k_ntt64_pass", k_ntt64_pass<uint64_t>, dim3(blocks), dim3(kNttThreads), (uint32_t)lds, st, d);
    ZK_LAUNCH_CHECK("64-bit NTT pass launch");
    src = dst;
  }
  return 0;
}

}  // extern "C"
