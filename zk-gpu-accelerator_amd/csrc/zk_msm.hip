// Pippenger MSM over BN254 G1 (include/zkgpu.h, "Pippenger MSM").
//
// Computes sum_i s_i * P_i with the bucket method, entirely on the device:
//
//  1. prep     points -> 261-form coordinates packed in 80-byte records.
//  2. digits   every scalar (reduced mod r) is recoded into W signed digits of
//              c bits, d in [-2^(c-1), 2^(c-1)], W = floor(254 / c) + 1; a digit
//              d != 0 sends +-P_i to bucket (window, |d|). Signed digits halve the
//              buckets per window against the reference's unsigned ones
//              (src/zkgpu/msm_gpu.py:89-105).
//  3. sort     counting sort of the (window, bucket) keys: counts, an exclusive
//              scan and a scatter. Witness-like scalars put a large share of all
//              entries into the few smallest digits (booleans make bucket 1 of
//              window 0 hold ~15% of them), so those keys are counted and ranked
//              in LDS first and reach global atomics once per workgroup.
//  4. buckets  every thread takes exactly K consecutive sorted entries and adds
//              them with mixed Jacobian + affine additions, writing each bucket
//              that lies wholly inside its range. A bucket cut by a range
//              boundary leaves partial sums, at most two per thread, in a slot
//              array that stays sorted by key; the same segmented pass over the
//              slots (full Jacobian additions) repeats until no bucket is split.
//              The load is balanced whatever the bucket sizes, with no host sync
//              (the reference splits buckets into chunks level by level and syncs
//              the host per level, msm_gpu.py:287-333).
//  5. reduce   per window, sum_v v * B_v: threads take segments of L buckets with
//              running sums (T = sum (i+1) B, R = sum B); a window's total is
//              sum_s T_s + L * sum_j 2^j Q_j with Q_j = sum of R_s over segments s
//              whose index has bit j set, each Q_j a block-wide tree sum.
//  6. combine  one Horner pass over all bit positions of all windows gives the
//              Jacobian result on the device (the reference combines windows on
//              the CPU, msm_gpu.py:449 and msm.py:156-169).
//
// The result is a unique group element, so it equals the reference's
// msm_pippenger for every input; intermediate coordinates never leave the device.
#include <algorithm>
#include <vector>

#include "../../include/zkgpu.h"
#include "zk_curve.hpp"
#include "zk_fields.hpp"
#include "zk_msm.hpp"
#include "zk_scan.hpp"

namespace zk {

using Fq = FqConst;
constexpr uint32_t kInvalid = 0x80000000u;   // slot flag: no value
constexpr uint32_t kKeyMask = 0x7fffffffu;
constexpr int kHot = 32;                      // digits 1..kHot are aggregated in LDS

// ---- packed records -------------------------------------------------------

struct alignas(16) AffRec {
  uint32_t w[20];  // x limbs 0..8, y limbs 9..17
};
struct alignas(16) JacRec {
  uint32_t w[28];  // x 0..8, y 9..17, z 18..26
};

__device__ __forceinline__ Aff load_aff_rec(const AffRec* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint32_t w[20];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    uint4 v = q[i];
    w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
  }
  Aff a;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    a.x.l[i] = w[i];
    a.y.l[i] = w[NL + i];
  }
  return a;
}

__device__ __forceinline__ void store_aff_rec(AffRec* p, const Aff& a) {
  uint32_t w[20];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    w[i] = a.x.l[i];
    w[NL + i] = a.y.l[i];
  }
  w[18] = 0;
  w[19] = 0;
  uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
  for (int i = 0; i < 5; ++i) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

__device__ __forceinline__ Jac load_jac_rec(const JacRec* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint32_t w[28];
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    uint4 v = q[i];
    w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
  }
  Jac a;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    a.x.l[i] = w[i];
    a.y.l[i] = w[NL + i];
    a.z.l[i] = w[2 * NL + i];
  }
  return a;
}

__device__ __forceinline__ void store_jac_rec(JacRec* p, const Jac& a) {
  uint32_t w[28];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    w[i] = a.x.l[i];
    w[NL + i] = a.y.l[i];
    w[2 * NL + i] = a.z.l[i];
  }
  w[27] = 0;
  uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
  for (int i = 0; i < 7; ++i) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// ---- scalars and digits -----------------------------------------------------

// r, little-endian 32-bit words
__constant__ uint32_t kR32[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                                 0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};

__device__ __forceinline__ void load_scalar(const uint64_t* s, uint32_t w[8]) {
  const uint4* q = reinterpret_cast<const uint4*>(s);
  uint4 a = q[0], b = q[1];
  w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w; w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
  // reduce mod r: any 256-bit value is below 6r
  for (int rep = 0; rep < 6; ++rep) {
    uint32_t d[8], borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      uint64_t t = (uint64_t)w[i] - kR32[i] - borrow;
      d[i] = (uint32_t)t;
      borrow = (uint32_t)(t >> 63);
    }
    if (borrow) break;  // w < r
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = d[i];
  }
}

__device__ __forceinline__ uint32_t raw_window(const uint32_t w[8], int bit, int c) {
  const int word = bit >> 5, sh = bit & 31;
  if (word >= 8) return 0;
  uint64_t v = w[word];
  if (word + 1 < 8) v |= (uint64_t)w[word + 1] << 32;
  return (uint32_t)(v >> sh) & ((1u << c) - 1u);
}

// Signed digit of window `win` given the carry in; updates the carry.
__device__ __forceinline__ int signed_digit(const uint32_t w[8], int win, int c, uint32_t& carry) {
  uint32_t raw = raw_window(w, win * c, c) + carry;
  const uint32_t half = 1u << (c - 1);
  if (raw > half) {
    carry = 1;
    return (int)raw - (int)(1u << c);
  }
  carry = 0;
  return (int)raw;
}

// ---- kernels ----------------------------------------------------------------

__global__ void __launch_bounds__(256) k_prep_points(const uint64_t* __restrict__ pts, int64_t n, AffRec* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Fq f;
  store_aff_rec(out + i, load_aff(f, pts + 8 * i));
}

// mode 0: count keys; mode 1: scatter (key, ref) into sorted position via cursors.
template <int mode>
__global__ void __launch_bounds__(256) k_digits(const uint64_t* __restrict__ scalars, int64_t n, int c, int windows,
                                                uint32_t* __restrict__ counts_or_cursor, uint32_t* __restrict__ skey,
                                                uint32_t* __restrict__ sref) {
  __shared__ uint32_t hot_cnt[kHot];
  __shared__ uint32_t hot_base[kHot];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < n;
  uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (active) load_scalar(scalars + 4 * i, w);
  const uint32_t buckets = 1u << (c - 1);
  const int hot = (int)min((uint32_t)kHot, buckets);
  uint32_t carry = 0;
  for (int win = 0; win < windows; ++win) {
    if (threadIdx.x < kHot) hot_cnt[threadIdx.x] = 0;
    __syncthreads();
    int d = active ? signed_digit(w, win, c, carry) : 0;
    uint32_t mag = (uint32_t)(d < 0 ? -d : d);
    uint32_t key = (uint32_t)win * buckets + (mag - 1);
    bool is_hot = mag != 0 && (int)mag <= hot;
    uint32_t local = 0;
    if (is_hot) local = atomicAdd(&hot_cnt[mag - 1], 1u);
    if (mode == 0 && mag != 0 && !is_hot) atomicAdd(&counts_or_cursor[key], 1u);
    __syncthreads();
    if (threadIdx.x < hot && hot_cnt[threadIdx.x]) {
      uint32_t k = (uint32_t)win * buckets + threadIdx.x;
      uint32_t base = atomicAdd(&counts_or_cursor[k], hot_cnt[threadIdx.x]);
      hot_base[threadIdx.x] = base;
    }
    __syncthreads();
    if (mode == 1 && mag != 0) {
      uint32_t pos = is_hot ? hot_base[mag - 1] + local : atomicAdd(&counts_or_cursor[key], 1u);
      skey[pos] = key;
      sref[pos] = (uint32_t)i | (d < 0 ? 0x80000000u : 0u);
    }
    __syncthreads();
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_fill_zero(T* __restrict__ p, int64_t words) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) p[i] = 0;
}

// Level sizes: m_{l+1} = 2 * ceil(m_l / K_l).
__device__ __forceinline__ int64_t level_size(uint32_t m0, int level, int k0, int k1) {
  int64_t m = m0;
  for (int l = 0; l < level; ++l) {
    int k = l == 0 ? k0 : k1;
    m = 2 * ((m + k - 1) / k);
  }
  return m;
}

// Level 0: K consecutive sorted entries per thread, mixed additions.
__global__ void __launch_bounds__(256) k_accumulate(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sref,
                                                    const uint32_t* __restrict__ m_dev, int k0,
                                                    const AffRec* __restrict__ pts, JacRec* __restrict__ buckets,
                                                    uint32_t* __restrict__ slot_key, JacRec* __restrict__ slot_val) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t m = *m_dev;
  const int64_t s = t * k0;
  if (s >= m) return;
  const int64_t e = min(s + (int64_t)k0, m);
  const Fq f;
  const uint32_t first_key = skey[s], last_key = skey[e - 1];
  const bool head_partial = s > 0 && skey[s - 1] == first_key;
  const bool tail_partial = e < m && skey[e] == last_key;
  slot_key[2 * t] = first_key | kInvalid;
  slot_key[2 * t + 1] = last_key | kInvalid;
  Jac acc = jac_identity(f);
  uint32_t cur = first_key;
  bool first_seg = true;
  for (int64_t x = s; x <= e; ++x) {
    const uint32_t key = x < e ? skey[x] : 0xffffffffu;
    if (key != cur) {
      // flush segment `cur`
      const bool is_last = x == e;
      if (first_seg && head_partial) {
        slot_key[2 * t] = cur;
        store_jac_rec(slot_val + 2 * t, acc);
      } else if (is_last && tail_partial) {
        slot_key[2 * t + 1] = cur;
        store_jac_rec(slot_val + 2 * t + 1, acc);
      } else {
        store_jac_rec(buckets + cur, acc);
      }
      if (is_last) break;
      first_seg = false;
      cur = key;
      acc = jac_identity(f);
    }
    const uint32_t ref = sref[x];
    Aff q = load_aff_rec(pts + (ref & 0x7fffffffu));
    if (ref >> 31) q.y = fneg(f, q.y);
    acc = jac_madd(f, acc, q);
  }
}

// Levels >= 1: K consecutive slots per thread, full additions; invalid slots add nothing.
__global__ void __launch_bounds__(256) k_combine(const uint32_t* __restrict__ in_key, const JacRec* __restrict__ in_val,
                                                 const uint32_t* __restrict__ m_dev, int level, int k0, int k1,
                                                 JacRec* __restrict__ buckets, uint32_t* __restrict__ out_key,
                                                 JacRec* __restrict__ out_val) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t m = level_size(*m_dev, level, k0, k1);
  const int64_t s = t * k1;
  if (s >= m) return;
  const int64_t e = min(s + (int64_t)k1, m);
  const Fq f;
  const uint32_t first_key = in_key[s] & kKeyMask, last_key = in_key[e - 1] & kKeyMask;
  const bool head_partial = s > 0 && (in_key[s - 1] & kKeyMask) == first_key;
  const bool tail_partial = e < m && (in_key[e] & kKeyMask) == last_key;
  out_key[2 * t] = first_key | kInvalid;
  out_key[2 * t + 1] = last_key | kInvalid;
  Jac acc = jac_identity(f);
  bool has = false;
  uint32_t cur = first_key;
  bool first_seg = true;
  for (int64_t x = s; x <= e; ++x) {
    const uint32_t raw = x < e ? in_key[x] : 0xffffffffu;
    const uint32_t key = x < e ? (raw & kKeyMask) : 0xffffffffu;
    if (key != cur) {
      const bool is_last = x == e;
      if (first_seg && head_partial) {
        out_key[2 * t] = cur | (has ? 0u : kInvalid);
        if (has) store_jac_rec(out_val + 2 * t, acc);
      } else if (is_last && tail_partial) {
        out_key[2 * t + 1] = cur | (has ? 0u : kInvalid);
        if (has) store_jac_rec(out_val + 2 * t + 1, acc);
      } else if (has) {
        store_jac_rec(buckets + cur, acc);
      }
      if (is_last) break;
      first_seg = false;
      cur = key;
      acc = jac_identity(f);
      has = false;
    }
    if (!(raw & kInvalid)) {
      acc = jac_add(f, acc, load_jac_rec(in_val + x));
      has = true;
    }
  }
}

// Leaf of the bucket reduction: segment s of window w covers buckets [sL, sL + L)
// (bucket b holds the digit value b + 1): T = sum (i + 1) B_{sL+i}, R = sum B_{sL+i}.
__global__ void __launch_bounds__(256) k_reduce_leaf(const JacRec* __restrict__ buckets, int64_t segments, int seg_len,
                                                     JacRec* __restrict__ tot, JacRec* __restrict__ run) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // global segment index (window-major)
  if (g >= segments) return;
  const Fq f;
  Jac r = jac_identity(f), t = jac_identity(f);
  for (int i = seg_len - 1; i >= 0; --i) {
    r = jac_add(f, r, load_jac_rec(buckets + g * seg_len + i));
    t = jac_add(f, t, r);
  }
  store_jac_rec(tot + g, t);
  store_jac_rec(run + g, r);
}

// One workgroup per (window, j): j < log2(S) gives Q_j = sum of run[s] over s with bit j set;
// j = log2(S) gives sum of tot[s].
__global__ void __launch_bounds__(256) k_reduce_bits(const JacRec* __restrict__ tot, const JacRec* __restrict__ run,
                                                     int64_t segs_per_window, int log_segs, JacRec* __restrict__ out) {
  __shared__ JacRec sh[256];
  const int w = blockIdx.x / (log_segs + 1), j = blockIdx.x % (log_segs + 1);
  const Fq f;
  Jac acc = jac_identity(f);
  const JacRec* src = (j == log_segs ? tot : run) + (int64_t)w * segs_per_window;
  for (int64_t s = threadIdx.x; s < segs_per_window; s += 256) {
    if (j == log_segs || ((s >> j) & 1)) acc = jac_add(f, acc, load_jac_rec(src + s));
  }
  store_jac_rec(&sh[threadIdx.x], acc);
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      Jac a = load_jac_rec(&sh[threadIdx.x]);
      a = jac_add(f, a, load_jac_rec(&sh[threadIdx.x + off]));
      store_jac_rec(&sh[threadIdx.x], a);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// Horner over every bit position: window w contributes sum_tot at 2^(cw) and Q_j at 2^(cw + log2 L + j).
__global__ void k_final(const JacRec* __restrict__ sums, int windows, int c, int log_seg, int log_segs,
                        uint64_t* __restrict__ out_jac) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Fq f;
  Jac acc = jac_identity(f);
  for (int w = windows - 1; w >= 0; --w) {
    const JacRec* ws = sums + (int64_t)w * (log_segs + 1);
    for (int e = c - 1; e >= 0; --e) {
      acc = jac_dbl(f, acc);
      const int j = e - log_seg;
      if (j >= 0 && j < log_segs) acc = jac_add(f, acc, load_jac_rec(ws + j));
      if (e == 0) acc = jac_add(f, acc, load_jac_rec(ws + log_segs));
    }
  }
  store_jac(f, out_jac, acc);
}

// ---- host ---------------------------------------------------------------------

int default_window(int64_t n) {
  // Signed digits: about n additions per window plus ~2 * 2^(c-1) for the reduction.
  int best = 1;
  double best_cost = 1e300;
  for (int c = 1; c <= 20; ++c) {
    const int windows = 254 / c + 1;
    double cost = (double)windows * ((double)n + 3.0 * (double)(1 << (c - 1)));
    if (cost < best_cost) {
      best_cost = cost;
      best = c;
    }
  }
  return best;
}

int msm_g1(const uint64_t* points, const uint64_t* scalars, int64_t n, int c, int k0, uint64_t* out_jac,
           hipStream_t stream) {
  if (n < 0 || n > 0x7fffffff) return fail("MSM size must be in [0, 2^31)");
  if (c <= 0) c = default_window(n);
  if (c < 1 || c > 20) return fail("window must be in 1..20");
  if (k0 <= 0) k0 = 16;
  if (k0 < 2) return fail("chunk_len must be at least 2");
  const int k1 = std::max(k0, 4);
  const int windows = 254 / c + 1;
  const int64_t buckets_per_window = 1LL << (c - 1);
  const int64_t nkeys = (int64_t)windows * buckets_per_window;
  const int64_t max_entries = (int64_t)windows * n;
  Scratch sc(stream);

  // Level plan (upper bounds) and sizes
  // m_l upper bounds: a level with more entries than one thread takes feeds another level.
  std::vector<int64_t> level_max;
  level_max.push_back(max_entries);
  while (level_max.back() > (level_max.size() == 1 ? k0 : k1)) {
    const int k = level_max.size() == 1 ? k0 : k1;
    level_max.push_back(2 * ((level_max.back() + k - 1) / k));
    if (level_max.size() > 64) return fail("internal: MSM level plan did not converge");
  }
  int64_t max_slots = 2;
  for (size_t l = 1; l < level_max.size(); ++l) max_slots = std::max(max_slots, level_max[l]);

  AffRec* pts = sc.get<AffRec>((size_t)std::max<int64_t>(n, 1));
  uint32_t* counts = sc.get<uint32_t>((size_t)nkeys + 1);
  uint32_t* offsets = sc.get<uint32_t>((size_t)nkeys + 1);
  uint32_t* m_dev = sc.get<uint32_t>(1);
  uint32_t* skey = sc.get<uint32_t>((size_t)std::max<int64_t>(max_entries, 1));
  uint32_t* sref = sc.get<uint32_t>((size_t)std::max<int64_t>(max_entries, 1));
  JacRec* buckets = sc.get<JacRec>((size_t)nkeys);
  uint32_t* slot_key[2] = {sc.get<uint32_t>((size_t)max_slots), sc.get<uint32_t>((size_t)max_slots)};
  JacRec* slot_val[2] = {sc.get<JacRec>((size_t)max_slots), sc.get<JacRec>((size_t)max_slots)};
  if (!pts || !counts || !offsets || !m_dev || !skey || !sref || !buckets || !slot_key[0] || !slot_key[1] || !slot_val[0] ||
      !slot_val[1])
    return 1;

  const unsigned nb = grid_for(std::max<int64_t>(n, 1), 256);
  if (n > 0) {
    hipLaunchKernelGGL(k_prep_points, dim3(nb), dim3(256), 0, stream, points, n, pts);
    ZK_LAUNCH_CHECK("msm prep");
  }
  hipLaunchKernelGGL(k_fill_zero<uint32_t>, dim3(grid_for(nkeys, 256)), dim3(256), 0, stream, counts, nkeys);
  hipLaunchKernelGGL(k_fill_zero<uint32_t>, dim3(grid_for(nkeys * 28, 256)), dim3(256), 0, stream,
                     reinterpret_cast<uint32_t*>(buckets), nkeys * 28);
  ZK_LAUNCH_CHECK("msm clear");
  if (n > 0) {
    hipLaunchKernelGGL(k_digits<0>, dim3(nb), dim3(256), 0, stream, scalars, n, c, windows, counts, nullptr, nullptr);
    ZK_LAUNCH_CHECK("msm count");
  }
  if (int rc = scan_u32(counts, offsets, nkeys, m_dev, stream, sc)) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(k_digits<1>, dim3(nb), dim3(256), 0, stream, scalars, n, c, windows, offsets, skey, sref);
    ZK_LAUNCH_CHECK("msm scatter");
  }
  // Level 0 and the combine levels
  {
    const int64_t threads0 = (max_entries + k0 - 1) / k0;
    if (threads0 > 0) {
      hipLaunchKernelGGL(k_accumulate, dim3(grid_for(threads0, 256)), dim3(256), 0, stream, skey, sref, m_dev, k0, pts,
                         buckets, slot_key[0], slot_val[0]);
      ZK_LAUNCH_CHECK("msm accumulate");
    }
    int cur = 0;
    for (size_t l = 1; l < level_max.size(); ++l) {
      const int64_t threads = (level_max[l] + k1 - 1) / k1;
      hipLaunchKernelGGL(k_combine, dim3(grid_for(threads, 256)), dim3(256), 0, stream, slot_key[cur], slot_val[cur], m_dev,
                         (int)l, k0, k1, buckets, slot_key[cur ^ 1], slot_val[cur ^ 1]);
      ZK_LAUNCH_CHECK("msm combine");
      cur ^= 1;
    }
  }
  // Bucket reduction and window combination
  const int log_seg = std::min(3, c - 1);
  const int seg_len = 1 << log_seg;
  const int log_segs = (c - 1) - log_seg;
  const int64_t segs_per_window = 1LL << log_segs;
  const int64_t segments = (int64_t)windows * segs_per_window;
  JacRec* tot = sc.get<JacRec>((size_t)segments);
  JacRec* run = sc.get<JacRec>((size_t)segments);
  JacRec* sums = sc.get<JacRec>((size_t)windows * (log_segs + 1));
  if (!tot || !run || !sums) return 1;
  hipLaunchKernelGGL(k_reduce_leaf, dim3(grid_for(segments, 256)), dim3(256), 0, stream, buckets, segments, seg_len, tot, run);
  ZK_LAUNCH_CHECK("msm reduce leaf");
  hipLaunchKernelGGL(k_reduce_bits, dim3((unsigned)(windows * (log_segs + 1))), dim3(256), 0, stream, tot, run,
                     segs_per_window, log_segs, sums);
  ZK_LAUNCH_CHECK("msm reduce bits");
  hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, stream, sums, windows, c, log_seg, log_segs, out_jac);
  ZK_LAUNCH_CHECK("msm final");
  return 0;
}

}  // namespace zk

using namespace zk;

extern "C" {

int zk_msm_g1(const uint64_t* points_affine, const uint64_t* scalars, int64_t n, int window, int chunk_len,
              uint64_t* out_jac, void* stream) {
  return msm_g1(points_affine, scalars, n, window, chunk_len, out_jac, (hipStream_t)stream);
}

}  // extern "C"
