// Pippenger MSM over BN254 G1 (include/zkgpu.h, "Pippenger MSM").
//
// Computes sum_i s_i * P_i with the bucket method, entirely on the device:
//
//  1. prep     points -> 64-byte records (two 32-byte rows of 261-form coordinates),
//              plus phi(P_i) = (beta x_i, y_i) = lambda P_i, the curve's endomorphism.
//  2. digits   every scalar (reduced mod r) is split k = k1 + k2 lambda with
//              |k1|, |k2| < 2^127 (GLV: a lattice basis of {(a, b): a + b lambda
//              = 0 mod r}, zk_constants.hpp), and |k1| (for P_i) and |k2| (for
//              phi(P_i)) are recoded into W = floor(127 / c) + 1 signed digits of
//              c bits, d in [-(2^(c-1) - 1), 2^(c-1)] with the signs of k1, k2 folded
//              in, stored as 16-bit codes for c <= 16. A digit d != 0 sends +-P to
//              bucket (window, |d|). Against the
//              reference's unsigned digits of the full scalar (src/zkgpu/msm_gpu.py
//              :89-105) this halves the buckets per window and the windows, so the
//              bucket reduction and the final combination are a quarter and a half
//              as long; the entries grow from W254 n to 2 W127 n (about +12%).
//  3. sort     two counting sorts with no global atomics, so that writes stay local.
//              Pass A sorts the digits by partition (a window's buckets in groups of
//              512): LDS histograms per (window, tile of scalars), per-partition
//              offsets, and placement of packed (key, point index | sign) records
//              from LDS cursors, a batch of digits ordered by partition in LDS before
//              its records leave in runs. Pass B sorts each partition by
//              bucket in chunks of 8192 records, staged and sorted in LDS and written
//              one bucket run at a time. Witness-like scalars put a large share of all
//              entries into a few buckets (booleans send ~15% of them to bucket 1 of
//              window 0); those hot keys only ever meet LDS atomics.
//  4. buckets  a persistent grid splits the m sorted entries evenly, ceil(m / threads)
//              consecutive entries per thread, added to an XYZZ accumulator with
//              affine points (madd-2008-s: 10 products instead of Jacobian's 11);
//              a run of equal keys strictly inside a thread's range is a whole
//              bucket and is written at once, in the accumulator's coordinates (no
//              conversion inside the loop: some lane of a wave ends a run in most
//              iterations, so a conversion there would cost the whole wave 4 products
//              per entry). The runs cut by range boundaries leave as one run summary
//              per thread; boundary passes join the fragments of short chains, and
//              trees of 64 or 256 summaries per workgroup merge the rest, level after
//              level, each merge at most one addition, each level writing the runs
//              that end inside its workgroups. The load is balanced whatever the
//              bucket sizes, with no host sync (the reference splits buckets into
//              chunks level by level and syncs the host per level,
//              msm_gpu.py:287-333).
//  5. reduce   per window, sum_v v * B_v: threads take segments of L buckets (2 or 4,
//              read in XYZZ coordinates) with running sums (T = sum (i+1) B,
//              R = sum B); a window's total is sum_s T_s + L * sum_j 2^j Q_j with
//              Q_j = sum of R_s over segments s whose index has bit j set, each Q_j
//              (and sum T_s, in two halves) a block-wide tree sum with the same number
//              of additions per block.
//  6. combine  per window, the terms 2^(j + log L) Q_j are doubled side by side and
//              summed; the windows then meet in sum_w 2^(cw) U_w, Horner over four
//              groups of windows side by side, all with the limb-sliced formulas of
//              zk_curve_sliced.hpp, and the Jacobian result stays on the device (the
//              reference combines windows on the CPU, msm_gpu.py:449 and
//              msm.py:156-169).
//
// Steps 1 and 2 and the clearing of the buckets share the plan's first launch (k_prepare).
//
// The result is a unique group element, so it equals the reference's
// msm_pippenger for every input; intermediate coordinates never leave the device.
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/zkgpu.h"
#include "zk_curve.hpp"
#include "zk_curve_sliced.hpp"
#include "zk_curve_wave.hpp"
#include "zk_curve_xyzz.hpp"
#include "zk_fields.hpp"
#include "zk_msm.hpp"
#include "zk_scalar.hpp"
#include "zk_scan.hpp"

namespace zk {

// The base field for this file's kernels, with unmasked quotient digits in the thread-serial
// products (zk_field.hpp, FreeQuotient). Every product operand here is a coordinate or product
// below 2p with normalized limbs, a sum of two of them (limbs below 2^30), a p4r difference
// (limbs below 3 * 2^29) against a normalized operand, the two-product sum, or an lc_carry
// result of a combination below 18p (top limb below 2^27): none takes an a b column part above
// 9 * 2^60. q's limbs sum below 2^30.6, so the m p part of a column stays below 2^62.6.
struct FqMsm : FqConst {
  static constexpr bool kFreeQuotient = true;
};
struct FqMsmIlp : FqIlp {
  static constexpr bool kFreeQuotient = true;
};
using Fq = FqMsm;

// ---- packed records -------------------------------------------------------

// A finite affine point as two 32-byte rows (the 261-form coordinates are canonical, below
// 2^256): 64 bytes, so a gather touches half a cache line.
struct alignas(64) AffRec {
  uint64_t x[4], y[4];
};
struct alignas(16) JacRec {
  uint32_t w[28];  // x 0..8, y 9..17, z 18..26 (29-bit limbs)
};

// An XYZZ point (x, y, zz, zzz: nine 29-bit limbs each), 144 bytes.
struct alignas(16) XyzzRec {
  uint32_t w[36];
};

__device__ __forceinline__ Xyzz load_xyzz_rec(const XyzzRec* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint32_t w[36];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const uint4 v = q[i];
    w[4 * i] = v.x, w[4 * i + 1] = v.y, w[4 * i + 2] = v.z, w[4 * i + 3] = v.w;
  }
  Xyzz a;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    a.x.l[i] = w[i];
    a.y.l[i] = w[NL + i];
    a.zz.l[i] = w[2 * NL + i];
    a.zzz.l[i] = w[3 * NL + i];
  }
  return a;
}

__device__ __forceinline__ void store_xyzz_rec(XyzzRec* p, const Xyzz& a) {
  uint4* q = reinterpret_cast<uint4*>(p);
  uint32_t w[36];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    w[i] = a.x.l[i];
    w[NL + i] = a.y.l[i];
    w[2 * NL + i] = a.zz.l[i];
    w[3 * NL + i] = a.zzz.l[i];
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// The raw words of an AffRec, loaded ahead of their use.
struct AffRaw {
  uint4 a, b, c, d;
};

__device__ __forceinline__ AffRaw load_aff_raw(const AffRec* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  return AffRaw{q[0], q[1], q[2], q[3]};
}

__device__ __forceinline__ Aff unpack_aff(const AffRaw& r) {
  const uint32_t wx[8] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y, r.b.z, r.b.w};
  const uint32_t wy[8] = {r.c.x, r.c.y, r.c.z, r.c.w, r.d.x, r.d.y, r.d.z, r.d.w};
  Aff a;
  a.x = from_words(wx);
  a.y = from_words(wy);
  return a;
}

__device__ __forceinline__ void store_aff_rec(AffRec* p, const Aff& a) {
  store_row(p->x, a.x);
  store_row(p->y, a.y);
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

// A bucket sum: XYZZ coordinates (x, y, zz, zzz: 29-bit limbs, words 0..35) as the accumulation
// leaves them, or a Jacobian point (x, y, z in words 0..26) marked by bit 31 of word 35 (the top
// limb of a coordinate below 2p never has it). All-zero words are the XYZZ identity (zz = 0), so a
// cleared bucket that receives no entries reads as the identity.
struct alignas(16) BucketRec {
  uint32_t w[36];
};
constexpr uint32_t kJacMark = 0x80000000u;

__device__ __forceinline__ void store_bucket_words(BucketRec* p, const uint32_t (&w)[36]) {
  uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
  for (int i = 0; i < 9; ++i) q[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

__device__ __forceinline__ void store_bucket(BucketRec* p, const Xyzz& a) {
  uint32_t w[36];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    w[i] = a.x.l[i];
    w[NL + i] = a.y.l[i];
    w[2 * NL + i] = a.zz.l[i];
    w[3 * NL + i] = a.zzz.l[i];
  }
  store_bucket_words(p, w);
}

__device__ __forceinline__ void store_bucket(BucketRec* p, const Jac& a) {
  uint32_t w[36];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    w[i] = a.x.l[i];
    w[NL + i] = a.y.l[i];
    w[2 * NL + i] = a.z.l[i];
  }
#pragma unroll
  for (int i = 3 * NL; i < 35; ++i) w[i] = 0;
  w[35] = kJacMark;
  store_bucket_words(p, w);
}

// The bucket in XYZZ coordinates (Jacobian records converted: 2 products).
template <class F>
__device__ __forceinline__ Xyzz load_bucket_xyzz(const F& f, const BucketRec* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint32_t w[36];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const uint4 v = q[i];
    w[4 * i] = v.x, w[4 * i + 1] = v.y, w[4 * i + 2] = v.z, w[4 * i + 3] = v.w;
  }
  Xyzz a;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    a.x.l[i] = w[i];
    a.y.l[i] = w[NL + i];
    a.zz.l[i] = w[2 * NL + i];
    a.zzz.l[i] = w[3 * NL + i];
  }
  if (w[35] & kJacMark) {
    Jac j;
    j.x = a.x;
    j.y = a.y;
    j.z = a.zz;
    return jac_to_xyzz(f, j);
  }
  return a;
}

// The bucket as a Jacobian point (XYZZ sums converted: 4 products).
template <class F>
__device__ __forceinline__ Jac load_bucket(const F& f, const BucketRec* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint32_t w[36];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const uint4 v = q[i];
    w[4 * i] = v.x, w[4 * i + 1] = v.y, w[4 * i + 2] = v.z, w[4 * i + 3] = v.w;
  }
  Xyzz a;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    a.x.l[i] = w[i];
    a.y.l[i] = w[NL + i];
    a.zz.l[i] = w[2 * NL + i];
    a.zzz.l[i] = w[3 * NL + i];
  }
  if (w[35] & kJacMark) {
    Jac j;
    j.x = a.x;
    j.y = a.y;
    j.z = a.zz;
    return j;
  }
  return xyzz_to_jac(f, a);
}

// ---- scalars and digits -----------------------------------------------------

// ---- kernels ----------------------------------------------------------------

// Point i (a 256-form row pair) -> its 261-form record; with phi, also phi(P_i) = (beta x_i, y_i).
__device__ __forceinline__ void prep_point(const uint64_t* __restrict__ pts, int64_t i, AffRec* __restrict__ out,
                                           AffRec* __restrict__ phi) {
  const Fq f;
  Aff a = load_aff(f, pts + 8 * i);
  store_aff_rec(out + i, a);
  if (phi) {
    a.x = mont(f, a.x, fe_const<Fq>(Glv::beta261));
    store_aff_rec(phi + i, a);
  }
}

__global__ void __launch_bounds__(256) k_prep_points(const uint64_t* __restrict__ pts, int64_t n, AffRec* __restrict__ out,
                                                     AffRec* __restrict__ phi) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) prep_point(pts, i, out, phi);
}

// Per-MSM parameters of a batch (see MsmPlan).
struct MsmJob {
  const uint64_t* scalars;  // (n, 4) plain limbs
  int64_t n;
  int64_t block_base;       // first 256-scalar digit block of this job
  int64_t key_base;         // first bucket key
  int64_t seg_base;         // first reduction segment
  int64_t part_base;        // first (window, bit, chunk) partial of the bucket reduction
  int64_t win_base;         // first per-window sum
  int64_t dig_base;         // first signed digit (window-major: window w at dig_base + w * n)
  int64_t sort_base;        // first sort block (window, scalar tile)
  int64_t cnt_base;         // first (partition, tile) count
  int64_t pkey_base;        // first partition key (a partition: 2^pbits consecutive buckets of a window)
  int64_t tile;             // scalars per sort tile
  uint64_t* out;            // 12 limbs, the Jacobian result
  int64_t nv;               // digit streams: n, or 2n with the GLV split (k1 for P_i, k2 for phi(P_i))
  int64_t phi_base;         // record index of phi(P_0) in the point array
  int64_t table_base;       // with a window table: record index of 2^(c w) P_i at table_base + w * stride + i
  int64_t table_stride;
  int c, windows, log_seg, log_segs, chunks, tiles, ranges, glv;
  int pbits;                // bucket bits inside a partition: min(c - 1, kPartBits)
  int table;                // 1: every window's digits share one set of buckets (the SRS's window table)
  int rw;                   // windows the bucket reduction and the final Horner chain see (1 with a table)
};

// Largest j with base(j) <= x for a field of the job table (bases increase).
template <int64_t MsmJob::*base>
__device__ __forceinline__ int find_job(const MsmJob* jobs, int njobs, int64_t x) {
  int lo = 0, hi = njobs;  // jobs[lo].base <= x < jobs[hi].base
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (jobs[mid].*base <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// A signed digit d as the sort reads it: (|d| - 1) with the sign in the top bit, every bit set
// for d = 0. No nonzero digit's code has every bit set: below 16-bit windows |d| - 1 < 2^14, and
// for 16-bit windows the stored digits (a GLV half's sign folded in, its ties chosen for that) lie
// in [-(2^15 - 1), 2^15]. So 16-bit codes hold windows of up to 16 bits (a plan with wider ones
// stores 32-bit codes): the digit array is written once and read twice, and at 2^20 points 16-bit
// codes move 100 MB less.
template <class D>
__device__ __forceinline__ D digit_code(int d) {
  if (d == 0) return (D)~(D)0;
  const uint32_t mag = (uint32_t)(d < 0 ? -d : d) - 1u;
  return (D)(mag | (d < 0 ? 1u << (8 * sizeof(D) - 1) : 0u));
}
template <class D>
__device__ __forceinline__ bool digit_any(D v) { return v != (D)~(D)0; }
template <class D>
__device__ __forceinline__ int64_t digit_bucket(D v) { return (int64_t)((uint32_t)v & ((1u << (8 * sizeof(D) - 1)) - 1u)); }
template <class D>
__device__ __forceinline__ bool digit_neg(D v) { return (v >> (8 * sizeof(D) - 1)) & 1; }

// Signed digits of every scalar: digits[dig_base + w * nv + v] = the code of d, v = i for the
// scalar itself, or with the GLV split v = i for k1 (point P_i) and v = n + i for k2 (phi(P_i)),
// the sign of k1 / k2 folded into the digits. Block b of the digit grid (256 scalars).
template <class D>
__device__ __forceinline__ void signed_digits_block(const MsmJob* __restrict__ jobs, int njobs, int64_t b,
                                                    D* __restrict__ digits) {
  const MsmJob& job = jobs[find_job<&MsmJob::block_base>(jobs, njobs, b)];
  const int64_t i = (b - job.block_base) * 256 + threadIdx.x;
  if (i >= job.n) return;
  uint32_t w[8];
  {
    const uint4* q = reinterpret_cast<const uint4*>(job.scalars + 4 * i);
    const uint4 a = q[0], b = q[1];
    w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w; w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
  }
  reduce_mod_r(w);
  D* d = digits + job.dig_base;
  if (!job.glv) {
    uint32_t carry = 0;
    for (int win = 0; win < job.windows; ++win) d[(int64_t)win * job.nv + i] = digit_code<D>(signed_digit(w, win, job.c, carry));
    return;
  }
  uint32_t m1[8], m2[8];
  bool neg1, neg2;
  glv_split(w, m1, neg1, m2, neg2);
  uint32_t c1 = 0, c2 = 0;
  for (int win = 0; win < job.windows; ++win) {
    // a negative half's ties go negative below the top window, so that its folded digits stay in
    // [-(2^(c-1) - 1), 2^(c-1)] like the positive half's wherever c = 16 needs it (the top window
    // of a half below 2^126.97 stays below 2^15 there; narrower windows have room in the codes)
    const bool top = win + 1 == job.windows;
    const int d1 = signed_digit(m1, win, job.c, c1, neg1 && !top), d2 = signed_digit(m2, win, job.c, c2, neg2 && !top);
    d[(int64_t)win * job.nv + i] = digit_code<D>(neg1 ? -d1 : d1);
    d[(int64_t)win * job.nv + job.n + i] = digit_code<D>(neg2 ? -d2 : d2);
  }
}

// The first launch of a plan, three independent jobs in one grid: the signed digits (blocks below
// `digit_blocks`), the bucket records cleared to zero (one 32-bit word per thread), and, for a plan
// over raw points, the point records (one point per thread). Grid-stride, so any grid covers them.
template <class D>
__global__ void __launch_bounds__(256) k_prepare(const MsmJob* __restrict__ jobs, int njobs, int64_t digit_blocks,
                                                 D* __restrict__ digits, uint32_t* __restrict__ clear,
                                                 int64_t clear_words, const uint64_t* __restrict__ pts, int64_t npts,
                                                 AffRec* __restrict__ out, AffRec* __restrict__ phi) {
  const int64_t threads = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < npts; i += threads) prep_point(pts, i, out, phi);
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < clear_words; i += threads) clear[i] = 0;
  for (int64_t b = blockIdx.x; b < digit_blocks; b += gridDim.x) signed_digits_block(jobs, njobs, b, digits);
}

// The sort runs in two passes so that its writes stay local. Pass A is a counting sort of the
// digits by partition (a window's buckets in groups of 2^kPartBits): a workgroup's tile of digits
// goes to a few dozen partitions, so its records leave in runs of hundreds. Pass B sorts each
// partition by bucket in chunks of kChunkB records, writing inside the partition's own region.
// (One counting sort by bucket would scatter single 8-byte records over the whole array: each
// costs a memory transaction of its own.)
constexpr int kPartBits = 9;              // buckets per partition: 512
constexpr int kPartMax = 1 << 11;         // partitions per window at most (c <= 20)
constexpr int kChunkB = 8192;             // records per pass-B workgroup
// Digits or records a thread loads before it updates the LDS counters or cursors: the loads of one
// batch are in flight together (one load at a time left these passes waiting on memory latency).
constexpr int kSortBatch = 8;
constexpr int kPlaceThreads = 512;        // pass-B placement workgroup (two per CU, by its LDS)

// Sort block geometry: (window, scalar tile) of one job.
struct SortBlock {
  const MsmJob* job;
  int win, tile;
};

__device__ __forceinline__ SortBlock sort_block(const MsmJob* __restrict__ jobs, int njobs) {
  const MsmJob* job = &jobs[find_job<&MsmJob::sort_base>(jobs, njobs, blockIdx.x)];
  const int64_t local = (int64_t)blockIdx.x - job->sort_base;
  SortBlock b;
  b.job = job;
  b.tile = (int)(local % job->tiles);
  b.win = (int)(local / job->tiles);
  return b;
}

__device__ __forceinline__ int parts_per_window(const MsmJob& job) { return 1 << (job.c - 1 - job.pbits); }

// Pass A, counts: per (partition, tile), the nonzero digits of one tile of one window. Counts are
// laid out tile-major (window w, tile t: a row of the window's partitions).
template <class D>
__global__ void __launch_bounds__(256) k_sort_count(const MsmJob* __restrict__ jobs, int njobs,
                                                    const D* __restrict__ digits, uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[kPartMax];
  const SortBlock sb = sort_block(jobs, njobs);
  const MsmJob& job = *sb.job;
  const int np = parts_per_window(job);
  for (int k = threadIdx.x; k < np; k += 256) hist[k] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)sb.tile * job.tile, hi = min(job.nv, lo + job.tile);
  const D* d = digits + job.dig_base + (int64_t)sb.win * job.nv;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += 256 * kSortBatch) {
    D v[kSortBatch];
#pragma unroll
    for (int u = 0; u < kSortBatch; ++u) v[u] = i0 + 256 * u < hi ? d[i0 + 256 * u] : (D)~(D)0;
#pragma unroll
    for (int u = 0; u < kSortBatch; ++u)
      if (digit_any(v[u])) atomicAdd(&hist[digit_bucket(v[u]) >> job.pbits], 1u);
  }
  __syncthreads();
  uint32_t* out = counts + job.cnt_base + ((int64_t)sb.win * job.tiles + sb.tile) * np;
  for (int k = threadIdx.x; k < np; k += 256) out[k] = hist[k];
}

// One wavefront per partition key: turns the partition's per-tile counts into per-tile offsets
// inside the partition (an exclusive scan down the tiles, 64 tiles per step across the lanes) and
// writes its total, whose scan over all partitions gives every partition its first sorted position.
__global__ void __launch_bounds__(256) k_sort_columns(const MsmJob* __restrict__ jobs, int njobs, int64_t nparts,
                                                      uint32_t* __restrict__ counts, uint32_t* __restrict__ totals) {
  const int64_t key = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (key >= nparts) return;
  const MsmJob& job = jobs[find_job<&MsmJob::pkey_base>(jobs, njobs, key)];
  const int np = parts_per_window(job);
  const int64_t local = key - job.pkey_base;
  const int64_t win = local / np, q = local % np;
  // a partition's rows: the tiles of its window, or of every window when they share the buckets
  uint32_t* col = counts + job.cnt_base + (job.table ? 0 : win * job.tiles * np) + q;
  const int rows = job.table ? job.windows * job.tiles : job.tiles;
  uint32_t run = 0;  // the same in every lane
  for (int t0 = 0; t0 < rows; t0 += 64) {
    const int t = t0 + lane;
    const uint32_t v = t < rows ? col[(int64_t)t * np] : 0u;
    uint32_t incl = v;  // inclusive scan across the lanes
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(incl, d, 64);
      if (lane >= d) incl += o;
    }
    if (t < rows) col[(int64_t)t * np] = run + incl - v;
    run += __shfl(incl, 63, 64);
  }
  if (lane == 0) totals[key] = run;
}

// Pass A, placement: each digit's (bucket key, point ref) record goes to its partition's next
// free position (LDS cursors per partition, no global atomics). A batch of 256 x kSortBatch digits
// is first ordered by partition in LDS (a histogram, its scan, each record at its partition's
// offset plus its rank), so that a partition's records of the batch leave as one run of
// consecutive positions instead of as single 8-byte stores.
template <class D>
__global__ void __launch_bounds__(256) k_sort_place(const MsmJob* __restrict__ jobs, int njobs,
                                                    const D* __restrict__ digits, const uint32_t* __restrict__ counts,
                                                    const uint32_t* __restrict__ starts, uint2* __restrict__ staged) {
  constexpr int kBatch = 256 * kSortBatch;
  __shared__ uint32_t cur[kPartMax];  // the partition's next position in staged
  __shared__ uint32_t at[kPartMax];   // the batch's count per partition, then its first slot in buf
  __shared__ uint2 buf[kBatch];
  __shared__ uint32_t wsum[4];
  const SortBlock sb = sort_block(jobs, njobs);
  const MsmJob& job = *sb.job;
  const int np = parts_per_window(job);
  const int64_t buckets = 1LL << (job.c - 1);
  const uint32_t* off = counts + job.cnt_base + ((int64_t)sb.win * job.tiles + sb.tile) * np;
  const int64_t pkey0 = job.pkey_base + (job.table ? 0 : (int64_t)sb.win * np);
  const int64_t key0 = job.key_base + (job.table ? 0 : (int64_t)sb.win * buckets);
  for (int k = threadIdx.x; k < np; k += 256) cur[k] = starts[pkey0 + k] + off[k];
  const int64_t lo = (int64_t)sb.tile * job.tile, hi = min(job.nv, lo + job.tile);
  const D* d = digits + job.dig_base + (int64_t)sb.win * job.nv;
  const int per = (np + 255) / 256;  // partitions per thread in the scan (at most kPartMax / 256)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t b0 = lo; b0 < hi; b0 += kBatch) {
    for (int k = threadIdx.x; k < np; k += 256) at[k] = 0;
    D v[kSortBatch];
#pragma unroll
    for (int u = 0; u < kSortBatch; ++u) {
      const int64_t i = b0 + threadIdx.x + 256 * u;
      v[u] = i < hi ? d[i] : (D)~(D)0;
    }
    __syncthreads();
    uint32_t rank[kSortBatch];
#pragma unroll
    for (int u = 0; u < kSortBatch; ++u)
      if (digit_any(v[u])) rank[u] = atomicAdd(&at[digit_bucket(v[u]) >> job.pbits], 1u);
    __syncthreads();
    // exclusive scan of at[0, np): per partitions per thread, then across the lanes and the waves
    uint32_t mine = 0;
    for (int e = 0; e < per; ++e) {
      const int k = threadIdx.x * per + e;
      mine += k < np ? at[k] : 0u;
    }
    uint32_t incl = mine;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
      const uint32_t o = __shfl_up(incl, s, 64);
      if (lane >= s) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t run = incl - mine;
    for (int w = 0; w < wave; ++w) run += wsum[w];
    const uint32_t total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    for (int e = 0; e < per; ++e) {
      const int k = threadIdx.x * per + e;
      if (k < np) {
        const uint32_t c = at[k];
        at[k] = run;
        run += c;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kSortBatch; ++u) {
      if (!digit_any(v[u])) continue;
      const int64_t i = b0 + threadIdx.x + 256 * u;
      const int64_t b = digit_bucket(v[u]);
      // the point record: 2^(c w) P_i from the window table, P_i, or phi(P_i) for the k2 stream
      const int64_t rec = job.table ? job.table_base + (int64_t)sb.win * job.table_stride + i
                                    : (i < job.n ? i : job.phi_base + (i - job.n));
      buf[at[b >> job.pbits] + rank[u]] = make_uint2((uint32_t)(key0 + b), (uint32_t)rec | (digit_neg(v[u]) ? 0x80000000u : 0u));
    }
    __syncthreads();
    // buf is in partition order: slot x of partition k goes to cur[k] plus its rank in the batch
    for (int x = threadIdx.x; x < (int)total; x += 256) {
      const uint2 r = buf[x];
      const int k = (int)(((int64_t)r.x - key0) >> job.pbits);
      staged[cur[k] + ((uint32_t)x - at[k])] = r;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < np; k += 256) cur[k] += (k + 1 < np ? at[k + 1] : total) - at[k];
    __syncthreads();
  }
}

// Pass B geometry: chunks of kChunkB records per partition, numbered partition after partition.
__global__ void __launch_bounds__(256) k_part_chunks(const uint32_t* __restrict__ totals, int64_t nparts,
                                                     uint32_t* __restrict__ chunks) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nparts) chunks[p] = (totals[p] + kChunkB - 1) / kChunkB;
}

// The sort's offsets in one workgroup when the partitions are few (a single MSM has a few hundred):
// starts = the exclusive scan of the partition totals and *m their sum, chunk_base = the exclusive
// scan of the partitions' pass-B chunk counts and chunk_base[nparts] their sum. The same as two
// device-wide scans and k_part_chunks, in one launch instead of five.
constexpr int kOffThreads = 1024;
constexpr int64_t kOffMaxParts = 16 * kOffThreads;

__global__ void __launch_bounds__(kOffThreads) k_part_offsets(const uint32_t* __restrict__ totals, int64_t nparts,
                                                              uint32_t* __restrict__ starts, uint32_t* __restrict__ m,
                                                              uint32_t* __restrict__ chunk_base) {
  __shared__ uint32_t sa[kOffThreads], sb[kOffThreads];
  uint32_t carry_a = 0, carry_b = 0;  // the same in every thread
  for (int64_t base = 0; base < nparts; base += kOffThreads) {
    const int64_t i = base + threadIdx.x;
    const uint32_t v = i < nparts ? totals[i] : 0u;
    const uint32_t c = (v + kChunkB - 1) / kChunkB;
    sa[threadIdx.x] = v;
    sb[threadIdx.x] = c;
    __syncthreads();
    for (int d = 1; d < kOffThreads; d <<= 1) {  // inclusive scans
      const uint32_t xa = (int)threadIdx.x >= d ? sa[threadIdx.x - d] : 0u;
      const uint32_t xb = (int)threadIdx.x >= d ? sb[threadIdx.x - d] : 0u;
      __syncthreads();
      sa[threadIdx.x] += xa;
      sb[threadIdx.x] += xb;
      __syncthreads();
    }
    if (i < nparts) {
      starts[i] = carry_a + sa[threadIdx.x] - v;
      chunk_base[i] = carry_b + sb[threadIdx.x] - c;
    }
    carry_a += sa[kOffThreads - 1];
    carry_b += sb[kOffThreads - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *m = carry_a;
    chunk_base[nparts] = carry_b;
  }
}

// The (partition, chunk) of pass-B workgroup g, or false past the last chunk.
struct PartChunk {
  int64_t p, lo, hi, bkey0;
  int pbits;
};

__device__ __forceinline__ bool part_chunk(const MsmJob* __restrict__ jobs, int njobs, const uint32_t* __restrict__ chunk_base,
                                           int64_t nparts, const uint32_t* __restrict__ starts,
                                           const uint32_t* __restrict__ totals, int64_t g, PartChunk* pc) {
  if (g >= (int64_t)chunk_base[nparts]) return false;
  int64_t lo = 0, hi = nparts;  // chunk_base[lo] <= g < chunk_base[hi]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)chunk_base[mid] <= g) lo = mid;
    else hi = mid;
  }
  const MsmJob& job = jobs[find_job<&MsmJob::pkey_base>(jobs, njobs, lo)];
  pc->p = lo;
  pc->pbits = job.pbits;
  pc->bkey0 = job.key_base + ((lo - job.pkey_base) << job.pbits);
  const int64_t first = (int64_t)starts[lo] + (g - (int64_t)chunk_base[lo]) * kChunkB;
  pc->lo = first;
  pc->hi = min(first + (int64_t)kChunkB, (int64_t)starts[lo] + (int64_t)totals[lo]);
  return true;
}

// atomicAdd(&ctr[k], 1) for every active lane, returning the old value: the lanes whose key equals
// the first active lane's are served by one atomic of their count (their ranks follow the lane
// order), the others by one each. A run of one bucket, as a hot bucket's chunks are, would
// otherwise queue 64 atomics on one LDS word per wavefront instruction.
__device__ __forceinline__ uint32_t lds_claim(uint32_t* ctr, uint32_t k) {
  const uint32_t k0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)k);
  const uint64_t same = __ballot(k == k0);
  uint32_t pos;
  if (k == k0) {
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(same >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)same, 0u));
    uint32_t base = 0;
    if (rank == 0) base = atomicAdd(&ctr[k0], (uint32_t)__popcll(same));
    pos = (uint32_t)__builtin_amdgcn_readlane((int)base, __ffsll((unsigned long long)same) - 1) + rank;
  } else {
    pos = atomicAdd(&ctr[k], 1u);
  }
  return pos;
}

// Pass B, counts: per (chunk, bucket of its partition).
__global__ void __launch_bounds__(256) k_part_count(const MsmJob* __restrict__ jobs, int njobs,
                                                    const uint32_t* __restrict__ chunk_base, int64_t nparts,
                                                    const uint32_t* __restrict__ starts, const uint32_t* __restrict__ totals,
                                                    const uint2* __restrict__ staged, uint32_t* __restrict__ counts,
                                                    uint32_t* __restrict__ raw) {
  __shared__ uint32_t hist[1 << kPartBits];
  PartChunk pc;
  if (!part_chunk(jobs, njobs, chunk_base, nparts, starts, totals, blockIdx.x, &pc)) return;
  const int nb = 1 << pc.pbits;
  for (int k = threadIdx.x; k < nb; k += 256) hist[k] = 0;
  __syncthreads();
  for (int64_t x0 = pc.lo + threadIdx.x; x0 < pc.hi; x0 += 256 * kSortBatch) {
    uint32_t k[kSortBatch];
#pragma unroll
    for (int u = 0; u < kSortBatch; ++u) k[u] = x0 + 256 * u < pc.hi ? staged[x0 + 256 * u].x : 0u;
#pragma unroll
    for (int u = 0; u < kSortBatch; ++u)
      if (x0 + 256 * u < pc.hi) lds_claim(hist, (uint32_t)(k[u] - pc.bkey0));
  }
  __syncthreads();
  uint32_t* out = counts + (int64_t)blockIdx.x * (1 << kPartBits);
  uint32_t* keep = raw + (int64_t)blockIdx.x * (1 << kPartBits);
  for (int k = threadIdx.x; k < (1 << kPartBits); k += 256) {
    const uint32_t v = k < nb ? hist[k] : 0u;
    out[k] = v;
    keep[k] = v;  // the chunk's own histogram, for its placement
  }
}

// Pass B, offsets: one workgroup per partition. Per bucket an exclusive scan down the partition's
// chunks, then a scan over the buckets; every (chunk, bucket) count becomes the absolute position
// of the chunk's first record of that bucket.
__global__ void __launch_bounds__(256) k_part_columns(const MsmJob* __restrict__ jobs, int njobs,
                                                      const uint32_t* __restrict__ chunk_base, int64_t nparts,
                                                      const uint32_t* __restrict__ starts, uint32_t* __restrict__ counts) {
  __shared__ uint32_t part[256];
  const int64_t p = blockIdx.x;
  if (p >= nparts) return;
  const MsmJob& job = jobs[find_job<&MsmJob::pkey_base>(jobs, njobs, p)];
  const int nb = 1 << job.pbits;
  const int64_t g0 = chunk_base[p], g1 = chunk_base[p + 1];
  constexpr int kPer = (1 << kPartBits) / 256;  // buckets per thread
  uint32_t run[kPer];
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int j = threadIdx.x * kPer + e;
    run[e] = 0;
    if (j < nb) {
      for (int64_t g = g0; g < g1; g += kSortBatch) {
        uint32_t v[kSortBatch];
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) v[u] = g + u < g1 ? counts[(g + u) * (1 << kPartBits) + j] : 0u;
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) {
          if (g + u < g1) counts[(g + u) * (1 << kPartBits) + j] = run[e];
          run[e] += v[u];
        }
      }
    }
  }
  // exclusive scan of the bucket totals: per thread, then over the 256 threads
  uint32_t sum = 0;
#pragma unroll
  for (int e = 0; e < kPer; ++e) sum += run[e];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const uint32_t v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t base = starts[p] + part[threadIdx.x] - sum;  // exclusive
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int j = threadIdx.x * kPer + e;
    if (j < nb)
      for (int64_t g = g0; g < g1; g += kSortBatch) {
        uint32_t v[kSortBatch];
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) v[u] = g + u < g1 ? counts[(g + u) * (1 << kPartBits) + j] : 0u;
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u)
          if (g + u < g1) counts[(g + u) * (1 << kPartBits) + j] = v[u] + base;
      }
    base += run[e];
  }
}

// Pass B, placement: the chunk's records are sorted by bucket in LDS first (a histogram, its scan,
// and LDS cursors), then leave in that order, so that a bucket's records of the chunk are
// consecutive in memory and written as runs rather than one by one.
__global__ void __launch_bounds__(kPlaceThreads) k_part_place(const MsmJob* __restrict__ jobs, int njobs,
                                                    const uint32_t* __restrict__ chunk_base, int64_t nparts,
                                                    const uint32_t* __restrict__ starts, const uint32_t* __restrict__ totals,
                                                    const uint2* __restrict__ staged, const uint32_t* __restrict__ counts,
                                                    const uint32_t* __restrict__ raw, uint2* __restrict__ sorted) {
  constexpr int kB = 1 << kPartBits;
  __shared__ uint2 buf[kChunkB];
  __shared__ uint32_t first[kB];  // the bucket's first slot in buf
  __shared__ uint32_t cur[kB];
  __shared__ uint32_t part[kPlaceThreads];
  PartChunk pc;
  if (!part_chunk(jobs, njobs, chunk_base, nparts, starts, totals, blockIdx.x, &pc)) return;
  const int count = (int)(pc.hi - pc.lo);
  // exclusive scan of the chunk's histogram (kept by the count pass): kB / kPlaceThreads buckets per thread,
  // then over the threads
  const uint32_t* hist = raw + (int64_t)blockIdx.x * kB;
  constexpr int kPer = kB > kPlaceThreads ? kB / kPlaceThreads : 1;
  uint32_t loc[kPer], sum = 0;
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int j = threadIdx.x * kPer + e;
    loc[e] = j < kB ? hist[j] : 0u;
    sum += loc[e];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int d = 1; d < kPlaceThreads; d <<= 1) {
    const uint32_t v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t at = part[threadIdx.x] - sum;
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int j = threadIdx.x * kPer + e;
    if (j < kB) {
      first[j] = at;
      cur[j] = at;
    }
    at += loc[e];
  }
  __syncthreads();
  for (int x0 = threadIdx.x; x0 < count; x0 += kPlaceThreads * kSortBatch) {
    uint2 r[kSortBatch];
#pragma unroll
    for (int u = 0; u < kSortBatch; ++u) r[u] = x0 + kPlaceThreads * u < count ? staged[pc.lo + x0 + kPlaceThreads * u] : make_uint2(0u, 0u);
#pragma unroll
    for (int u = 0; u < kSortBatch; ++u)
      if (x0 + kPlaceThreads * u < count) buf[lds_claim(cur, (uint32_t)(r[u].x - pc.bkey0))] = r[u];
  }
  __syncthreads();
  // buf is now in bucket order: slot x of bucket j goes to the chunk's position for j plus its rank
  const uint32_t* base = counts + (int64_t)blockIdx.x * kB;
  for (int x = threadIdx.x; x < count; x += kPlaceThreads) {
    const uint2 r = buf[x];
    const uint32_t j = r.x - (uint32_t)pc.bkey0;
    sorted[base[j] + (x - first[j])] = r;
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_fill_zero(T* __restrict__ p, int64_t words) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) p[i] = 0;
}

// Runs of equal bucket keys, summarised per range of sorted entries: the sums of the run that
// touches the range's first entry (fk) and of the run touching its last entry (lk); runs
// strictly inside a range are complete and are written to their bucket immediately.
struct alignas(16) RunSummary {
  uint32_t fk, lk;
  uint32_t flags;  // bit 0: any entries; bit 1: a single run (fk == lk covers the whole range)
  uint32_t pad;
  JacRec fsum, lsum;
};
constexpr uint32_t kValid = 1, kSingle = 2;
static_assert(sizeof(RunSummary) % 16 == 0 && sizeof(JacRec) % 16 == 0, "summaries are copied in 16-byte words");

// Who performs a merge's addition: one thread, the 16 lanes of a row (wave-parallel formulas), or
// a whole wavefront (limb-sliced formulas); the first lane of a row or wavefront does the writes.
enum MergeBy { kByThread, kByRow, kByWave };

// Merges the summary at *l with the one at *r (the range right after it), writing every run that
// becomes bounded on both sides. At most one point addition.
template <MergeBy kBy>
__device__ __forceinline__ void merge_runs(RunSummary* l, const RunSummary* r, BucketRec* __restrict__ buckets) {
  const FqMsmIlp f;
  const bool lead = kBy == kByThread || (threadIdx.x & (kBy == kByRow ? 15 : 63)) == 0;
  const uint32_t lf = l->flags, rf = r->flags;
  if (!(rf & kValid)) return;
  if (!(lf & kValid)) {
    if (lead) *l = *r;
    return;
  }
  const bool ls = lf & kSingle, rs = rf & kSingle;
  const uint32_t lk = ls ? l->fk : l->lk;
  if (lk == r->fk) {
    // the run crossing the boundary
    const Jac a = load_jac_rec(ls ? &l->fsum : &l->lsum), b = load_jac_rec(&r->fsum);
    Jac joined;
    if constexpr (kBy == kByRow) {
      joined = jac_canonical(f, jac_add_wave(f, a, b));
    } else if constexpr (kBy == kByWave) {
      // canonical: the sliced identity leaves with Z = p, and the thread-serial formulas and the
      // bucket loads recognise the identity by Z == 0 exactly
      const SlLane c = sl_lane<FqConst>();
      joined = jac_canonical(f, sl_to_jac<FqConst>(c, sl_add<FqConst>(c, sl_from_jac(a), sl_from_jac(b))));
    } else {
      joined = jac_add_lazy(f, a, b);
    }
    if (!lead) return;
    if (ls && rs) {
      store_jac_rec(&l->fsum, joined);
    } else if (ls) {
      store_jac_rec(&l->fsum, joined);
      l->lk = r->lk;
      l->lsum = r->lsum;
      l->flags = kValid;
    } else if (rs) {
      store_jac_rec(&l->lsum, joined);
    } else {
      store_bucket(buckets + lk, joined);  // bounded by both ranges' interiors
      l->lk = r->lk;
      l->lsum = r->lsum;
    }
    return;
  }
  if (!lead) return;
  // different keys meet at the boundary
  if (!ls) store_bucket(buckets + l->lk, load_jac_rec(&l->lsum));  // l's last run is now closed on both sides
  if (!rs) store_bucket(buckets + r->fk, load_jac_rec(&r->fsum));  // r's first run likewise
  // (a single run of l stays open on the left)
  l->lk = rs ? r->fk : r->lk;
  l->lsum = rs ? r->fsum : r->lsum;
  l->flags = kValid;
}

// Tree over the first `width` (a power of two) summaries of a workgroup in LDS; the result is left
// in sh[0]. Levels with more merges than the workgroup has 16-lane rows run one merge per thread;
// the upper levels, where a thread-serial addition (9 µs) would set the level's latency, one merge
// per row (wave-parallel, 6.7 µs), and the last levels, with no more merges than wavefronts, one
// per wavefront (limb-sliced, 2.7 µs).
__device__ __forceinline__ void merge_tree(RunSummary* sh, int width, BucketRec* __restrict__ buckets) {
  const int rows = (int)blockDim.x >> 4, waves = (int)blockDim.x >> 6;
  for (int d = 1; d < width; d <<= 1) {
    __syncthreads();
    const int merges = width / (2 * d);
    if (merges > rows) {
      // merge i on thread i: the active lanes fill the leading waves, and the rest skip the addition
      const int i = (int)threadIdx.x;
      if (i < merges) merge_runs<kByThread>(&sh[2 * d * i], &sh[2 * d * i + d], buckets);
    } else if (merges > waves) {
      const int row = (int)threadIdx.x >> 4;
      if (row < merges) merge_runs<kByRow>(&sh[2 * d * row], &sh[2 * d * row + d], buckets);
    } else {
      const int wave = (int)threadIdx.x >> 6;
      if (wave < merges) merge_runs<kByWave>(&sh[2 * d * wave], &sh[2 * d * wave + d], buckets);
    }
  }
  __syncthreads();
}

// Writes the open runs of the final summary (they are bounded by the ends of the array).
__device__ __forceinline__ void close_runs(const RunSummary* sm, BucketRec* __restrict__ buckets) {
  if (!(sm->flags & kValid)) return;
  store_bucket(buckets + sm->fk, load_jac_rec(&sm->fsum));
  if (!(sm->flags & kSingle)) store_bucket(buckets + sm->lk, load_jac_rec(&sm->lsum));
}

constexpr int kAccThreads = 64;      // one wave per accumulation workgroup
// Summaries per combine workgroup, chosen per plan (MsmPlan::group): 64 lets the tree's widest
// level run one merge per 16-lane row (32 rows) instead of one per thread, which pays when the
// level's workgroups are few or mostly empty (a hot bucket's chain through long ranges); 256 keeps
// the workgroups few when most buckets reach the trees (ranges shorter than a bucket).
constexpr int kCombineThreads = 256;  // the most summaries per combine workgroup
// Threads of a combine workgroup: the tree's upper levels pick their mode from the workgroup's rows
// and wavefronts, so 512 threads (32 rows, 8 wavefronts) run the level of 32 merges one per row and
// the last four one per wavefront (limb-sliced) instead of one per thread and per row. At 228
// registers a CU holds one such workgroup, but past the boundary pass most workgroups of a level
// hold no summary and end at once, so even a level with more workgroups than CUs waits only on
// the few that merge a hot bucket's pieces (1024 threads would cap the registers at 128, below the
// thread-serial addition's).
constexpr int kCombineLanes = 512;

// Level 0, a persistent grid: the m sorted entries are split evenly, ceil(m / threads) consecutive
// entries per thread; complete runs leave as bucket records in the accumulator's coordinates, the
// range's first and last runs as one Jacobian run summary per thread, merged by k_combine (no tree
// here: its point additions would raise the kernel's registers past three waves per SIMD).
// The running bucket sum of the accumulation, in Jacobian (madd-2007-bl) or XYZZ (madd-2008-s)
// coordinates; F picks the product schedule (serial columns, or FqMsmIlp's independent partial sums).
template <class F, bool kXyzz>
struct BucketSum;

template <class F>
struct BucketSum<F, false> {
  Jac a;
  __device__ __forceinline__ void reset(const F& f) { a = jac_identity(f); }
  __device__ __forceinline__ void add(const F& f, Aff q, bool neg) {
    if (neg) q.y = fneg(f, q.y);
    a = jac_madd(f, a, q);
  }
  __device__ __forceinline__ Jac jac(const F&) const { return a; }
  __device__ __forceinline__ void store(BucketRec* p) const { store_bucket(p, a); }
};

template <class F>
struct BucketSum<F, true> {
  Xyzz a;
  __device__ __forceinline__ void reset(const F& f) { a = xyzz_identity(f); }
  __device__ __forceinline__ void add(const F& f, const Aff& q, bool neg) { a = xyzz_madd(f, a, q, neg); }
  __device__ __forceinline__ Jac jac(const F& f) const { return xyzz_to_jac(f, a); }
  __device__ __forceinline__ void store(BucketRec* p) const { store_bucket(p, a); }
};

template <class F, bool kXyzz>
__global__ void __launch_bounds__(kAccThreads)
k_accumulate(const uint2* __restrict__ sorted, const uint32_t* __restrict__ m_dev,
             const AffRec* __restrict__ pts, BucketRec* __restrict__ buckets, RunSummary* __restrict__ out) {
  const int64_t threads = (int64_t)gridDim.x * kAccThreads;
  const int64_t t = (int64_t)blockIdx.x * kAccThreads + threadIdx.x;
  const int64_t m = *m_dev;
  const int64_t k = (m + threads - 1) / threads;
  const int64_t s = t * k;
  const F f;
  RunSummary me;
  me.flags = 0;
  me.fk = me.lk = 0;
  if (s < m) {
    const int64_t e = min(s + k, m), last = e - 1;
    BucketSum<F, kXyzz> acc;
    acc.reset(f);
    // software pipeline: the point of entry x and the record of entry x + 1 are in flight while
    // entry x - 1 is added (indices past the range are clamped to its last entry, and unused)
    uint2 rec = sorted[s];
    uint2 rec1 = sorted[min(s + 1, last)];
    AffRaw pt = load_aff_raw(pts + (rec.y & 0x7fffffffu));
    uint32_t cur = rec.x;
    bool first = true;
    me.fk = cur;
    for (int64_t x = s; x < e; ++x) {
      const uint2 rec2 = sorted[min(x + 2, last)];
      const AffRaw pt1 = load_aff_raw(pts + (rec1.y & 0x7fffffffu));
      if (rec.x != cur) {
        // a run ends and leaves as it is (no conversion inside the loop, where some lane ends a
        // run in most iterations). The range's first run may continue in the previous range; it
        // is parked in its own bucket, which nothing else writes before the run summaries are
        // merged, and taken back below.
        acc.store(buckets + cur);
        first = false;
        cur = rec.x;
        acc.reset(f);
      }
      acc.add(f, unpack_aff(pt), rec.y >> 31);
      rec = rec1;
      rec1 = rec2;
      pt = pt1;
    }
    if (first) {
      store_jac_rec(&me.fsum, acc.jac(f));
      me.flags = kValid | kSingle;
    } else {
      store_jac_rec(&me.lsum, acc.jac(f));
      store_jac_rec(&me.fsum, load_bucket(f, buckets + me.fk));
      me.lk = cur;
      me.flags = kValid;
    }
  }
  out[t] = me;
}

// The boundaries between consecutive summaries of one tree level, resolved before the next level
// (one thread per summary t; its neighbours' keys decide). Each bucket cut by the boundaries is a
// chain of fragments: the last run of a summary, the runs of summaries that hold nothing else, the
// first run of a summary. A run bounded on both sides is written as it is; a chain of two fragments
// is added up by the thread of its right fragment (one addition) and written; only chains of three
// or more fragments (hot buckets, through whole summaries) stay for the trees, as reduced
// summaries in `red` that hold just those fragments (one open run is marked single). Every other
// reduced summary is empty, so the trees merge nothing else. When a thread's range is longer than
// a typical bucket (the tuned windows make it so), and past the first level in any case, chains of
// three or more come from hot buckets only.
__global__ void __launch_bounds__(256) k_fixup(const RunSummary* __restrict__ in, int64_t count,
                                               BucketRec* __restrict__ buckets, RunSummary* __restrict__ red) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= count) return;
  // keys and flags of the summaries t - 2 .. t + 2 (index 2 is t)
  uint32_t fk[5], ke[5];
  bool valid[5], single[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const int64_t i = t - 2 + k;
    uint4 m = make_uint4(0, 0, 0, 0);
    if (i >= 0 && i < count) m = *reinterpret_cast<const uint4*>(in + i);
    fk[k] = m.x;
    valid[k] = m.z & kValid;
    single[k] = m.z & kSingle;
    ke[k] = single[k] ? m.x : m.y;  // the key of the summary's last run
  }
  bool jl[5], jr[5];  // the summary's first (last) run continues in the previous (next) summary
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    jl[k] = k > 0 && valid[k - 1] && valid[k] && ke[k - 1] == fk[k];
    jr[k] = k < 4 && valid[k] && valid[k + 1] && ke[k] == fk[k + 1];
  }
  RunSummary* o = red + t;
  if (!valid[2]) {
    o->flags = 0;
    return;
  }
  const FqMsmIlp f;
  const RunSummary& me = in[t];
  // What goes on to the trees: nothing, the whole summary, or one of its runs as a single run.
  // (The output is written field by field from registers: a RunSummary assembled on the stack
  // would live in scratch memory.)
  enum { kNone, kWhole, kFirst, kLast } pass = kNone;
  // the one addition a thread may owe: the first fragment of summary t (its single run or its
  // first run) as the right end of a chain of two, added to the fragment that ends summary t - 1;
  // every case below reaches the same call, so lanes that take different cases do not each pay
  // for an addition of their own
  bool add = false;
  if (single[2]) {
    const bool lo = jl[2], hi = jr[2];
    if (!lo && !hi) {
      store_bucket(buckets + fk[2], load_jac_rec(&me.fsum));
    } else if ((lo && hi) || (lo && single[1] && jl[1]) || (hi && single[3] && jr[3])) {
      pass = kFirst;
    } else {
      add = lo;  // else: a chain of two that the next summary's thread adds up
    }
  } else {
    bool first_in_tree = false, last_in_tree = false;
    if (!jl[2]) store_bucket(buckets + fk[2], load_jac_rec(&me.fsum));
    else if (single[1] && jl[1]) first_in_tree = true;
    else add = true;
    if (!jr[2]) store_bucket(buckets + me.lk, load_jac_rec(&me.lsum));
    else if (single[3] && jr[3]) last_in_tree = true;
    pass = first_in_tree && last_in_tree ? kWhole : first_in_tree ? kFirst : last_in_tree ? kLast : kNone;
  }
  if (add) {
    const JacRec* left = single[1] ? &in[t - 1].fsum : &in[t - 1].lsum;
    store_bucket(buckets + fk[2], jac_add_lazy(f, load_jac_rec(left), load_jac_rec(&me.fsum)));
  }
  if (pass == kNone) {
    o->flags = 0;
  } else if (pass == kWhole) {
    const uint4* src = reinterpret_cast<const uint4*>(&me);
    uint4* dst = reinterpret_cast<uint4*>(o);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(RunSummary) / 16); ++i) dst[i] = src[i];
  } else {
    const uint32_t key = pass == kFirst ? fk[2] : me.lk;
    *reinterpret_cast<uint4*>(o) = make_uint4(key, key, kValid | kSingle, 0u);
    const uint4* src = reinterpret_cast<const uint4*>(pass == kFirst ? &me.fsum : &me.lsum);
    uint4* dst = reinterpret_cast<uint4*>(&o->fsum);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(JacRec) / 16); ++i) dst[i] = src[i];
  }
}

// The `count` summaries of the previous level, `group` per workgroup, merged by a tree. The non-empty
// summaries are packed to the front in their order first (an empty summary holds no entries, so
// the ones on either side of it are neighbours), and the tree is as wide as they are many: past the
// boundary passes most summaries are empty, and a workgroup holding a few pieces of a hot bucket's
// chain then merges them in a few levels instead of eight.
__global__ void __launch_bounds__(kCombineLanes) k_combine(const RunSummary* __restrict__ in, int64_t count,
                                                           BucketRec* __restrict__ buckets, RunSummary* __restrict__ out,
                                                           int last_level, int group) {
  __shared__ RunSummary sh[kCombineThreads];
  __shared__ uint32_t wcount[kCombineThreads / 64];
  const int64_t t = (int64_t)blockIdx.x * group + threadIdx.x;
  const bool keep = (int)threadIdx.x < group && t < count && in[t].flags != 0;
  const uint64_t mask = __ballot(keep);
  const int wave = (int)threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0 && wave < kCombineThreads / 64) wcount[wave] = (uint32_t)__popcll(mask);
  __syncthreads();
  int before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kCombineThreads / 64; ++w) {
    before += w < wave ? (int)wcount[w] : 0;
    total += (int)wcount[w];
  }
  if (keep) sh[before + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u))] = in[t];
  int width = 1;
  while (width < total) width <<= 1;
  if ((int)threadIdx.x >= total && (int)threadIdx.x < width) sh[threadIdx.x].flags = 0;
  merge_tree(sh, width, buckets);
  if (threadIdx.x != 0) return;
  RunSummary* s = &sh[0];
  if (last_level) {
    close_runs(s, buckets);
    return;
  }
  // Runs that do not continue past the workgroup's range are complete: written now, so that the
  // next level merges only what continues. A chain's pieces sit at consecutive, non-empty indices,
  // so a run continues only into the summary right outside the range with the same key there.
  if (s->flags & kValid) {
    const int64_t base = (int64_t)blockIdx.x * group;
    const bool single = s->flags & kSingle;
    const uint32_t last = single ? s->fk : s->lk;
    bool cl = false, cr = false;
    if (base > 0) {
      const uint4 m = *reinterpret_cast<const uint4*>(in + base - 1);  // fk, lk, flags of the left neighbour
      cl = (m.z & kValid) && ((m.z & kSingle) ? m.x : m.y) == s->fk;
    }
    if (base + group < count) {
      const uint4 m = *reinterpret_cast<const uint4*>(in + base + group);
      cr = (m.z & kValid) && m.x == last;
    }
    if (single) {
      if (!cl && !cr) {
        store_bucket(buckets + s->fk, load_jac_rec(&s->fsum));
        s->flags = 0;
      }
    } else {
      if (!cl) store_bucket(buckets + s->fk, load_jac_rec(&s->fsum));
      if (!cr) store_bucket(buckets + s->lk, load_jac_rec(&s->lsum));
      if (!cl && !cr) {
        s->flags = 0;
      } else if (!cl) {
        // only the last run goes on, as a single run
        s->fk = s->lk;
        s->fsum = s->lsum;
        s->flags = kValid | kSingle;
      } else if (!cr) {
        s->lk = s->fk;
        s->flags = kValid | kSingle;
      }
    }
  }
  out[blockIdx.x] = *s;
}

// Leaf of the bucket reduction: segment s of a window covers buckets [sL, sL + L)
// (bucket b holds the digit value b + 1): T = sum (i + 1) B_{sL+i}, R = sum B_{sL+i}.
__global__ void __launch_bounds__(256) k_reduce_leaf(const MsmJob* __restrict__ jobs, int njobs, int64_t segments,
                                                     const BucketRec* __restrict__ buckets, XyzzRec* __restrict__ tot,
                                                     XyzzRec* __restrict__ run) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= segments) return;
  const MsmJob& job = jobs[find_job<&MsmJob::seg_base>(jobs, njobs, g)];
  const int64_t local = g - job.seg_base;
  const int seg_len = 1 << job.log_seg;
  const int64_t win = local >> job.log_segs, s = local & ((1LL << job.log_segs) - 1);
  const BucketRec* b = buckets + job.key_base + (win << (job.c - 1)) + s * seg_len;
  const FqMsmIlp f;
  // XYZZ running sums (14 products an addition, no conversion per bucket)
  Xyzz r = xyzz_identity(f), t = xyzz_identity(f);
  for (int i = seg_len - 1; i >= 0; --i) {
    r = xyzz_add_lazy(f, r, load_bucket_xyzz(f, b + i));
    t = xyzz_add_lazy(f, t, r);
  }
  store_xyzz_rec(tot + g, t);
  store_xyzz_rec(run + g, r);
}

// Segments each thread adds serially. The rows' additions are a latency chain (the serial sums,
// then a tree), so the kernel wants about one wavefront per SIMD rather than more threads: with
// 8, a 2^20-point MSM (16-bit windows: 8 windows x 15 rows x 2 chunks) runs 240 workgroups of
// four wavefronts.
constexpr int kRedPerThread = 8;
constexpr int kRedChunk = 256 * kRedPerThread;          // segments per workgroup

// Rows of a window's reduction: Q_j for j < log_segs (the segments whose index has bit j set,
// half of them), then the T sum in two halves (segments below and above segs / 2), so that every
// row has the same number of additions.
__device__ __forceinline__ int red_rows(const MsmJob& job) { return job.log_segs + 2; }

// One workgroup per (job, window, row, chunk): 256 threads x kRedPerThread selected segments, then a tree.
__global__ void __launch_bounds__(256) k_reduce_bits(const MsmJob* __restrict__ jobs, int njobs,
                                                     const XyzzRec* __restrict__ tot, const XyzzRec* __restrict__ run,
                                                     JacRec* __restrict__ parts) {
  __shared__ XyzzRec sh[256];
  __shared__ JacRec shj[16];
  const MsmJob& job = jobs[find_job<&MsmJob::part_base>(jobs, njobs, blockIdx.x)];
  const int64_t local = (int64_t)blockIdx.x - job.part_base;
  const int per = red_rows(job);
  const int64_t row = local / job.chunks;
  const int64_t chunk = local % job.chunks;
  const int64_t win = row / per;
  const int j = (int)(row % per);
  const int64_t segs = 1LL << job.log_segs;
  const FqMsmIlp f;
  Xyzz acc = xyzz_identity(f);
  const int64_t lo = chunk * kRedChunk;
  if (j < job.log_segs) {
    // the k-th segment with bit j set, k in this chunk: whole runs of 2^j segments, so that
    // neighbouring lanes never skip additions the wavefront executes anyway
    const XyzzRec* src = run + job.seg_base + win * segs;
    const int64_t hi = min(segs >> 1, lo + (int64_t)kRedChunk), low = (1LL << j) - 1;
    auto seg = [&](int64_t k) { return ((k >> j) << (j + 1)) + (1LL << j) + (k & low); };
    int64_t k = lo + threadIdx.x;
    if (k < hi) {
      Xyzz next = load_xyzz_rec(src + seg(k));  // the next segment's load overlaps this addition
      for (; k < hi; k += 256) {
        const Xyzz cur = next;
        if (k + 256 < hi) next = load_xyzz_rec(src + seg(k + 256));
        acc = xyzz_add_lazy(f, acc, cur);
      }
    }
  } else {
    // the T sum over one half of the segments
    const int64_t first = (segs + 1) >> 1;
    const int64_t base = j == job.log_segs ? 0 : first, count = j == job.log_segs ? first : segs - first;
    const XyzzRec* src = tot + job.seg_base + win * segs + base;
    const int64_t hi = min(count, lo + (int64_t)kRedChunk);
    int64_t k = lo + threadIdx.x;
    if (k < hi) {
      Xyzz next = load_xyzz_rec(src + k);
      for (; k < hi; k += 256) {
        const Xyzz cur = next;
        if (k + 256 < hi) next = load_xyzz_rec(src + k + 256);
        acc = xyzz_add_lazy(f, acc, cur);
      }
    }
  }
  store_xyzz_rec(&sh[threadIdx.x], acc);
  __syncthreads();
  // a tree over the 256 partial sums: one XYZZ addition per thread down to 16 sums, which become
  // Jacobian points for the last levels on the four wavefronts with the limb-sliced formulas (an
  // addition in about a third of one lane's time)
  for (int off = 128; off > 8; off >>= 1) {
    if ((int)threadIdx.x < off) {
      Xyzz a = load_xyzz_rec(&sh[threadIdx.x]);
      a = xyzz_add_lazy(f, a, load_xyzz_rec(&sh[threadIdx.x + off]));
      store_xyzz_rec(&sh[threadIdx.x], a);
    }
    __syncthreads();
  }
  if (threadIdx.x < 16) store_jac_rec(&shj[threadIdx.x], xyzz_to_jac(f, load_xyzz_rec(&sh[threadIdx.x])));
  __syncthreads();
  const SlLane c = sl_lane<FqConst>();
  const int wave = (int)threadIdx.x >> 6;
  for (int off = 8; off > 0; off >>= 1) {
    for (int i = wave; i < off; i += 4) {
      const SlJac a = sl_add<FqConst>(c, sl_from_jac(load_jac_rec(&shj[i])), sl_from_jac(load_jac_rec(&shj[i + off])));
      const Jac r = sl_to_jac<FqConst>(c, a);
      if ((threadIdx.x & 63) == 0) store_jac_rec(&shj[i], off == 1 ? jac_canonical(f, r) : r);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) parts[blockIdx.x] = shj[0];
}

// Window sums U = sum T_s + sum_j 2^(j + log_seg) Q_j, in two launches. k_window_terms: four
// workgroups of four wavefronts per (job, window), one wavefront per SIMD (the limb-sliced formulas
// keep a SIMD's issue busy with one wavefront, so a second one there would only wait). Wavefront
// slot i of the window takes the terms i, i + 16, ...: a term is the sum of its reduction row's
// chunk partials; term j < log_segs is Q_j doubled j + log_seg times, the last two are the halves
// of sum T_s. The doublings of different terms run side by side (a Horner pass over j would chain
// all of them with the additions); each workgroup sums its wavefronts' terms, and k_window_total
// sums a window's four partials.
constexpr int kWinGroups = 4;
constexpr int kWinWaves = 4;

template <int kWaves>
__device__ __forceinline__ SlJac sl_workgroup_sum(JacRec* sh, const SlLane& c, const SlJac& mine) {
  const int wave = (int)threadIdx.x >> 6;
  const bool lead = (threadIdx.x & 63) == 0;
  const Jac v = sl_to_jac<FqConst>(c, mine);
  if (lead) store_jac_rec(&sh[wave], v);
  for (int d = 1; d < kWaves; d <<= 1) {
    __syncthreads();
    if ((wave & (2 * d - 1)) == 0) {
      const SlJac a = sl_add<FqConst>(c, sl_from_jac(load_jac_rec(&sh[wave])), sl_from_jac(load_jac_rec(&sh[wave + d])));
      const Jac r = sl_to_jac<FqConst>(c, a);
      if (lead) store_jac_rec(&sh[wave], r);
    }
  }
  __syncthreads();
  return sl_from_jac(load_jac_rec(&sh[0]));
}

__global__ void __launch_bounds__(64 * kWinWaves) k_window_terms(const MsmJob* __restrict__ jobs, int njobs,
                                                                 int64_t total_windows, const JacRec* __restrict__ parts,
                                                                 JacRec* __restrict__ wparts) {
  __shared__ JacRec sh[kWinWaves];
  const int64_t g = blockIdx.x / kWinGroups;
  const int group = (int)(blockIdx.x % kWinGroups);
  if (g >= total_windows) return;
  const MsmJob& job = jobs[find_job<&MsmJob::win_base>(jobs, njobs, g)];
  const int64_t win = g - job.win_base;
  const int per = red_rows(job);
  const JacRec* p = parts + job.part_base + win * per * job.chunks;
  const FqConst f;
  const SlLane c = sl_lane<FqConst>();
  SlJac acc = sl_from_jac(jac_identity(f));
  for (int t = group * kWinWaves + ((int)threadIdx.x >> 6); t < per; t += kWinGroups * kWinWaves) {
    SlJac d = sl_from_jac(load_jac_rec(p + (int64_t)t * job.chunks));
    for (int k = 1; k < job.chunks; ++k) d = sl_add<FqConst>(c, d, sl_from_jac(load_jac_rec(p + (int64_t)t * job.chunks + k)));
    const int dbl = t < job.log_segs ? t + job.log_seg : 0;
    for (int e = 0; e < dbl; ++e) d = sl_dbl<FqConst>(c, d);
    acc = sl_add<FqConst>(c, acc, d);
  }
  acc = sl_workgroup_sum<kWinWaves>(sh, c, acc);
  const Jac r = sl_to_jac<FqConst>(c, acc);
  if (threadIdx.x == 0) store_jac_rec(wparts + blockIdx.x, r);
}

// One workgroup per (job, window): the sum of the window's kWinGroups partials.
__global__ void __launch_bounds__(64 * kWinGroups) k_window_total(int64_t total_windows, const JacRec* __restrict__ wparts,
                                                                  JacRec* __restrict__ wsum) {
  __shared__ JacRec sh[kWinGroups];
  const int64_t g = blockIdx.x;
  if (g >= total_windows) return;
  const SlLane c = sl_lane<FqConst>();
  const SlJac mine = sl_from_jac(load_jac_rec(wparts + g * kWinGroups + ((int)threadIdx.x >> 6)));
  const Jac r = sl_to_jac<FqConst>(c, sl_workgroup_sum<kWinGroups>(sh, c, mine));
  if (threadIdx.x == 0) store_jac_rec(wsum + g, r);
}

// One workgroup per job: sum_w 2^(cw) U_w with the limb-sliced formulas. The windows split into
// kFinalWaves contiguous groups, the top one the smallest; wavefront g runs Horner over its group
// [lo, hi) from the top down, then doubles its sum c * lo times, and the groups' results meet in a
// two-level tree. Every path keeps the (rw - 1) c doublings a single Horner chain has (the top
// window's), but only the top group's additions and the tree's stay on it: at 2^10 points (22
// windows of 6 bits) 5 additions against 21.
constexpr int kFinalWaves = 4;

__global__ void __launch_bounds__(64 * kFinalWaves) k_final(const MsmJob* __restrict__ jobs, int njobs,
                                                            const JacRec* __restrict__ wsum) {
  __shared__ JacRec sh[kFinalWaves];
  const int jb = blockIdx.x;
  if (jb >= njobs) return;
  const MsmJob& job = jobs[jb];
  const FqConst f;
  const SlLane c = sl_lane<FqConst>();
  const int per = (job.rw + kFinalWaves - 1) / kFinalWaves;
  const int lo = min(job.rw, ((int)threadIdx.x >> 6) * per), hi = min(job.rw, lo + per);
  SlJac acc = sl_from_jac(hi > lo ? load_jac_rec(wsum + job.win_base + hi - 1) : jac_identity(f));
  for (int w = hi - 2; w >= lo; --w) {
    for (int e = 0; e < job.c; ++e) acc = sl_dbl<FqConst>(c, acc);
    acc = sl_add<FqConst>(c, acc, sl_from_jac(load_jac_rec(wsum + job.win_base + w)));
  }
  if (hi > lo)
    for (int e = job.c * lo; e > 0; --e) acc = sl_dbl<FqConst>(c, acc);
  Jac r = sl_to_jac<FqConst>(c, sl_workgroup_sum<kFinalWaves>(sh, c, acc));
  if (fe_is_zero_lazy(f, r.z)) r = jac_identity(f);
  if (threadIdx.x == 0) store_jac(f, job.out, r);
}

// Window table of an SRS: out[w * n + i] = 2^(c w) P_i as an affine record, w < windows (one
// thread per point: c doublings between windows, one inversion per stored point).
__global__ void __launch_bounds__(256) k_srs_window_table(const AffRec* __restrict__ pts, int64_t n, int c, int windows,
                                                          AffRec* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Fq f;
  const Aff p = unpack_aff(load_aff_raw(pts + i));
  out[i] = pts[i];
  Jac acc;
  acc.x = p.x;
  acc.y = p.y;
  acc.z = fe_const<Fq>(f.one);
  for (int w = 1; w < windows; ++w) {
    for (int e = 0; e < c; ++e) acc = jac_dbl(f, acc);
    store_aff_rec(out + (int64_t)w * n + i, jac_to_aff(f, acc));  // 2^(c w) P_i is finite: r is an odd prime
  }
}

// ---- host ---------------------------------------------------------------------

// Whether new plans split scalars with the GLV endomorphism (zk_msm_set_glv).
static int g_glv = 1;

// log2 of the reduction's segment length for new plans (zk_msm_set_segment); 0: automatic, 2
// for windows of 14 bits and more and 1 below (min c - 1). Narrow windows come with small MSMs,
// whose reduction is latency-bound: two-bucket leaves halve its first chain (2^10..2^17 points:
// 5% faster per MSM), while 2^18 and 2^20 points are 5% faster with four-bucket leaves.
static int g_log_seg = 0;

// The bucket accumulation kernel new plans use (zk_msm_set_accumulate): bit 0 picks the
// low-latency product schedule, bit 1 XYZZ coordinates.
static int g_acc_variant = 0;

using AccKernel = void (*)(const uint2*, const uint32_t*, const AffRec*, BucketRec*, RunSummary*);

static AccKernel acc_kernel(int variant) {
  switch (variant & 3) {
    case 1: return k_accumulate<FqMsmIlp, false>;
    case 2: return k_accumulate<Fq, true>;
    case 3: return k_accumulate<FqMsmIlp, true>;
    default: return k_accumulate<Fq, false>;
  }
}

int default_window(int64_t n) {
  // Signed digits: about n additions per window (2n over half as many windows with the GLV
  // split) plus ~3 * 2^(c-1) for each window's bucket reduction.
  const int bits = g_glv ? Glv::kBits : 254;
  const double streams = g_glv ? 2.0 : 1.0;
  int best = 1;
  double best_cost = 1e300;
  for (int c = 1; c <= 20; ++c) {
    const int windows = bits / c + 1;
    double cost = (double)windows * (streams * (double)n + 3.0 * (double)(1 << (c - 1)));
    if (cost < best_cost) {
      best_cost = cost;
      best = c;
    }
  }
  return best;
}

// Every buffer one batch of MSMs needs, sized at creation, so that launching
// allocates nothing (and the launch sequence can be captured into a graph).
struct MsmPlan {
  int njobs = 0;
  hipStream_t alloc_stream = nullptr;  // the stream whose allocator cache the memory returns to
  bool escaped = false;                // launched on another stream or recorded into a graph
  std::vector<MsmJob> jobs;
  std::vector<int64_t> level_max;
  int64_t blocks = 0, nkeys = 0, max_entries = 0, max_slots = 0, max_slots1 = 0, segments = 0, parts = 0,
          windows_total = 0;
  int64_t sort_blocks = 0, acc_blocks = 0;
  int64_t nparts = 0, max_chunks = 0;  // sort partitions; pass-B chunks at most
  size_t bytes = 0;
  void* srs = nullptr;  // the SRS handle the plan reads, if any
  int acc_variant = 0;  // the accumulation kernel (g_acc_variant when planned)
  // A plan over raw (n, 2, 4) point rows converts them in its first launch: prep_n points from
  // prep_src into prep_out[0, prep_n), and phi(P_i) after them when prep_phi.
  const uint64_t* prep_src = nullptr;
  int64_t prep_n = 0;
  AffRec* prep_out = nullptr;
  bool prep_phi = false;
  void* digits = nullptr;
  bool d16 = true;  // 16-bit digit codes (every job's window at most 16 bits), else 32-bit
  int group = kCombineThreads;  // summaries per combine workgroup
  const AffRec* points = nullptr;
  // device buffers
  void* mem = nullptr;
  MsmJob* d_jobs = nullptr;
  uint32_t *counts = nullptr, *totals = nullptr, *starts = nullptr, *m_dev = nullptr, *scan_ws = nullptr;
  uint32_t *chunks = nullptr, *chunk_base = nullptr, *counts_b = nullptr, *counts_raw = nullptr;
  uint2 *staged = nullptr, *sorted = nullptr;
  BucketRec* buckets = nullptr;
  RunSummary* runs[2] = {nullptr, nullptr};
  XyzzRec *tot = nullptr, *run = nullptr;
  JacRec *parts_buf = nullptr, *wparts = nullptr, *wsum = nullptr;
};

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets of every buffer inside a plan's single device allocation.
struct MsmLayout {
  size_t jobs, counts, totals, starts, digits, m, staged, sorted, scan, chunks, chunk_base, counts_b, counts_raw, buckets, r0, r1,
      tot, run,
      parts, wparts, wsum, total;
};

// jobs: scalars, n, c (0 = default) and out filled in by the caller; the rest is planned here,
// without touching the device.
static int msm_plan_geometry(MsmPlan* plan, std::vector<MsmJob>& jobs, int chunk, MsmLayout* L) {
  plan->acc_variant = g_acc_variant;
  if (chunk < 0) return fail("chunk_len must be positive (or 0 for automatic)");
  int64_t blocks = 0, keys = 0, entries = 0, segs = 0, parts = 0, wins = 0, sort_blocks = 0, cnts = 0, pkeys = 0;
  for (auto& j : jobs) {
    if (j.n < 0 || j.n > 0x3fffffff) return fail("MSM size must be in [0, 2^30)");
    if (j.c <= 0) j.c = default_window(j.n);
    if (j.c > 20) return fail("window must be in 1..20");
    j.glv = j.table ? 0 : g_glv;
    j.nv = j.glv ? 2 * j.n : j.n;
    j.windows = (j.glv ? Glv::kBits : 254) / j.c + 1;
    j.rw = j.table ? 1 : j.windows;
    j.log_seg = std::min(g_log_seg > 0 ? g_log_seg : (j.c >= 14 ? 2 : 1), j.c - 1);
    j.log_segs = (j.c - 1) - j.log_seg;
    // chunks per reduction row: a row adds half the segments, or (the first T half) ceil(segs / 2)
    j.chunks = (int)((((1LL << j.log_segs) + 1) / 2 + kRedChunk - 1) / kRedChunk);
    j.block_base = blocks;
    j.key_base = keys;
    j.seg_base = segs;
    j.part_base = parts;
    j.win_base = wins;
    // sort geometry: windows x scalar tiles, about a thousand blocks for large jobs (pass A), and
    // partitions of 2^pbits buckets (pass B)
    j.pbits = std::min(j.c - 1, kPartBits);
    const int64_t np = 1LL << (j.c - 1 - j.pbits);
    j.ranges = 1;
    const int64_t want = (1024 + (int64_t)j.windows - 1) / j.windows;
    j.tiles = (int)std::max<int64_t>(1, std::min<int64_t>(want, (j.nv + 2047) / 2048));
    j.tile = std::max<int64_t>(1, (j.nv + j.tiles - 1) / j.tiles);
    j.dig_base = entries;
    j.sort_base = sort_blocks;
    j.cnt_base = cnts;
    j.pkey_base = pkeys;
    sort_blocks += (int64_t)j.windows * j.tiles;
    cnts += (int64_t)j.windows * np * j.tiles;
    pkeys += (int64_t)j.rw * np;
    blocks += std::max<int64_t>(1, (j.n + 255) / 256);
    keys += (int64_t)j.rw << (j.c - 1);
    entries += (int64_t)j.windows * j.nv;
    segs += (int64_t)j.rw << j.log_segs;
    parts += (int64_t)j.rw * (j.log_segs + 2) * j.chunks;
    wins += j.rw;
  }
  if (keys >= 0x7fffffff || entries >= 0xffffffffLL) return fail("MSM batch too large for 32-bit bucket keys");
  // Accumulation grid: every resident workgroup slot of the GPU once (the kernel splits the sorted
  // entries evenly over its threads), or, with chunk_len > 0, enough workgroups for chunk_len
  // entries per thread at the largest possible entry count.
  int64_t grid0;
  if (chunk > 0) {
    grid0 = (entries + (int64_t)chunk * kAccThreads - 1) / ((int64_t)chunk * kAccThreads);
  } else {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, acc_kernel(plan->acc_variant), kAccThreads, 0) !=
            hipSuccess ||
        per_cu < 1) {
      (void)hipGetLastError();
      per_cu = 1;
    }
    grid0 = std::min<int64_t>((int64_t)device_cus() * per_cu, (entries + 4 * kAccThreads - 1) / (4 * kAccThreads));
  }
  grid0 = std::max<int64_t>(1, grid0);
  plan->jobs = jobs;
  plan->njobs = (int)jobs.size();
  plan->d16 = true;
  for (const auto& j : jobs) plan->d16 = plan->d16 && j.c <= 16;
  plan->blocks = blocks;
  plan->nkeys = keys;
  plan->nparts = pkeys;
  plan->max_chunks = (entries + kChunkB - 1) / kChunkB + pkeys;  // every partition adds at most one partial chunk
  plan->max_entries = entries;
  plan->segments = segs;
  plan->sort_blocks = sort_blocks;
  plan->parts = parts;
  plan->windows_total = wins;
  // Accumulation levels: level 0 leaves one run summary per thread; each further level merges 256
  // summaries per workgroup, until a single workgroup remains. Levels alternate between two buffers.
  plan->acc_blocks = grid0;
  plan->level_max.assign(1, grid0 * kAccThreads);
  const int64_t lvl0 = grid0 * kAccThreads;
  plan->group = lvl0 <= keys || (lvl0 + 63) / 64 <= device_cus() ? 64 : kCombineThreads;
  while (plan->level_max.back() > 1) plan->level_max.push_back((plan->level_max.back() + plan->group - 1) / plan->group);
  int64_t slots[2] = {1, 1};
  for (size_t l = 0; l < plan->level_max.size(); ++l) slots[l & 1] = std::max(slots[l & 1], plan->level_max[l]);
  // the boundary pass before the first tree level reads level 0 from one buffer and writes as many
  // reduced summaries into the other, so both hold the level-0 count
  const int64_t max_slots = slots[0], max_slots1 = std::max(slots[1], slots[0]);
  plan->max_slots = max_slots;
  plan->max_slots1 = max_slots1;
  const size_t e1 = (size_t)std::max<int64_t>(entries, 1);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t at = off;
    off += align_up(bytes);
    return at;
  };
  L->jobs = take(sizeof(MsmJob) * jobs.size());
  L->counts = take(4 * (size_t)std::max<int64_t>(cnts, 1));
  L->totals = take(4 * (size_t)(pkeys + 1));
  L->starts = take(4 * (size_t)(pkeys + 1));
  L->digits = take(4 * e1);
  L->m = take(16);
  L->staged = take(8 * e1);
  L->sorted = take(8 * e1);
  L->scan = take(4 * scan_workspace_words(pkeys));
  L->chunks = take(4 * (size_t)(pkeys + 1));
  L->chunk_base = take(4 * (size_t)(pkeys + 1));
  L->counts_b = take(4 * ((size_t)plan->max_chunks << kPartBits));
  L->counts_raw = take(4 * ((size_t)plan->max_chunks << kPartBits));
  L->buckets = take(sizeof(BucketRec) * (size_t)keys);
  L->r0 = take(sizeof(RunSummary) * (size_t)max_slots);
  L->r1 = take(sizeof(RunSummary) * (size_t)max_slots1);
  L->tot = take(sizeof(XyzzRec) * (size_t)segs);
  L->run = take(sizeof(XyzzRec) * (size_t)segs);
  L->parts = take(sizeof(JacRec) * (size_t)parts);
  L->wparts = take(sizeof(JacRec) * (size_t)wins * kWinGroups);
  L->wsum = take(sizeof(JacRec) * (size_t)wins);
  L->total = off;
  return 0;
}

// Points every buffer of the plan at its offset in the block at base (the plan's memory).
static void msm_plan_bind(MsmPlan* plan, const MsmLayout& L, void* mem) {
  plan->mem = mem;
  plan->bytes = L.total;
  char* base = static_cast<char*>(mem);
  plan->d_jobs = reinterpret_cast<MsmJob*>(base + L.jobs);
  plan->counts = reinterpret_cast<uint32_t*>(base + L.counts);
  plan->totals = reinterpret_cast<uint32_t*>(base + L.totals);
  plan->starts = reinterpret_cast<uint32_t*>(base + L.starts);
  plan->digits = base + L.digits;
  plan->m_dev = reinterpret_cast<uint32_t*>(base + L.m);
  plan->staged = reinterpret_cast<uint2*>(base + L.staged);
  plan->sorted = reinterpret_cast<uint2*>(base + L.sorted);
  plan->chunks = reinterpret_cast<uint32_t*>(base + L.chunks);
  plan->chunk_base = reinterpret_cast<uint32_t*>(base + L.chunk_base);
  plan->counts_b = reinterpret_cast<uint32_t*>(base + L.counts_b);
  plan->counts_raw = reinterpret_cast<uint32_t*>(base + L.counts_raw);
  plan->scan_ws = reinterpret_cast<uint32_t*>(base + L.scan);
  plan->buckets = reinterpret_cast<BucketRec*>(base + L.buckets);
  plan->runs[0] = reinterpret_cast<RunSummary*>(base + L.r0);
  plan->runs[1] = reinterpret_cast<RunSummary*>(base + L.r1);
  plan->tot = reinterpret_cast<XyzzRec*>(base + L.tot);
  plan->run = reinterpret_cast<XyzzRec*>(base + L.run);
  plan->parts_buf = reinterpret_cast<JacRec*>(base + L.parts);
  plan->wparts = reinterpret_cast<JacRec*>(base + L.wparts);
  plan->wsum = reinterpret_cast<JacRec*>(base + L.wsum);
}

int msm_plan_create(const AffRec* points, std::vector<MsmJob> jobs, int chunk, hipStream_t stream, MsmPlan** out) {
  auto plan = new MsmPlan();
  plan->points = points;
  MsmLayout L;
  if (int rc = msm_plan_geometry(plan, jobs, chunk, &L)) {
    delete plan;
    return rc;
  }
  void* mem = dev_alloc(L.total, stream);
  if (!mem) {
    delete plan;
    return 1;
  }
  plan->alloc_stream = stream;
  msm_plan_bind(plan, L, mem);
  // Stream-ordered upload from the plan's own job table, which lives as long as the plan.
  if (!plan->jobs.empty()) {
    hipError_t e = hipMemcpyAsync(plan->d_jobs, plan->jobs.data(), sizeof(MsmJob) * plan->jobs.size(),
                                  hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) {
      dev_free(plan->mem);
      delete plan;
      return fail_hip(e, "MSM plan upload");
    }
  }
  *out = plan;
  return 0;
}

void msm_plan_destroy(MsmPlan* plan) {
  if (!plan) return;
  if (plan->escaped) (void)hipDeviceSynchronize();  // work the allocating stream does not order
  dev_free(plan->mem);
  delete plan;
}

// The plan's first launch: digits, the buckets cleared (a fill kernel, not a memset node, so that
// graph replays clear them too), and the raw points converted when the plan has any.
static int msm_plan_prepare(MsmPlan* plan, hipStream_t stream) {
  static_assert(sizeof(BucketRec) == 36 * 4, "a bucket record is 36 words");
  const int64_t clear_words = plan->nkeys * 36;
  const int64_t grid = std::min<int64_t>(
      std::max<int64_t>({plan->blocks, (int64_t)grid_for(clear_words, 256), (int64_t)grid_for(plan->prep_n, 256), 1}),
      (int64_t)1 << 16);
  AffRec* phi = plan->prep_phi ? plan->prep_out + plan->prep_n : nullptr;
  uint32_t* clear = reinterpret_cast<uint32_t*>(plan->buckets);
  if (plan->d16)
    hipLaunchKernelGGL(k_prepare<uint16_t>, dim3((unsigned)grid), dim3(256), 0, stream, plan->d_jobs, plan->njobs,
                       plan->blocks, static_cast<uint16_t*>(plan->digits), clear, clear_words, plan->prep_src,
                       plan->prep_n, plan->prep_out, phi);
  else
    hipLaunchKernelGGL(k_prepare<uint32_t>, dim3((unsigned)grid), dim3(256), 0, stream, plan->d_jobs, plan->njobs,
                       plan->blocks, static_cast<uint32_t*>(plan->digits), clear, clear_words, plan->prep_src,
                       plan->prep_n, plan->prep_out, phi);
  ZK_LAUNCH_CHECK("msm digits");
  return 0;
}

// MSM launches run one at a time on the device, whatever stream issues them: each launch waits
// for the previous one's completion event and records its own. Two MSM launch sequences executing
// side by side on two streams gave wrong (still on-curve) results for about 1 MSM in 10^4, while an
// MSM beside NTT or copy kernels on another stream, and MSMs on two streams one after the other,
// gave none in several 10^5; the cause was not found, so MSMs are kept from overlapping each other.
// Work on other streams (quotients, NTTs) still overlaps them. Launches recorded into a HIP graph
// take no part in this order (a replay must not run beside other MSM work).
static std::mutex g_order_mu;
static hipEvent_t g_order_ev = nullptr;

static int msm_order_enter(hipStream_t stream, bool* ordered) {
  *ordered = false;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  ZK_HIP(hipStreamIsCapturing(stream, &cap));
  if (cap != hipStreamCaptureStatusNone) return 0;
  std::lock_guard<std::mutex> lk(g_order_mu);
  if (!g_order_ev) ZK_HIP(hipEventCreateWithFlags(&g_order_ev, hipEventDisableTiming));
  ZK_HIP(hipStreamWaitEvent(stream, g_order_ev, 0));
  *ordered = true;
  return 0;
}

static int msm_order_leave(hipStream_t stream) {
  std::lock_guard<std::mutex> lk(g_order_mu);
  ZK_HIP(hipEventRecord(g_order_ev, stream));
  return 0;
}

static int msm_plan_launch_kernels(MsmPlan* plan, hipStream_t stream);

// Enqueues the whole batch after every earlier MSM launch; allocates nothing.
int msm_plan_launch(MsmPlan* plan, hipStream_t stream) {
  if (plan->njobs == 0) return 0;
  if (plan->mem && launch_escapes(plan->alloc_stream, stream)) plan->escaped = true;
  bool ordered = false;
  if (int rc = msm_order_enter(stream, &ordered)) return rc;
  const int rc = msm_plan_launch_kernels(plan, stream);
  if (ordered) {
    // recorded even after a failed launch, so that later launches do not wait on a stale order
    if (int rc2 = msm_order_leave(stream)) return rc ? rc : rc2;
  }
  return rc;
}

static int msm_plan_launch_kernels(MsmPlan* plan, hipStream_t stream) {
  const int nj = plan->njobs;
  if (int rc = msm_plan_prepare(plan, stream)) return rc;
  {
    if (plan->d16)
      hipLaunchKernelGGL(k_sort_count<uint16_t>, dim3((unsigned)plan->sort_blocks), dim3(256), 0, stream, plan->d_jobs, nj,
                         static_cast<const uint16_t*>(plan->digits), plan->counts);
    else
      hipLaunchKernelGGL(k_sort_count<uint32_t>, dim3((unsigned)plan->sort_blocks), dim3(256), 0, stream, plan->d_jobs, nj,
                         static_cast<const uint32_t*>(plan->digits), plan->counts);
    ZK_LAUNCH_CHECK("msm sort count");
    hipLaunchKernelGGL(k_sort_columns, dim3(grid_for(plan->nparts * 64, 256)), dim3(256), 0, stream, plan->d_jobs, nj,
                       plan->nparts, plan->counts, plan->totals);
    ZK_LAUNCH_CHECK("msm sort columns");
    const bool few = plan->nparts <= kOffMaxParts;
    if (few) {
      hipLaunchKernelGGL(k_part_offsets, dim3(1), dim3(kOffThreads), 0, stream, plan->totals, plan->nparts, plan->starts,
                         plan->m_dev, plan->chunk_base);
      ZK_LAUNCH_CHECK("msm sort offsets");
    } else if (int rc = scan_u32_ws(plan->totals, plan->starts, plan->nparts, plan->m_dev, plan->scan_ws, stream)) {
      return rc;
    }
    if (plan->d16)
      hipLaunchKernelGGL(k_sort_place<uint16_t>, dim3((unsigned)plan->sort_blocks), dim3(256), 0, stream, plan->d_jobs, nj,
                         static_cast<const uint16_t*>(plan->digits), plan->counts, plan->starts, plan->staged);
    else
      hipLaunchKernelGGL(k_sort_place<uint32_t>, dim3((unsigned)plan->sort_blocks), dim3(256), 0, stream, plan->d_jobs, nj,
                         static_cast<const uint32_t*>(plan->digits), plan->counts, plan->starts, plan->staged);
    ZK_LAUNCH_CHECK("msm sort place");
    // pass B: every partition sorted by bucket, chunk by chunk
    if (!few) {
      hipLaunchKernelGGL(k_part_chunks, dim3(grid_for(plan->nparts, 256)), dim3(256), 0, stream, plan->totals, plan->nparts,
                         plan->chunks);
      ZK_LAUNCH_CHECK("msm sort chunks");
      if (int rc = scan_u32_ws(plan->chunks, plan->chunk_base, plan->nparts, plan->chunk_base + plan->nparts, plan->scan_ws,
                               stream))
        return rc;
    }
    hipLaunchKernelGGL(k_part_count, dim3((unsigned)plan->max_chunks), dim3(256), 0, stream, plan->d_jobs, nj, plan->chunk_base,
                       plan->nparts, plan->starts, plan->totals, plan->staged, plan->counts_b, plan->counts_raw);
    ZK_LAUNCH_CHECK("msm sort partition count");
    hipLaunchKernelGGL(k_part_columns, dim3((unsigned)plan->nparts), dim3(256), 0, stream, plan->d_jobs, nj, plan->chunk_base,
                       plan->nparts, plan->starts, plan->counts_b);
    ZK_LAUNCH_CHECK("msm sort partition columns");
    hipLaunchKernelGGL(k_part_place, dim3((unsigned)plan->max_chunks), dim3(kPlaceThreads), 0, stream, plan->d_jobs, nj, plan->chunk_base,
                       plan->nparts, plan->starts, plan->totals, plan->staged, plan->counts_b, plan->counts_raw,
                       plan->sorted);
    ZK_LAUNCH_CHECK("msm sort partition place");
  }
  const int levels = (int)plan->level_max.size();
  if (plan->level_max[0] > plan->max_slots) return fail("internal: MSM run-summary buffer smaller than a level");
  launch_timed("msm_accumulate", acc_kernel(plan->acc_variant), dim3((unsigned)plan->acc_blocks), dim3(kAccThreads), 0,
               stream, plan->sorted, plan->m_dev, plan->points, plan->buckets, plan->runs[0]);
  ZK_LAUNCH_CHECK("msm accumulate");
  {
    // every pass writes the other summary buffer: check the count against its capacity first
    const int64_t cap[2] = {plan->max_slots, plan->max_slots1};
    // A second boundary pass, before the second tree level, pays when the threads outnumber the
    // buckets: ranges are then shorter than a typical bucket, most buckets reach the trees, and the
    // first level leaves many two-piece chains across its workgroups' edges, which the pass adds up
    // side by side. With longer ranges only the hot buckets' chains remain there, and the trees'
    // own closing of complete runs leaves the pass nothing to do.
    const bool second_pass = plan->level_max[0] > plan->nkeys;
    int cur = 0;
    for (int l = 1; l < levels; ++l) {
      const bool pass = l == 1 || (l == 2 && second_pass);
      if ((pass && plan->level_max[(size_t)l - 1] > cap[cur ^ 1]) || plan->level_max[(size_t)l] > cap[pass ? cur : cur ^ 1])
        return fail("internal: MSM run-summary buffer smaller than a level");
      if (pass) {
        // resolve the summaries' boundaries in one pass
        hipLaunchKernelGGL(k_fixup, dim3(grid_for(plan->level_max[l - 1], 256)), dim3(256), 0, stream,
                           plan->runs[cur], plan->level_max[l - 1], plan->buckets, plan->runs[cur ^ 1]);
        ZK_LAUNCH_CHECK("msm boundaries");
        cur ^= 1;
      }
      const int lanes = kCombineLanes;
      hipLaunchKernelGGL(k_combine, dim3((unsigned)plan->level_max[(size_t)l]), dim3(lanes), 0, stream,
                         plan->runs[cur], plan->level_max[(size_t)l - 1], plan->buckets, plan->runs[cur ^ 1],
                         l + 1 == levels ? 1 : 0, plan->group);
      ZK_LAUNCH_CHECK("msm combine");
      cur ^= 1;
    }
  }
  {
    hipLaunchKernelGGL(k_reduce_leaf, dim3(grid_for(plan->segments, 256)), dim3(256), 0, stream, plan->d_jobs, nj,
                       plan->segments, plan->buckets, plan->tot, plan->run);
    ZK_LAUNCH_CHECK("msm reduce leaf");
    hipLaunchKernelGGL(k_reduce_bits, dim3((unsigned)plan->parts), dim3(256), 0, stream, plan->d_jobs, nj, plan->tot,
                       plan->run, plan->parts_buf);
    ZK_LAUNCH_CHECK("msm reduce bits");
    hipLaunchKernelGGL(k_window_terms, dim3((unsigned)(plan->windows_total * kWinGroups)), dim3(64 * kWinWaves), 0, stream,
                       plan->d_jobs, nj, plan->windows_total, plan->parts_buf, plan->wparts);
    ZK_LAUNCH_CHECK("msm window terms");
    hipLaunchKernelGGL(k_window_total, dim3((unsigned)plan->windows_total), dim3(64 * kWinGroups), 0, stream,
                       plan->windows_total, plan->wparts, plan->wsum);
    ZK_LAUNCH_CHECK("msm window sums");
  }
  hipLaunchKernelGGL(k_final, dim3((unsigned)nj), dim3(64 * kFinalWaves), 0, stream, plan->d_jobs, nj, plan->wsum);
  ZK_LAUNCH_CHECK("msm final");
  return 0;
}

// Points (n, 2, 4) 256-form rows -> packed 261-form records at out[0, n), and phi(P_i) at
// out[n, 2n) when with_phi.
int convert_points(const uint64_t* points, int64_t n, AffRec* out, bool with_phi, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_prep_points, dim3(grid_for(n, 256)), dim3(256), 0, stream, points, n, out,
                     with_phi ? out + n : nullptr);
  ZK_LAUNCH_CHECK("msm prep");
  return 0;
}

int msm_g1(const uint64_t* points, const uint64_t* scalars, int64_t n, int c, int k0, uint64_t* out_jac,
           hipStream_t stream, uint64_t* entries) {
  if (n < 0 || n > 0x3fffffff) return fail("MSM size must be in [0, 2^30)");
  if (c < 0 || c > 20) return fail("window must be in 1..20");
  const bool glv = g_glv != 0;
  Scratch sc(stream);
  AffRec* pts = sc.get<AffRec>((size_t)std::max<int64_t>(glv ? 2 * n : n, 1));
  if (!pts) return 1;
  MsmJob job{};
  job.scalars = scalars;
  job.n = n;
  job.c = c;
  job.out = out_jac;
  job.phi_base = n;
  MsmPlan* plan = nullptr;
  if (int rc = msm_plan_create(pts, {job}, k0, stream, &plan)) return rc;
  plan->prep_src = points;  // the points are converted in the plan's first launch
  plan->prep_n = n;
  plan->prep_out = pts;
  plan->prep_phi = glv;
  int rc = msm_plan_launch(plan, stream);
  if (rc == 0 && entries) {  // the sorted-entry count the accumulation processed
    uint32_t m = 0;
    hipError_t err = hipMemcpyAsync(&m, plan->m_dev, sizeof(m), hipMemcpyDeviceToHost, stream);
    if (err == hipSuccess) err = hipStreamSynchronize(stream);
    rc = err == hipSuccess ? 0 : fail_hip(err, "reading the entry count");
    *entries = m;
  }
  msm_plan_destroy(plan);  // blocks return to this stream's cache: later work on it is ordered after the launch
  return rc;
}

// ---- the MSM entry that allocates nothing (zk_msm_g1_noalloc) ------------------------------

// Device memory of zk_msm_g1_noalloc: point records first, then the plan's buffers, laid out as
// zk_msm_g1 lays out its workspace. Each GPU holds its own copy of this array from the module's
// load until the process ends. On the host the name only stands for the symbol: host code uses
// the address hipGetSymbolAddress gives for the current device, never the name itself.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kMaxDevices];  // per device: the device address of scratch

// Looks up the current device's copy of scratch once per device, with that device current
// (zk_init calls it; zk_msm_g1_noalloc calls it again on a device it has not seen).
int msm_scratch_init() {
  int dev = 0;
  ZK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("device index out of range for the MSM scratch");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* addr = nullptr;
    ZK_HIP(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
    if (!addr) return fail("no device address for the MSM scratch");
    g_scratch_addr[dev] = addr;
  }
  return 0;
}

static void* scratch_addr() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_addr[dev];
}

// Writes the job table from a kernel argument (no host-to-device copy).
__global__ void k_store_job(MsmJob job, MsmJob* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = job;
}

// zk_msm_g1 with every device buffer in scratch: the same plan geometry, kernels and launch order
// for the same arguments, so the same Jacobian result.
int msm_g1_noalloc(const uint64_t* points, const uint64_t* scalars, int64_t n, int c, int k0, uint64_t* out_jac,
                   hipStream_t stream) {
  if (n < 0 || n > 0x3fffffff) return fail("MSM size must be in [0, 2^30)");
  if (c < 0 || c > 20) return fail("window must be in 1..20");
  if (int rc = msm_scratch_init()) return rc;
  char* base = static_cast<char*>(scratch_addr());
  if (!base) return fail("the MSM scratch has no device address");
  const bool glv = g_glv != 0;
  const size_t rec_bytes = align_up(sizeof(AffRec) * (size_t)std::max<int64_t>(glv ? 2 * n : n, 1));
  MsmJob job{};
  job.scalars = scalars;
  job.n = n;
  job.c = c;
  job.out = out_jac;
  job.phi_base = n;
  MsmPlan plan;
  MsmLayout L;
  std::vector<MsmJob> jobs{job};
  if (int rc = msm_plan_geometry(&plan, jobs, k0, &L)) return rc;
  if (rec_bytes + L.total > kNoAllocScratchBytes)
    return fail("zk_msm_g1_noalloc: an MSM of " + std::to_string(n) + " points with these parameters needs " +
                std::to_string(rec_bytes + L.total) + " bytes of scratch, more than its " +
                std::to_string(kNoAllocScratchBytes) + " (use zk_msm_g1)");
  AffRec* pts = reinterpret_cast<AffRec*>(base);
  plan.points = pts;
  plan.prep_src = points;  // converted in the plan's first launch
  plan.prep_n = n;
  plan.prep_out = pts;
  plan.prep_phi = glv;
  msm_plan_bind(&plan, L, base + rec_bytes);
  hipLaunchKernelGGL(k_store_job, dim3(1), dim3(64), 0, stream, plan.jobs[0], plan.d_jobs);
  ZK_LAUNCH_CHECK("msm job table");
  return msm_plan_launch(&plan, stream);
}

}  // namespace zk

using namespace zk;

extern "C" {

int zk_msm_default_window(int64_t n, int* window) {
  *window = default_window(n);
  return 0;
}

int zk_msm_set_glv(int on) {
  g_glv = on ? 1 : 0;
  return 0;
}

int zk_msm_get_glv(int* on) {
  *on = g_glv;
  return 0;
}

int zk_msm_set_accumulate(int variant) {
  if (variant < 0 || variant > 3) return fail("accumulation variant must be in 0..3");
  g_acc_variant = variant;
  return 0;
}

int zk_msm_get_accumulate(int* variant) {
  *variant = g_acc_variant;
  return 0;
}

int zk_msm_set_segment(int log_seg) {
  if (log_seg < 0 || log_seg > 6) return fail("segment length must be 2^0..2^6 (0: automatic)");
  g_log_seg = log_seg;
  return 0;
}

int zk_msm_get_segment(int* log_seg) {
  *log_seg = g_log_seg;
  return 0;
}

int zk_msm_g1(const uint64_t* points_affine, const uint64_t* scalars, int64_t n, int window, int chunk_len,
              uint64_t* out_jac, void* stream) {
  return msm_g1(points_affine, scalars, n, window, chunk_len, out_jac, (hipStream_t)stream);
}

int zk_msm_g1_entries(const uint64_t* points_affine, const uint64_t* scalars, int64_t n, int window, int chunk_len,
                      uint64_t* out_jac, uint64_t* entries, void* stream) {
  return msm_g1(points_affine, scalars, n, window, chunk_len, out_jac, (hipStream_t)stream, entries);
}

int zk_msm_g1_noalloc(const uint64_t* points_affine, const uint64_t* scalars, int64_t n, int window, int chunk_len,
                      uint64_t* out_jac, void* stream) {
  return msm_g1_noalloc(points_affine, scalars, n, window, chunk_len, out_jac, (hipStream_t)stream);
}

int zk_msm_noalloc_max_bytes(uint64_t* bytes) {
  *bytes = kNoAllocScratchBytes;
  return 0;
}

// An SRS on the device: records [0, n) hold P_i, [n, 2n) phi(P_i), and each window table of
// width c holds 2^(c w) P_i at base + w * n + i.
struct SrsTable {
  int c, windows;
  int64_t base;
};

struct ZkSrs {
  AffRec* pts;
  int64_t n;
  int64_t records;
  std::vector<SrsTable> tables;
  int plans = 0;  // live plans hold pts: no table may be added meanwhile
};

int zk_srs_create(const uint64_t* points_affine, int64_t n, void** srs, void* stream) {
  if (n <= 0 || n > 0x3fffffff) return fail("SRS size must be in [1, 2^30)");
  AffRec* p = nullptr;
  ZK_HIP(hipMalloc(&p, sizeof(AffRec) * 2 * (size_t)n));  // P_i, then phi(P_i)
  if (int rc = convert_points(points_affine, n, p, true, (hipStream_t)stream)) {
    (void)hipFree(p);
    return rc;
  }
  ZK_HIP(hipStreamSynchronize((hipStream_t)stream));
  auto* s = new ZkSrs{p, n, 2 * n, {}, 0};
  *srs = s;
  return 0;
}

int zk_srs_precompute(void* srs, int window, void* stream_) {
  auto* s = static_cast<ZkSrs*>(srs);
  hipStream_t stream = (hipStream_t)stream_;
  if (!s) return fail("an SRS handle is required");
  if (window < 4 || window > 20) return fail("window table width must be in 4..20");
  if (s->plans > 0) return fail("window tables must be added before plans are created over the SRS");
  for (const SrsTable& t : s->tables)
    if (t.c == window) return 0;
  const int windows = 254 / window + 1;
  const size_t records = (size_t)s->records + (size_t)windows * (size_t)s->n;
  if (records >= ((size_t)1 << 31)) return fail("window tables would exceed 2^31 point records (31-bit point references)");
  AffRec* p = nullptr;
  ZK_HIP(hipMalloc(&p, sizeof(AffRec) * records));
  hipError_t e = hipMemcpyAsync(p, s->pts, sizeof(AffRec) * (size_t)s->records, hipMemcpyDeviceToDevice, stream);
  if (e != hipSuccess) {
    (void)hipFree(p);
    return fail_hip(e, "SRS copy");
  }
  hipLaunchKernelGGL(k_srs_window_table, dim3(grid_for(s->n, 256)), dim3(256), 0, stream, s->pts, s->n, window, windows,
                     p + s->records);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    (void)hipFree(p);
    return fail_hip(e, "SRS window table");
  }
  (void)hipFree(s->pts);
  s->pts = p;
  s->tables.push_back(SrsTable{window, windows, s->records});
  s->records = (int64_t)records;
  return 0;
}

int zk_srs_destroy(void* srs) {
  if (!srs) return 0;
  auto* s = static_cast<ZkSrs*>(srs);
  ZK_HIP(hipDeviceSynchronize());
  (void)hipFree(s->pts);
  delete s;
  return 0;
}

// The window table an MSM of n points uses: the one with the least work (entries plus
// reduction) among the SRS's tables, or the one of the requested width; nullptr for none.
static const SrsTable* pick_table(const ZkSrs* s, int64_t n, int window) {
  const SrsTable* best = nullptr;
  double best_cost = 1e300;
  for (const SrsTable& t : s->tables) {
    if (window > 0 && t.c != window) continue;
    const double cost = (double)t.windows * (double)n + 3.0 * (double)(1LL << (t.c - 1));
    if (cost < best_cost) {
      best_cost = cost;
      best = &t;
    }
  }
  return best;
}

// The job table of count MSMs over prefixes of an SRS (window tables picked as plans pick them).
static int srs_jobs(const ZkSrs* s, int count, const uint64_t* const* scalars, const int64_t* n, const int32_t* window,
                    uint64_t* out_jac, std::vector<MsmJob>& jobs) {
  jobs.assign((size_t)std::max(count, 0), MsmJob{});
  for (int j = 0; j < count; ++j) {
    if (n[j] < 0 || n[j] > s->n) return fail("MSM size exceeds the SRS");
    MsmJob& job = jobs[(size_t)j];
    job.scalars = scalars ? scalars[j] : nullptr;
    job.phi_base = s->n;
    job.n = n[j];
    job.c = window ? window[j] : 0;
    job.out = out_jac ? out_jac + 12 * j : nullptr;
    if (const SrsTable* t = pick_table(s, n[j], job.c)) {
      job.table = 1;
      job.c = t->c;
      job.table_base = t->base;
      job.table_stride = s->n;
    }
  }
  return 0;
}

int zk_msm_plan_create(void* srs, int count, const uint64_t* const* scalars, const int64_t* n, const int32_t* window,
                       int chunk_len, uint64_t* out_jac, void** plan, void* stream) {
  auto* s = static_cast<ZkSrs*>(srs);
  if (!s) return fail("an SRS handle is required");
  std::vector<MsmJob> jobs;
  if (int rc = srs_jobs(s, count, scalars, n, window, out_jac, jobs)) return rc;
  MsmPlan* p = nullptr;
  if (int rc = msm_plan_create(s->pts, jobs, chunk_len, (hipStream_t)stream, &p)) return rc;
  p->srs = s;
  ++s->plans;
  *plan = p;
  return 0;
}

int zk_msm_workspace_bytes(void* srs, int count, const int64_t* n, const int32_t* window, int chunk_len, uint64_t* bytes) {
  std::vector<MsmJob> jobs;
  if (srs) {
    if (int rc = srs_jobs(static_cast<ZkSrs*>(srs), count, nullptr, n, window, nullptr, jobs)) return rc;
  } else {
    jobs.assign((size_t)std::max(count, 0), MsmJob{});
    for (int j = 0; j < count; ++j) {
      jobs[(size_t)j].n = n[j];
      jobs[(size_t)j].c = window ? window[j] : 0;
    }
  }
  MsmPlan plan;
  MsmLayout L;
  if (int rc = msm_plan_geometry(&plan, jobs, chunk_len, &L)) return rc;
  *bytes = L.total;
  return 0;
}

int zk_msm_plan_bytes(void* plan, uint64_t* bytes) {
  auto* p = static_cast<MsmPlan*>(plan);
  if (!p) return fail("a plan handle is required");
  *bytes = p->bytes;
  return 0;
}

int zk_msm_plan_launch(void* plan, void* stream) { return msm_plan_launch(static_cast<MsmPlan*>(plan), (hipStream_t)stream); }

int zk_msm_plan_destroy(void* plan) {
  auto* p = static_cast<MsmPlan*>(plan);
  if (p && p->srs) --static_cast<ZkSrs*>(p->srs)->plans;
  msm_plan_destroy(p);
  return 0;
}

int zk_msm_g1_batch(void* srs, int count, const uint64_t* const* scalars, const int64_t* n, const int32_t* window,
                    uint64_t* out_jac, void* stream) {
  void* plan = nullptr;
  if (int rc = zk_msm_plan_create(srs, count, scalars, n, window, 0, out_jac, &plan, stream)) return rc;
  int rc = zk_msm_plan_launch(plan, stream);
  zk_msm_plan_destroy(plan);
  return rc;
}

}  // extern "C"
