// Radix-2 NTT over LDS-staged multi-stage passes (include/zkgpu.h, "NTT").
//
// Algorithm. Decimation in frequency (Gentleman-Sande): stage s pairs rows
// i and i + n/2^(s+1) and maps (u, v) to (u + v, (u - v) * w) with
// w = omega^(j * 2^s), j = i mod n/2^(s+1). After log2(n) stages row i holds
// the transform at bitrev(i); the last pass writes there directly, so input
// and output are both in natural order, as in the reference's transform
// (src/zkgpu/ntt.py:149-203 through field256_gpu.py:375-387). Every value is
// fully reduced, so the output equals the DFT bit for bit whatever the
// butterfly order.
//
// Passes. A pass runs k consecutive stages on independent groups of 2^k
// rows (stride S = n / 2^(s0 + k)); a workgroup stages a tile of C groups x
// 2^k rows (1024 rows) in LDS as nine 29-bit limb planes, so consecutive
// threads touch consecutive banks. Groups of a tile are C consecutive columns,
// so every global row access of a wavefront covers C consecutive rows. The
// last pass takes groups that are far apart in index, so that after bit
// reversal its writes are again C consecutive rows. 2^22 runs as 8 + 7 + 7
// stages: three read+write sweeps of HBM instead of the reference's 23.
//
// Reduction. Inside a pass values stay in [0, 2p): a sum is brought back under
// 2p, a difference u - v + 4p (limbs kept redundant, no carries) goes straight
// into the product, whose result is below 2p with no final subtraction
// (zk_field.hpp, fsub_for_mul); intermediate passes store such values,
// the last pass stores canonical ones. In a pair of stages done in registers
// the first stage's sums stay wide (below 4p, limbs without carries) and are
// reduced once, by the second stage: its sum by a quotient estimated from the
// top limb, its difference + 8p inside the product (fsub_wide). The last stage
// of a transform has the single twiddle 1 and does no product.
//
// Twiddles. One table per stage, w_s[j] = omega^(j * 2^s), so that a pass
// reads each stage's twiddles contiguously; values are in the 261-form, so
// (u - v) * w keeps the data's 256-form without conversions. Stage entries are
// 48-byte records of the nine limbs themselves (no repacking after the load);
// the other tables hold 32-byte rows like the data. Coset powers and
// n^-1 are tables in the same form, applied as the first pass loads and the
// last pass stores. Tables are built on the device from two short seed tables
// (host powers of omega) and cached per (field, log n, omega, coset).
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/zkgpu.h"
#include "zk_fields.hpp"
#include "zk_host_math.hpp"
#include "zk_ntt.hpp"

namespace zk {

// Rows per workgroup tile (2^kTileLog, staged in LDS as nine limb planes) and threads per
// workgroup; a pass fuses at most kTileLog stages.
#ifndef ZK_NTT_TILE_LOG
#define ZK_NTT_TILE_LOG 10
#endif
#ifndef ZK_NTT_THREADS
#define ZK_NTT_THREADS 256
#endif
constexpr int kTileLog = ZK_NTT_TILE_LOG;
constexpr int kTile = 1 << kTileLog;
constexpr int kThreads = ZK_NTT_THREADS;

__device__ __forceinline__ void lds_put(uint32_t* lds, int slot, const Fe& x) {
#pragma unroll
  for (int i = 0; i < NL; ++i) lds[i * kTile + slot] = x.l[i];
}

__device__ __forceinline__ Fe lds_get(const uint32_t* lds, int slot) {
  Fe x;
#pragma unroll
  for (int i = 0; i < NL; ++i) x.l[i] = lds[i * kTile + slot];
  return x;
}

// Stage-twiddle records: nine 29-bit limbs and three zero words, 48 bytes (16-byte aligned).
constexpr int kTwWords = 6;  // 64-bit words per record

__device__ __forceinline__ Fe load_tw(const uint64_t* tab, int64_t j) {
  const GlobalPtr<const U32x4> p = global_ptr<U32x4>(tab + kTwWords * j);
  const U32x4 a = p[0], b = p[1];
  const uint32_t c = global_ptr<uint32_t>(tab + kTwWords * j + 4)[0];
  Fe x;
#pragma unroll
  for (int i = 0; i < 4; ++i) x.l[i] = a[i], x.l[4 + i] = b[i];
  x.l[8] = c;
  return x;
}

__device__ __forceinline__ int64_t bitrev(int64_t i, int bits) {
  return (int64_t)(__builtin_bitreverse64((uint64_t)i) >> (64 - bits));
}

// Row index of (group, m) in this pass: group -> (hi, lo), i = hi*2^(L-s0) + m*S + lo.
struct PassGeom {
  int log_n, s0, k, log_c;
  int last;
  __device__ __forceinline__ int64_t row(int64_t wg, int g, int m) const {
    const int log_s = log_n - s0 - k;
    int64_t hi, lo;
    if (last) {  // S == 1: groups w + g * 2^(L-k-c) so that bit-reversed writes are contiguous
      hi = wg + ((int64_t)g << (log_n - k - log_c));
      lo = 0;
    } else {
      int64_t gid = (wg << log_c) + g;
      hi = gid >> log_s;
      lo = gid & ((1LL << log_s) - 1);
    }
    return (hi << (log_n - s0)) + ((int64_t)m << log_s) + lo;
  }
};

// Everything one workgroup of a pass needs (per transform in a batch).
struct NttPassDesc {
  const uint64_t* src;
  uint64_t* dst;
  const uint64_t* tw;
  const uint64_t* pre;
  const uint64_t* post;
  int post_is_table;
  PassGeom geo;
  QuotientArgs q;
  int64_t tile_base;  // first workgroup of this transform in a batched launch
};

template <class F>
__device__ __forceinline__ void ntt_tile(const F& f, const NttPassDesc& d, int64_t wg, uint32_t* lds) {
  const PassGeom geo = d.geo;
  const int C = 1 << geo.log_c;
  const int T = C << geo.k;
  const int64_t n = 1LL << geo.log_n;
  const int log_s = geo.log_n - geo.s0 - geo.k;

  for (int e = threadIdx.x; e < T; e += kThreads) {
    const int g = e & (C - 1), m = e >> geo.log_c;
    const int64_t i = geo.row(wg, g, m);
    Fe x;
    if (d.q.a) {
      // Groth16 numerator on the coset: (A*B' - C) * z^-1, with B' carrying an extra 2^5
      // so that mont(A, B') is in the 256-form.
      Fe av = load_row(d.q.a + 4 * i), bv = load_row(d.q.b + 4 * i), cv = load_row(d.q.c + 4 * i);
      Fe zi = load_row(d.q.z_inv);
      x = mont(f, fsub(f, mont(f, av, bv), cv), zi);
    } else {
      x = load_row(d.src + 4 * i);
    }
    if (d.pre) x = mont(f, x, load_row(d.pre + 4 * i));
    lds_put(lds, e, x);
  }
  __syncthreads();

  // Stages in pairs (radix 4): a thread takes rows m, m + H/2, m + H, m + 3H/2 of a group through
  // stage t (pairs at distance H = 2^(k-1-t)) and stage t + 1 (distance H/2) in registers:
  // 4 products, 3 twiddles, one LDS round trip and one barrier for two stages.
  int t = 0;
  for (; t + 1 < geo.k; t += 2) {
    const int hb = geo.k - 1 - t;  // >= 1
    const int h2 = 1 << (hb - 1);
    const int s = geo.s0 + t;
    const uint64_t* tw_a = d.tw + kTwWords * (n - (n >> s));        // stage s
    const uint64_t* tw_b = d.tw + kTwWords * (n - (n >> (s + 1)));  // stage s + 1
    const bool last_b = s + 1 == geo.log_n - 1;             // stage s + 1 has the single twiddle 1
    for (int u = threadIdx.x; u < (T >> 2); u += kThreads) {
      const int g = u & (C - 1), qv = u >> geo.log_c;
      const int mlow = qv & (h2 - 1);
      const int m00 = ((qv >> (hb - 1)) << (hb + 1)) | mlow;
      const int64_t lo = geo.row(wg, g, m00) & ((1LL << log_s) - 1);
      const int64_t ja = ((int64_t)mlow << log_s) + lo, jb = ((int64_t)(mlow + h2) << log_s) + lo;
      const int sl0 = (m00 << geo.log_c) + g, step = h2 << geo.log_c;
      const Fe x0 = lds_get(lds, sl0), x1 = lds_get(lds, sl0 + step), x2 = lds_get(lds, sl0 + 2 * step),
               x3 = lds_get(lds, sl0 + 3 * step);
      const Fe y2 = mont_raw(f, fsub_for_mul(f, x0, x2), load_tw(tw_a, ja));
      const Fe y3 = mont_raw(f, fsub_for_mul(f, x1, x3), load_tw(tw_a, jb));
      if constexpr (HasP4r<F>::value) {
        lds_put(lds, sl0 + 2 * step, fadd_wide_finish(f, y2, y3));
        // the first stage's sums stay wide (below 4p, no carries) and are reduced once, in the
        // second stage's sum, difference or product
        const Fe y0 = fadd_wide(x0, x2), y1 = fadd_wide(x1, x3);
        lds_put(lds, sl0, fadd_wide_finish(f, y0, y1));
        if (last_b) {
          lds_put(lds, sl0 + step, fsub_wide_finish(f, y0, y1));
          lds_put(lds, sl0 + 3 * step, fsub_lazy(f, y2, y3));
        } else {
          const Fe wc = load_tw(tw_b, ja);
          lds_put(lds, sl0 + step, mont_raw(f, fsub_wide(f, y0, y1), wc));
          lds_put(lds, sl0 + 3 * step, mont_raw(f, fsub_for_mul(f, y2, y3), wc));
        }
      } else {
        lds_put(lds, sl0 + 2 * step, fadd_lazy(f, y2, y3));
        const Fe y0 = fadd_lazy(f, x0, x2), y1 = fadd_lazy(f, x1, x3);
        lds_put(lds, sl0, fadd_lazy(f, y0, y1));
        if (last_b) {
          lds_put(lds, sl0 + step, fsub_lazy(f, y0, y1));
          lds_put(lds, sl0 + 3 * step, fsub_lazy(f, y2, y3));
        } else {
          const Fe wc = load_tw(tw_b, ja);
          lds_put(lds, sl0 + step, mont_raw(f, fsub_for_mul(f, y0, y1), wc));
          lds_put(lds, sl0 + 3 * step, mont_raw(f, fsub_for_mul(f, y2, y3), wc));
        }
      }
    }
    __syncthreads();
  }
  for (; t < geo.k; ++t) {  // an odd stage count ends with one radix-2 stage
    const int hb = geo.k - 1 - t;  // bit of m that this stage pairs
    const int s = geo.s0 + t;
    const uint64_t* tws = d.tw + kTwWords * (n - (n >> s));  // stage s table: n >> (s + 1) entries
    const bool last_stage = s == geo.log_n - 1;        // its only twiddle is 1
    for (int b = threadIdx.x; b < (T >> 1); b += kThreads) {
      const int g = b & (C - 1), qv = b >> geo.log_c;
      const int mlow = qv & ((1 << hb) - 1);
      const int m0 = ((qv >> hb) << (hb + 1)) | mlow;
      const int m1 = m0 | (1 << hb);
      const int64_t row0 = geo.row(wg, g, m0);
      const int64_t j = ((int64_t)mlow << log_s) + (row0 & ((1LL << log_s) - 1));
      const int s0slot = (m0 << geo.log_c) + g, s1slot = (m1 << geo.log_c) + g;
      Fe u = lds_get(lds, s0slot), v = lds_get(lds, s1slot);
      lds_put(lds, s0slot, fadd_lazy(f, u, v));
      if (last_stage) {
        lds_put(lds, s1slot, fsub_lazy(f, u, v));  // w = omega^0 = 1
      } else {
        Fe w = load_tw(tws, j);
        lds_put(lds, s1slot, mont_raw(f, fsub_for_mul(f, u, v), w));
      }
    }
    __syncthreads();
  }

  for (int e = threadIdx.x; e < T; e += kThreads) {
    const int g = e & (C - 1), m = e >> geo.log_c;
    const int64_t i = geo.row(wg, g, m);
    Fe x = lds_get(lds, e);  // below 2p: passes hand each other lazily reduced rows
    const int64_t o = geo.last ? bitrev(i, geo.log_n) : i;
    if (d.post) x = mont(f, x, load_row(d.post + (d.post_is_table ? 4 * o : 0)));
    else if (geo.last) x = canonical(f, x);
    store_row(d.dst + 4 * o, x);
  }
}

template <class F>
__global__ void __launch_bounds__(kThreads) k_ntt_pass(F f, NttPassDesc d) {
  __shared__ uint32_t lds[NL * kTile];
  ntt_tile(f, d, blockIdx.x, lds);
}

// A pass of several transforms at once: workgroup -> transform by binary search over tile_base.
template <class F>
__global__ void __launch_bounds__(kThreads) k_ntt_pass_multi(F f, const NttPassDesc* __restrict__ ds, int nd) {
  __shared__ uint32_t lds[NL * kTile];
  int lo = 0, hi = nd;
  const int64_t wg = blockIdx.x;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ds[mid].tile_base <= wg) lo = mid;
    else hi = mid;
  }
  const NttPassDesc d = ds[lo];
  ntt_tile(f, d, wg - d.tile_base, lds);
}

// Table expansion: out[j] = seed_lo[e & mask] * seed_hi[e >> h] for e = j << shift, in the 261-form.
// seed tables hold plain values; c522 lifts them to the 261-form. Entries are 32-byte rows, or
// 48-byte stage-twiddle records when records is set.
template <class F>
__global__ void __launch_bounds__(256) k_expand(F f, const uint64_t* __restrict__ lo_tab, const uint64_t* __restrict__ hi_tab,
                                                int h, int shift, uint64_t* __restrict__ out, int64_t count,
                                                const uint64_t* __restrict__ scale, int records) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int64_t e = j << shift;
  const Fe lift = fe_const<F>(f.c522);
  Fe a = mont(f, load_row(lo_tab + 4 * (e & ((1LL << h) - 1))), lift);
  Fe b = mont(f, load_row(hi_tab + 4 * (e >> h)), lift);
  Fe x = mont(f, a, b);
  if (scale) x = mont(f, x, mont(f, load_row(scale), lift));
  if (records) {
    const GlobalPtr<U32x4> p = global_ptr<U32x4>(out + kTwWords * j);
    p[0] = U32x4{x.l[0], x.l[1], x.l[2], x.l[3]};
    p[1] = U32x4{x.l[4], x.l[5], x.l[6], x.l[7]};
    p[2] = U32x4{x.l[8], 0u, 0u, 0u};
  } else {
    store_row(out + 4 * j, x);
  }
}

// ---- host side -----------------------------------------------------------

namespace {

struct TableKey {
  int field, log_n, kind;  // 0: stage twiddles; 1: coset powers; 2: inverse post-scale table;
                           // 3: one row extra * n^-1; 4: one row holding the constant a
  uint64_t a[4], b[4];
  bool operator<(const TableKey& o) const {
    return std::tie(field, log_n, kind, a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]) <
           std::tie(o.field, o.log_n, o.kind, o.a[0], o.a[1], o.a[2], o.a[3], o.b[0], o.b[1], o.b[2], o.b[3]);
  }
  bool operator==(const TableKey& o) const { return !(*this < o) && !(o < *this); }
};

struct Table {
  uint64_t* ptr;
  size_t rows;
  size_t bytes;
  int uses;  // pins held by plans and by calls whose kernels are not yet enqueued
};

std::mutex g_tab_mu;
std::map<TableKey, Table> g_tables;
std::list<TableKey> g_lru;  // most recently used first
size_t g_table_bytes = 0;
// Device bytes the cache keeps before it evicts unpinned tables, least recently used first.
constexpr size_t kTableCacheBytes = (size_t)4 << 30;
thread_local TableHold* g_hold = nullptr;

// Frees unpinned tables, least recently used first, until the cache holds at most `limit` bytes
// (0: every unpinned table). Waits for the device first: kernels already enqueued may still read
// any table. Called with g_tab_mu held.
int evict_locked(size_t limit) {
  bool synced = false;
  for (auto it = g_lru.end(); it != g_lru.begin() && g_table_bytes > limit;) {
    --it;
    auto t = g_tables.find(*it);
    if (t == g_tables.end() || t->second.uses > 0) continue;
    if (!synced) {
      ZK_HIP(hipDeviceSynchronize());
      synced = true;
    }
    (void)hipFree(t->second.ptr);
    g_table_bytes -= t->second.bytes;
    g_tables.erase(t);
    it = g_lru.erase(it);
  }
  return 0;
}

bool field_modulus(int field_id, U256* p) {
  static const uint64_t r[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
  static const uint64_t q[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
  if (field_id == 0) { *p = U256::from(r); return true; }
  if (field_id == 1) { *p = U256::from(q); return true; }
  FieldRT f;
  if (!lookup_field(field_id, &f)) return false;
  for (int i = 0; i < 4; ++i) p->w[i] = (uint64_t)f.p32[2 * i] | ((uint64_t)f.p32[2 * i + 1] << 32);
  return true;
}

// Expands count entries of base^(j << shift) (times scale, if given) into out, on the device.
int expand(int field_id, const HostField& hf, const U256& base_plain, int log_count, int shift, uint64_t* out, int64_t count,
           const U256* scale_plain, bool records, hipStream_t stream) {
  const int bits = log_count + shift;  // exponents are below 2^bits
  const int h = (bits + 1) / 2;
  const int64_t nlo = 1LL << h, nhi = (1LL << (bits - h)) + 1;
  std::vector<uint64_t> host((size_t)(nlo + nhi + 1) * 4);
  U256 bm = hf.to_mont(base_plain), acc = hf.one();
  for (int64_t i = 0; i < nlo; ++i) {
    U256 v = hf.from_mont(acc);
    memcpy(&host[(size_t)i * 4], v.w, 32);
    acc = hf.mul(acc, bm);
  }
  U256 step = acc;  // base^(2^h)
  acc = hf.one();
  for (int64_t i = 0; i < nhi; ++i) {
    U256 v = hf.from_mont(acc);
    memcpy(&host[(size_t)(nlo + i) * 4], v.w, 32);
    acc = hf.mul(acc, step);
  }
  if (scale_plain) memcpy(&host[(size_t)(nlo + nhi) * 4], scale_plain->w, 32);
  Scratch sc(stream);
  uint64_t* dev = sc.get<uint64_t>(host.size());
  if (!dev) return 1;
  ZK_HIP(hipMemcpyAsync(dev, host.data(), host.size() * 8, hipMemcpyHostToDevice, stream));
  int rc = with_field(field_id, [&](auto f) {
    hipLaunchKernelGGL((k_expand<decltype(f)>), dim3(grid_for(count, 256)), dim3(256), 0, stream, f, dev, dev + 4 * nlo, h,
                       shift, out, count, scale_plain ? dev + 4 * (nlo + nhi) : nullptr, records ? 1 : 0);
    ZK_LAUNCH_CHECK("table expansion");
    return 0;
  });
  if (rc) return rc;
  ZK_HIP(hipStreamSynchronize(stream));  // the host seed buffer is released on return
  return 0;
}

// A cached table's address (moved to the front of the LRU list and pinned by the thread's hold), or
// false if the key is not cached. Called with g_tab_mu held.
bool lookup_locked(const TableKey& key, const uint64_t** out) {
  auto it = g_tables.find(key);
  if (it == g_tables.end()) return false;
  g_lru.remove(key);
  g_lru.push_front(key);
  if (g_hold) {
    ++it->second.uses;
    g_hold->pin(it->second.ptr);
  }
  *out = it->second.ptr;
  return true;
}

// Adds a freshly built table (g_tab_mu held).
void insert_locked(const TableKey& key, uint64_t* buf, size_t rows, size_t bytes) {
  g_tables[key] = Table{buf, rows, bytes, g_hold ? 1 : 0};
  g_table_bytes += bytes;
  g_lru.push_front(key);
  if (g_hold) g_hold->pin(buf);
}

// Returns a cached table, building it on first use.
int get_table(const TableKey& key, hipStream_t stream, const uint64_t** out) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  if (lookup_locked(key, out)) return 0;
  U256 p;
  if (!field_modulus(key.field, &p)) return fail("unknown field id");
  HostField hf(p);
  const int64_t n = 1LL << key.log_n;
  size_t rows = key.kind == 0 ? (size_t)(n > 1 ? n - 1 : 1) : key.kind >= 3 ? 1 : (size_t)n;
  uint64_t* buf = nullptr;
  const size_t bytes = rows * (key.kind == 0 ? 8 * kTwWords : 32);
  if (int rc = evict_locked(kTableCacheBytes > bytes ? kTableCacheBytes - bytes : 0)) return rc;
  ZK_HIP(hipMalloc(&buf, bytes));
  int rc = 0;
  if (key.kind == 0) {
    U256 omega = U256::from(key.a);
    for (int s = 0; s < key.log_n && rc == 0; ++s) {
      const int64_t off = n - (n >> s), cnt = n >> (s + 1);
      rc = expand(key.field, hf, omega, key.log_n - 1 - s, s, buf + kTwWords * off, cnt, nullptr, true, stream);
    }
  } else if (key.kind == 1) {
    rc = expand(key.field, hf, U256::from(key.b), key.log_n, 0, buf, n, nullptr, false, stream);
  } else if (key.kind == 4) {
    U256 value = hf.reduce(U256::from(key.a));
    rc = expand(key.field, hf, U256::of(1), 0, 0, buf, 1, &value, false, stream);
  } else if (key.kind == 3) {
    U256 scale = hf.mulp(hf.invp(U256::of((uint64_t)n)), U256::from(key.a));
    rc = expand(key.field, hf, U256::of(1), 0, 0, buf, 1, &scale, false, stream);
  } else {
    // extra * n^-1 * g^-j
    U256 g = U256::from(key.b);
    U256 ginv = hf.invp(g);
    U256 ninv = hf.invp(U256::of((uint64_t)n));
    U256 scale = hf.mulp(ninv, U256::from(key.a));  // a carries an extra factor (1 or 2^5)
    rc = expand(key.field, hf, ginv, key.log_n, 0, buf, n, &scale, false, stream);
  }
  if (rc) {
    (void)hipFree(buf);
    return rc;
  }
  insert_locked(key, buf, rows, bytes);
  *out = buf;
  return 0;
}

// Drops one pin per listed table (g_tab_mu held).
void unpin_locked(const std::vector<const uint64_t*>& pins) {
  for (const uint64_t* p : pins)
    for (auto& kv : g_tables)
      if (kv.second.ptr == p) {
        if (kv.second.uses > 0) --kv.second.uses;
        break;
      }
}

struct PassPlan {
  int s0, k, log_c;
};


// Stages fused per pass, per transform size (0: the default of 8); set by zk_ntt_set_fusion.
int g_fusion[29] = {0};

std::vector<PassPlan> plan_passes(int log_n) {
  std::vector<PassPlan> out;
  const int max_k = g_fusion[log_n] > 0 ? g_fusion[log_n] : 8;
  if (log_n <= kTileLog && max_k >= log_n) {
    out.push_back({0, log_n, 0});
    return out;
  }
  int passes = (log_n + max_k - 1) / max_k;
  int s0 = 0;
  for (int p = 0; p < passes; ++p) {
    int k = (log_n - s0) / (passes - p);
    if ((log_n - s0) % (passes - p)) ++k;
    int log_c = kTileLog - k;
    if (log_c > log_n - k) log_c = log_n - k;  // no more groups than exist
    out.push_back({s0, k, log_c});
    s0 += k;
  }
  return out;
}

}  // namespace

TableHold::TableHold() : prev_(g_hold) { g_hold = this; }

TableHold::~TableHold() {
  g_hold = prev_;
  if (!pins_.empty()) tables_release(pins_);
}

std::vector<const uint64_t*> TableHold::take() {
  std::vector<const uint64_t*> out;
  out.swap(pins_);
  return out;
}

void tables_release(const std::vector<const uint64_t*>& pins) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  unpin_locked(pins);
}

// Frees every table no plan holds (pinned ones stay cached and valid).
int ntt_cache_clear() {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  return evict_locked(0);
}

int ntt_cache_stats(uint64_t* tables, uint64_t* bytes, uint64_t* pinned) {
  std::lock_guard<std::mutex> lk(g_tab_mu);
  uint64_t p = 0;
  for (const auto& kv : g_tables) p += kv.second.uses > 0 ? 1 : 0;
  *tables = g_tables.size();
  *bytes = g_table_bytes;
  *pinned = p;
  return 0;
}

int twiddles(int field_id, int log_n, const uint64_t omega[4], hipStream_t stream, const uint64_t** out) {
  TableKey key{field_id, log_n, 0, {omega[0], omega[1], omega[2], omega[3]}, {0, 0, 0, 0}};
  return get_table(key, stream, out);
}

int coset_table(int field_id, int log_n, const uint64_t g[4], bool inverse, const uint64_t extra[4], hipStream_t stream,
                const uint64_t** out) {
  TableKey key{field_id, log_n, inverse ? 2 : 1, {extra[0], extra[1], extra[2], extra[3]}, {g[0], g[1], g[2], g[3]}};
  return get_table(key, stream, out);
}

int scale_row(int field_id, int log_n, const uint64_t extra[4], hipStream_t stream, const uint64_t** out) {
  TableKey key{field_id, log_n, 3, {extra[0], extra[1], extra[2], extra[3]}, {0, 0, 0, 0}};
  return get_table(key, stream, out);
}

int const_row(int field_id, const uint64_t value[4], hipStream_t stream, const uint64_t** out) {
  TableKey key{field_id, 0, 4, {value[0], value[1], value[2], value[3]}, {0, 0, 0, 0}};
  return get_table(key, stream, out);
}

int host_table(int log_n, const uint64_t tag[4], size_t words, const std::function<void(uint64_t*)>& build,
               hipStream_t stream, const uint64_t** out) {
  TableKey key{-1, log_n, 5, {tag[0], tag[1], tag[2], tag[3]}, {0, 0, 0, 0}};
  std::lock_guard<std::mutex> lk(g_tab_mu);
  if (lookup_locked(key, out)) return 0;
  std::vector<uint64_t> host(words);
  build(host.data());
  const size_t bytes = words * 8;
  if (int rc = evict_locked(kTableCacheBytes > bytes ? kTableCacheBytes - bytes : 0)) return rc;
  uint64_t* buf = nullptr;
  ZK_HIP(hipMalloc(&buf, bytes));
  hipError_t err = hipMemcpyAsync(buf, host.data(), bytes, hipMemcpyHostToDevice, stream);
  if (err == hipSuccess) err = hipStreamSynchronize(stream);  // the host copy is released on return
  if (err != hipSuccess) {
    (void)hipFree(buf);
    return fail_hip(err, "table upload");
  }
  insert_locked(key, buf, words, bytes);
  *out = buf;
  return 0;
}

// Descriptors of every pass of one transform (tmp: n rows of scratch when there are several passes).
std::vector<NttPassDesc> pass_descs(const uint64_t* in, uint64_t* out, uint64_t* tmp, int log_n, const uint64_t* tw,
                                    const uint64_t* pre, const uint64_t* post, bool post_table, const QuotientArgs& q) {
  std::vector<PassPlan> plan = plan_passes(log_n);
  std::vector<NttPassDesc> ds;
  for (size_t p = 0; p < plan.size(); ++p) {
    const bool first = p == 0, last = p + 1 == plan.size();
    NttPassDesc d{};
    d.geo = PassGeom{log_n, plan[p].s0, plan[p].k, plan[p].log_c, last ? 1 : 0};
    d.src = first ? in : tmp;
    d.dst = last ? out : tmp;
    d.tw = tw;
    d.pre = first ? pre : nullptr;
    d.post = last ? post : nullptr;
    d.post_is_table = post_table ? 1 : 0;
    d.q = first ? q : QuotientArgs{};
    ds.push_back(d);
  }
  return ds;
}

int64_t pass_tiles(const NttPassDesc& d) { return 1LL << (d.geo.log_n - d.geo.k - d.geo.log_c); }

int ntt_run(int field_id, const uint64_t* in, uint64_t* out, int log_n, const uint64_t* tw, const uint64_t* pre,
            const uint64_t* post, bool post_table, const QuotientArgs& q, hipStream_t stream) {
  const int64_t n = 1LL << log_n;
  Scratch sc(stream);
  uint64_t* tmp = nullptr;
  if (plan_passes(log_n).size() > 1) {
    tmp = sc.get<uint64_t>((size_t)n * 4);
    if (!tmp) return 1;
  }
  std::vector<NttPassDesc> ds = pass_descs(in, out, tmp, log_n, tw, pre, post, post_table, q);
  return with_field(field_id, [&](auto f) {
    for (const NttPassDesc& d : ds) {
      launch_timed("ntt_pass", k_ntt_pass<decltype(f)>, dim3((unsigned)pass_tiles(d)), dim3(kThreads), 0, stream, f, d);
      ZK_LAUNCH_CHECK("ntt pass");
    }
    return 0;
  });
}

// ---- batched transforms ---------------------------------------------------------

int64_t transform_scratch_rows(int log_n) { return plan_passes(log_n).size() > 1 ? (1LL << log_n) : 0; }

int batch_plan_create(const std::vector<TransformSpec>& specs, hipStream_t stream, BatchPlan** result) {
  int stages = 0;
  for (const TransformSpec& t : specs) {
    if (t.log_n < 1 || t.log_n > 28) return fail("batched transform sizes must be 2^1..2^28");
    if (t.stage < 0) return fail("negative transform stage");
    stages = std::max(stages, t.stage + 1);
  }
  // pass scratch: transforms of one stage run together, stages one after another
  std::vector<size_t> stage_rows((size_t)stages, 0);
  for (const TransformSpec& t : specs) stage_rows[(size_t)t.stage] += (size_t)transform_scratch_rows(t.log_n);
  size_t tmp_rows = 0;
  for (size_t r : stage_rows) tmp_rows = std::max(tmp_rows, r);
  // descriptors per (stage, pass index)
  std::vector<std::vector<std::vector<NttPassDesc>>> launches((size_t)stages);
  std::vector<size_t> tmp_at((size_t)stages, 0);
  size_t desc_count = 0;
  const size_t tmp_bytes = (tmp_rows * 32 + 255) & ~(size_t)255;
  auto plan = new BatchPlan();
  std::vector<std::vector<NttPassDesc>> per;  // filled below once the scratch base is known
  for (const TransformSpec& t : specs) {
    per.push_back(pass_descs(t.in, t.out, nullptr, t.log_n, t.tw, t.pre, t.post, t.post_table, t.q));
    desc_count += per.back().size();
  }
  plan->bytes = tmp_bytes + desc_count * sizeof(NttPassDesc) + 256;
  plan->mem = dev_alloc(plan->bytes, stream);
  if (!plan->mem) {
    delete plan;
    return 1;
  }
  plan->alloc_stream = stream;
  uint64_t* tmp = static_cast<uint64_t*>(plan->mem);
  NttPassDesc* dtab = reinterpret_cast<NttPassDesc*>(static_cast<char*>(plan->mem) + tmp_bytes);
  for (size_t j = 0; j < specs.size(); ++j) {
    const TransformSpec& t = specs[j];
    auto& v = per[j];
    auto& st = launches[(size_t)t.stage];
    if (v.size() > 1) {
      uint64_t* mine = tmp + 4 * tmp_at[(size_t)t.stage];
      for (size_t p = 0; p < v.size(); ++p) {
        if (p > 0) v[p].src = mine;
        if (p + 1 < v.size()) v[p].dst = mine;
      }
      tmp_at[(size_t)t.stage] += (size_t)1 << t.log_n;
    }
    if (st.size() < v.size()) st.resize(v.size());
    for (size_t p = 0; p < v.size(); ++p) st[p].push_back(v[p]);
  }
  plan->host_desc.resize(desc_count * sizeof(NttPassDesc));
  NttPassDesc* host = reinterpret_cast<NttPassDesc*>(plan->host_desc.data());
  size_t at = 0;
  for (auto& st : launches) {
    for (auto& ds : st) {
      int64_t tiles = 0;
      for (NttPassDesc& d : ds) {
        d.tile_base = tiles;
        tiles += pass_tiles(d);
        host[at + (size_t)(&d - ds.data())] = d;
      }
      plan->counts.push_back((int)ds.size());
      plan->tiles.push_back(tiles);
      plan->d_desc.push_back(dtab + at);
      at += ds.size();
    }
  }
  // Stream-ordered upload from the plan's own host copy, which lives as long as the plan.
  if (desc_count) {
    hipError_t e = hipMemcpyAsync(dtab, host, desc_count * sizeof(NttPassDesc), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) {
      dev_free(plan->mem);
      delete plan;
      return fail_hip(e, "batched transform plan upload");
    }
  }
  *result = plan;
  return 0;
}

int batch_plan_launch(BatchPlan* plan, hipStream_t stream) {
  if (launch_escapes(plan->alloc_stream, stream)) plan->escaped = true;
  const FrConst f;
  for (size_t i = 0; i < plan->counts.size(); ++i) {
    launch_timed("ntt_pass_multi", k_ntt_pass_multi<FrConst>, dim3((unsigned)plan->tiles[i]), dim3(kThreads), 0, stream, f,
                 static_cast<const NttPassDesc*>(plan->d_desc[i]), plan->counts[i]);
    ZK_LAUNCH_CHECK("batched ntt pass");
  }
  return 0;
}

void batch_plan_destroy(BatchPlan* plan) {
  if (!plan) return;
  if (plan->escaped) (void)hipDeviceSynchronize();  // work the allocating stream does not order
  dev_free(plan->mem);
  for (void* p : plan->owned) dev_free(p);
  tables_release(plan->tables);
  delete plan;
}

int ntt_plan_create(int count, const uint64_t* const* in, uint64_t* const* out, const int32_t* log_n,
                    const int32_t* inverse, hipStream_t stream, BatchPlan** result) {
  TableHold hold;
  std::vector<TransformSpec> specs;
  const uint64_t one[4] = {1, 0, 0, 0};
  for (int j = 0; j < count; ++j) {
    if (log_n[j] < 1 || log_n[j] > 28) return fail("batched NTT sizes must be 2^1..2^28");
    U256 w, wi;
    fr_roots(log_n[j], &w, &wi);
    TransformSpec t;
    t.in = in[j];
    t.out = out[j];
    t.log_n = log_n[j];
    if (int rc = twiddles(0, log_n[j], inverse[j] ? wi.w : w.w, stream, &t.tw)) return rc;
    if (inverse[j]) {
      if (int rc = scale_row(0, log_n[j], one, stream, &t.post)) return rc;
    }
    specs.push_back(t);
  }
  if (int rc = batch_plan_create(specs, stream, result)) return rc;
  (*result)->tables = hold.take();
  return 0;
}

}  // namespace zk

using namespace zk;

extern "C" {

int zk_ntt_cache_clear(void) { return ntt_cache_clear(); }

int zk_ntt_cache_stats(uint64_t* tables, uint64_t* bytes, uint64_t* pinned) { return ntt_cache_stats(tables, bytes, pinned); }

int zk_ntt_set_fusion(int log_n, int max_stages) {
  if (log_n < 1 || log_n > 28) return fail("log_n must be in [1, 28]");
  if (max_stages < 0 || max_stages > kTileLog)
    return fail("stages per pass must be in 1.." + std::to_string(kTileLog) + " (0 = default)");
  g_fusion[log_n] = max_stages;
  return 0;
}

int zk_ntt_get_fusion(int log_n, int* max_stages) {
  if (log_n < 1 || log_n > 28) return fail("log_n must be in [1, 28]");
  *max_stages = g_fusion[log_n] > 0 ? g_fusion[log_n] : 8;
  return 0;
}

int zk_ntt_plan_create(int count, const uint64_t* const* in, uint64_t* const* out, const int32_t* log_n,
                       const int32_t* inverse, void** plan, void* stream) {
  BatchPlan* p = nullptr;
  if (int rc = ntt_plan_create(count, in, out, log_n, inverse, (hipStream_t)stream, &p)) return rc;
  *plan = p;
  return 0;
}

int zk_ntt_plan_launch(void* plan, void* stream) { return batch_plan_launch(static_cast<BatchPlan*>(plan), (hipStream_t)stream); }

int zk_ntt_plan_destroy(void* plan) {
  batch_plan_destroy(static_cast<BatchPlan*>(plan));
  return 0;
}

int zk_ntt_batch(int count, const uint64_t* const* in, uint64_t* const* out, const int32_t* log_n, const int32_t* inverse,
                 void* stream) {
  void* plan = nullptr;
  if (int rc = zk_ntt_plan_create(count, in, out, log_n, inverse, &plan, stream)) return rc;
  int rc = zk_ntt_plan_launch(plan, stream);
  zk_ntt_plan_destroy(plan);
  return rc;
}

int zk_ntt(int field_id, const uint64_t* in, uint64_t* out, int log_n, int inverse, const uint64_t omega[4],
           const uint64_t* coset, void* stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  if (log_n < 0 || log_n > 28) return fail("log_n must be in [0, 28]");
  const int64_t n = 1LL << log_n;
  if (log_n == 0) {
    // A single value is its own transform; the inverse scales by 1^-1 = 1 and g^0 = 1.
    if (in != out) ZK_HIP(hipMemcpyAsync(out, in, 32, hipMemcpyDeviceToDevice, stream));
    return 0;
  }
  TableHold hold;  // the tables stay pinned until the passes are enqueued
  const uint64_t* tw = nullptr;
  if (int rc = twiddles(field_id, log_n, omega, stream, &tw)) return rc;
  const uint64_t* pre = nullptr;
  const uint64_t* post = nullptr;
  const uint64_t one[4] = {1, 0, 0, 0};
  if (!inverse && coset) {
    if (int rc = coset_table(field_id, log_n, coset, false, one, stream, &pre)) return rc;
  }
  bool post_table = false;
  if (inverse && coset) {
    if (int rc = coset_table(field_id, log_n, coset, true, one, stream, &post)) return rc;
    post_table = true;
  } else if (inverse) {
    if (int rc = scale_row(field_id, log_n, one, stream, &post)) return rc;
  }
  (void)n;
  return ntt_run(field_id, in, out, log_n, tw, pre, post, post_table, QuotientArgs{}, stream);
}

}  // extern "C"
