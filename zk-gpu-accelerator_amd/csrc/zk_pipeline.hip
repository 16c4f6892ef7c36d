// The Groth16 quotient pipeline and the batched NTT / MSM entry points.
//
// Quotient (demo/prover_simulation.py:142-150, domain.py:112-135): for each of
// A, B, C one inverse NTT whose last pass multiplies by n^-1 * g^j (turning
// the coefficients into those of f(g x)), then a forward NTT: the coset
// evaluations. B's table also carries 2^5, so that one Montgomery product of
// A' and B' lands in the 256-form. The final inverse NTT's first pass loads
// (A' * B' - C') * (g^n - 1)^-1 directly from the three vectors, and its last
// pass multiplies by n^-1 * g^-j. Seven transforms, 21 HBM sweeps at 2^22,
// with no separate pointwise pass.
#include <vector>

#include "../../include/zkgpu.h"
#include "zk_fields.hpp"
#include "zk_host_math.hpp"
#include "zk_msm.hpp"
#include "zk_ntt.hpp"

namespace zk {

static const uint64_t kFrModulus[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL,
                                       0x30644e72e131a029ULL};

void fr_roots(int log_n, U256* omega, U256* omega_inv) {
  HostField hf(U256::from(kFrModulus));
  U256 e = HostField::sub_raw(U256::from(kFrModulus), U256::of(1));  // (r - 1) >> log_n
  for (int i = 0; i < log_n; ++i) {
    for (int w = 0; w < 4; ++w) e.w[w] = (e.w[w] >> 1) | (w < 3 ? e.w[w + 1] << 63 : 0);
  }
  *omega = hf.powp(U256::of(5), e);
  *omega_inv = hf.invp(*omega);
}

// Cached tables of one quotient of size 2^log_n on the coset g H.
struct QuotientTables {
  const uint64_t *tw_f, *tw_i, *post_a, *post_b, *post_h, *zrow;
};

int quotient_tables(int log_n, const uint64_t shift[4], hipStream_t stream, QuotientTables* t) {
  if (log_n < 1 || log_n > 28) return fail("log_n must be in [1, 28]");
  const int64_t n = 1LL << log_n;
  HostField hf(U256::from(kFrModulus));
  U256 g = hf.reduce(U256::from(shift));
  if (g.is_zero()) return fail("coset shift must be non-zero");
  U256 omega, omega_inv;
  fr_roots(log_n, &omega, &omega_inv);
  U256 gn = hf.powp(g, U256::of((uint64_t)n));
  U256 z = hf.sub(gn, U256::of(1));
  if (z.is_zero()) return fail("coset shift lies in the domain: g^n = 1");
  U256 z_inv = hf.invp(z);
  U256 g_inv = hf.invp(g);
  const uint64_t one[4] = {1, 0, 0, 0}, thirty_two[4] = {32, 0, 0, 0};
  if (int rc = twiddles(0, log_n, omega.w, stream, &t->tw_f)) return rc;
  if (int rc = twiddles(0, log_n, omega_inv.w, stream, &t->tw_i)) return rc;
  if (int rc = coset_table(0, log_n, g_inv.w, true, one, stream, &t->post_a)) return rc;         // n^-1 g^j
  if (int rc = coset_table(0, log_n, g_inv.w, true, thirty_two, stream, &t->post_b)) return rc;  // 2^5 n^-1 g^j
  if (int rc = coset_table(0, log_n, g.w, true, one, stream, &t->post_h)) return rc;             // n^-1 g^-j
  return const_row(0, z_inv.w, stream, &t->zrow);
}

int groth16_quotient(const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t* h, int log_n,
                     const uint64_t shift[4], hipStream_t stream) {
  TableHold hold;  // the tables stay pinned until every pass is enqueued
  QuotientTables t;
  if (int rc = quotient_tables(log_n, shift, stream, &t)) return rc;
  const int64_t n = 1LL << log_n;
  Scratch sc(stream);
  uint64_t* ev = sc.get<uint64_t>((size_t)n * 12);
  if (!ev) return 1;
  uint64_t* dst[3] = {ev, ev + 4 * n, ev + 8 * n};
  const uint64_t* src[3] = {a, b, c};
  const uint64_t* post[3] = {t.post_a, t.post_b, t.post_a};
  for (int k = 0; k < 3; ++k) {
    if (int rc = ntt_run(0, src[k], dst[k], log_n, t.tw_i, nullptr, post[k], true, QuotientArgs{}, stream)) return rc;
    if (int rc = ntt_run(0, dst[k], dst[k], log_n, t.tw_f, nullptr, nullptr, false, QuotientArgs{}, stream)) return rc;
  }
  QuotientArgs q{dst[0], dst[1], dst[2], t.zrow};
  return ntt_run(0, dst[0], h, log_n, t.tw_i, nullptr, t.post_h, true, q, stream);
}

// Device bytes a batched quotient over these sizes holds: three coset evaluation vectors per
// quotient, the pass scratch of the three transforms of a job that run side by side, and
// the descriptor tables.
uint64_t quotient_batch_bytes(int count, const int32_t* log_n) {
  uint64_t ev = 0, tmp = 0, descs = 0;
  for (int j = 0; j < count; ++j) {
    const uint64_t n = 1ULL << log_n[j];
    ev += 3 * n * 32;
    tmp += 3 * (uint64_t)transform_scratch_rows(log_n[j]) * 32;
    descs += 7 * 4 * 256;  // seven transforms of at most four passes, a generous bound per descriptor
  }
  return ev + tmp + descs + 1024;
}

// The quotient of `count` jobs of mixed sizes as one launch sequence: stage 0 runs the three
// inverse NTTs with the coset post-scale of every job, stage 1 the three forward NTTs, stage 2
// the final inverse NTT with the fused (A B - C) / Z load.
int quotient_plan_create(int count, const uint64_t* const* a, const uint64_t* const* b, const uint64_t* const* c,
                         uint64_t* const* h, const int32_t* log_n, const uint64_t* shifts, hipStream_t stream,
                         BatchPlan** out) {
  TableHold hold;  // handed to the plan, which keeps its tables pinned until destroyed
  std::vector<QuotientTables> tabs((size_t)std::max(count, 0));
  size_t ev_rows = 0;
  for (int j = 0; j < count; ++j) {
    if (int rc = quotient_tables(log_n[j], shifts + 4 * j, stream, &tabs[(size_t)j])) return rc;
    ev_rows += 3 * ((size_t)1 << log_n[j]);
  }
  uint64_t* ev = nullptr;
  if (ev_rows) {
    ev = static_cast<uint64_t*>(dev_alloc(ev_rows * 32, stream));
    if (!ev) return 1;
  }
  std::vector<TransformSpec> specs;
  size_t at = 0;
  for (int j = 0; j < count; ++j) {
    const QuotientTables& t = tabs[(size_t)j];
    const int64_t n = 1LL << log_n[j];
    uint64_t* dst[3] = {ev + 4 * at, ev + 4 * (at + n), ev + 4 * (at + 2 * n)};
    at += 3 * (size_t)n;
    const uint64_t* src[3] = {a[j], b[j], c[j]};
    const uint64_t* post[3] = {t.post_a, t.post_b, t.post_a};
    for (int k = 0; k < 3; ++k) {
      TransformSpec inv;
      inv.in = src[k];
      inv.out = dst[k];
      inv.log_n = log_n[j];
      inv.tw = t.tw_i;
      inv.post = post[k];
      inv.post_table = true;
      inv.stage = 0;
      specs.push_back(inv);
      TransformSpec fwd;
      fwd.in = dst[k];
      fwd.out = dst[k];
      fwd.log_n = log_n[j];
      fwd.tw = t.tw_f;
      fwd.stage = 1;
      specs.push_back(fwd);
    }
    TransformSpec fin;
    fin.in = dst[0];
    fin.out = h[j];
    fin.log_n = log_n[j];
    fin.tw = t.tw_i;
    fin.post = t.post_h;
    fin.post_table = true;
    fin.q = QuotientArgs{dst[0], dst[1], dst[2], t.zrow};
    fin.stage = 2;
    specs.push_back(fin);
  }
  BatchPlan* plan = nullptr;
  if (int rc = batch_plan_create(specs, stream, &plan)) {
    if (ev) dev_free(ev);
    return rc;
  }
  if (ev) {
    plan->owned.push_back(ev);
    plan->bytes += ev_rows * 32;
  }
  plan->tables = hold.take();
  *out = plan;
  return 0;
}

}  // namespace zk

using namespace zk;

extern "C" {

int zk_groth16_quotient(const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t* h, int log_n,
                        const uint64_t shift[4], void* stream) {
  return groth16_quotient(a, b, c, h, log_n, shift, (hipStream_t)stream);
}

int zk_groth16_plan_create(int count, const uint64_t* const* a, const uint64_t* const* b, const uint64_t* const* c,
                           uint64_t* const* h, const int32_t* log_n, const uint64_t* shifts, void** plan, void* stream) {
  if (count < 0) return fail("count must be non-negative");
  BatchPlan* p = nullptr;
  if (int rc = quotient_plan_create(count, a, b, c, h, log_n, shifts, (hipStream_t)stream, &p)) return rc;
  *plan = p;
  return 0;
}

int zk_groth16_plan_launch(void* plan, void* stream) {
  return batch_plan_launch(static_cast<BatchPlan*>(plan), (hipStream_t)stream);
}

int zk_groth16_plan_destroy(void* plan) {
  batch_plan_destroy(static_cast<BatchPlan*>(plan));
  return 0;
}

int zk_groth16_workspace_bytes(int count, const int32_t* log_n, uint64_t* bytes) {
  for (int j = 0; j < count; ++j)
    if (log_n[j] < 1 || log_n[j] > 28) return fail("log_n must be in [1, 28]");
  *bytes = quotient_batch_bytes(count, log_n);
  return 0;
}

}  // extern "C"
