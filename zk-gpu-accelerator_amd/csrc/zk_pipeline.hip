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

// The standard primitive 2^log_n-th root of unity of Fr (generator 5) and its inverse, plain.
void fr_roots(int log_n, U256* omega, U256* omega_inv) {
  HostField hf(U256::from(kFrModulus));
  U256 e = HostField::sub_raw(U256::from(kFrModulus), U256::of(1));  // (r - 1) >> log_n
  for (int i = 0; i < log_n; ++i) {
    for (int w = 0; w < 4; ++w) e.w[w] = (e.w[w] >> 1) | (w < 3 ? e.w[w + 1] << 63 : 0);
  }
  *omega = hf.powp(U256::of(5), e);
  *omega_inv = hf.invp(*omega);
}

int groth16_quotient(const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t* h, int log_n,
                     const uint64_t shift[4], hipStream_t stream) {
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
  const uint64_t *tw_f, *tw_i, *post_a, *post_b, *post_h, *zrow;
  if (int rc = twiddles(0, log_n, omega.w, stream, &tw_f)) return rc;
  if (int rc = twiddles(0, log_n, omega_inv.w, stream, &tw_i)) return rc;
  if (int rc = coset_table(0, log_n, g_inv.w, true, one, stream, &post_a)) return rc;         // n^-1 g^j
  if (int rc = coset_table(0, log_n, g_inv.w, true, thirty_two, stream, &post_b)) return rc;  // 2^5 n^-1 g^j
  if (int rc = coset_table(0, log_n, g.w, true, one, stream, &post_h)) return rc;             // n^-1 g^-j
  if (int rc = const_row(0, z_inv.w, stream, &zrow)) return rc;
  Scratch sc(stream);
  uint64_t* ev = sc.get<uint64_t>((size_t)n * 12);
  if (!ev) return 1;
  uint64_t* dst[3] = {ev, ev + 4 * n, ev + 8 * n};
  const uint64_t* src[3] = {a, b, c};
  const uint64_t* post[3] = {post_a, post_b, post_a};
  for (int k = 0; k < 3; ++k) {
    if (int rc = ntt_run(0, src[k], dst[k], log_n, tw_i, nullptr, post[k], true, QuotientArgs{}, stream)) return rc;
    if (int rc = ntt_run(0, dst[k], dst[k], log_n, tw_f, nullptr, nullptr, false, QuotientArgs{}, stream)) return rc;
  }
  QuotientArgs q{dst[0], dst[1], dst[2], zrow};
  return ntt_run(0, dst[0], h, log_n, tw_i, nullptr, post_h, true, q, stream);
}

}  // namespace zk

using namespace zk;

extern "C" {

int zk_groth16_quotient(const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t* h, int log_n,
                        const uint64_t shift[4], void* stream) {
  return groth16_quotient(a, b, c, h, log_n, shift, (hipStream_t)stream);
}

}  // extern "C"
