// Limb-sliced point formulas for latency-bound chains: one limb of a coordinate per lane.
//
// A Horner pass over an MSM's windows is a chain of about a hundred doublings
// on one point, so its time is the latency of one doubling times the chain.
// zk_curve_wave.hpp gives each product of a formula level its own lane, but one
// lane still runs a whole 9x9-limb product. Here a field element is one vector
// register: lane k of a 16-lane row holds limb k (lanes 9..15 hold 0), and a
// Montgomery product runs as 9 steps of CIOS across the row (broadcast limb
// a_i, every lane adds a_i b_k and m p_k to its column, the row shifts down one
// lane), about 11 instructions a step. The four rows of the wavefront compute
// four independent products, one per product of a formula level, and hand the
// results to each other with the gfx950 row swaps (v_permlane16_swap,
// v_permlane32_swap).
//
// Values are signed and lazily reduced: limbs 0..7 lie in [-1, 2^29] after
// normalization, limb 8 carries the sign and the rest. A product of a and b
// is (a b + M p) / 2^261 for some 0 <= M < 2^261, so it lies in
// (ab/2^261, ab/2^261 + p); with p / 2^261 < 0.0060 and |a|, |b| < 7.1p it is
// below 1.31p in magnitude. Linear combinations are formed limb by limb in 64
// bits and folded back below 2.02p by subtracting a quotient estimated from
// the top limb. Every coordinate is the same residue as in zk_curve.hpp.
#pragma once

#include <type_traits>

#include "zk_curve_wave.hpp"

namespace zk {

template <int ctrl>
__device__ __forceinline__ int32_t sl_dpp(int32_t x) {
  return __builtin_amdgcn_mov_dpp(x, ctrl, 0xf, 0xf, true);
}

// Limb i of the row's value, in every lane of the row (row_newbcast).
template <int i>
__device__ __forceinline__ int32_t sl_limb(int32_t x) {
  return sl_dpp<0x150 + i>(x);
}

template <int I, int N, class Fn>
__device__ __forceinline__ void sl_for(Fn&& fn) {
  if constexpr (I < N) {
    fn(std::integral_constant<int, I>{});
    sl_for<I + 1, N>(fn);
  }
}

// Per-lane constants: the lane's limb of p, its limb of 3p (unnormalized), and the masks of a
// normalization step (limbs 0..7 keep 29 bits and pass their carry up; limb 8 keeps everything;
// lanes 9..15 keep nothing).
struct SlLane {
  int32_t p, p3;
  uint32_t lmask, hmask;
  int row;
};

template <class F>
__device__ __forceinline__ SlLane sl_lane() {
  const int k = (int)(threadIdx.x & 15);
  SlLane c;
  uint32_t p = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) p = k == i ? F::p[i] : p;
  c.p = (int32_t)p;
  c.p3 = (int32_t)(3u * p);
  c.lmask = k < NL - 1 ? LMASK : (k == NL - 1 ? 0xffffffffu : 0u);
  c.hmask = k < NL - 1 ? 0xffffffffu : 0u;
  c.row = (int)(threadIdx.x >> 4) & 3;
  return c;
}

// One carry step: limb k keeps its low 29 bits and receives limb k-1's carry (row_shr:1).
__device__ __forceinline__ int32_t sl_carry(const SlLane& c, int32_t v) {
  const int32_t h = (v >> LB) & (int32_t)c.hmask;
  return (int32_t)((uint32_t)v & c.lmask) + sl_dpp<0x111>(h);
}

// The same from 64-bit limbs below 2^60 in magnitude whose top limb fits 32 bits: limbs 0..7
// come out in [-2^31, 2^31) (in [-2^6, 2^29 + 2^6] from limbs below 2^35).
__device__ __forceinline__ int32_t sl_carry64(const SlLane& c, int64_t v) {
  const int32_t h = (int32_t)(v >> LB) & (int32_t)c.hmask;
  return (int32_t)((uint32_t)v & c.lmask) + sl_dpp<0x111>(h);
}

// a b 2^-261 (CIOS across the row). a, b: limbs 0..7 in [-2^6, 2^29 + 2^6], |a|, |b| below 2^258.
// A step's quotient m only has to be right mod 2^29, so only the last one is masked to 29 bits:
// with the others below 2^32 the sum M of m_i 2^(29i) stays below 2^261 (1 + 2^-26), and a column
// below 2^33 + 2^58.1 + 2^61 in magnitude. The result is normalized.
template <class F>
__device__ __forceinline__ int32_t sl_mul(const SlLane& c, int32_t a, int32_t b) {
  int64_t t = 0;
  sl_for<0, NL>([&](auto i) {
    t += (int64_t)sl_limb<decltype(i)::value>(a) * (int64_t)b;
    uint32_t m = (uint32_t)sl_limb<0>((int32_t)t) * F::ninv;
    if constexpr (decltype(i)::value == NL - 1) m &= LMASK;
    t = (int64_t)((uint64_t)t + (uint64_t)m * (uint32_t)c.p);  // column 0 is now 0 mod 2^29
    // shift down one limb: lane k takes lane k+1's low bits (row_shl:1) and keeps its carry
    t = (t >> LB) + (int64_t)(uint32_t)sl_dpp<0x101>((int32_t)((uint32_t)t & LMASK));
  });
  return sl_carry(c, sl_carry64(c, t));
}

// v (64-bit limbs below 2^36 in magnitude, |v| < 40p) less q p for q estimated from the top limb,
// normalized: |result| < 2.02p. The top limb of a normalized value is v / 2^232 to within 2, so
// q = floor(top 2^232 / p) is floor(v / p) to within 1.
template <class F>
__device__ __forceinline__ int32_t sl_fold(const SlLane& c, int64_t v) {
  constexpr int32_t kInv = (int32_t)((1ull << 61) / (((uint64_t)F::p[NL - 1] << LB) + F::p[NL - 2]));
  const int32_t n = sl_carry64(c, v);
  const int32_t q = __mulhi(sl_limb<NL - 1>(n), kInv);
  return sl_carry(c, sl_carry64(c, (int64_t)n - (int64_t)q * (int64_t)c.p));
}

// Every row receives row 0's, 1's, 2's and 3's value.
__device__ __forceinline__ void sl_rows(int32_t x, int32_t& r0, int32_t& r1, int32_t& r2, int32_t& r3) {
  const auto s = __builtin_amdgcn_permlane16_swap((uint32_t)x, (uint32_t)x, false, false);  // (x0 x0 x2 x2), (x1 x1 x3 x3)
  const auto e = __builtin_amdgcn_permlane32_swap(s[0], s[0], false, false);               // (x0 x0 x0 x0), (x2 x2 x2 x2)
  const auto o = __builtin_amdgcn_permlane32_swap(s[1], s[1], false, false);               // (x1 x1 x1 x1), (x3 x3 x3 x3)
  r0 = (int32_t)e[0];
  r1 = (int32_t)o[0];
  r2 = (int32_t)e[1];
  r3 = (int32_t)o[1];
}

__device__ __forceinline__ void sl_rows(int32_t x, int32_t& r0, int32_t& r1, int32_t& r2) {
  int32_t r3;
  sl_rows(x, r0, r1, r2, r3);
}

// The operand of this lane's row: x0, x1, x2 or x3 by row.
__device__ __forceinline__ int32_t sl_pick(const SlLane& c, int32_t x0, int32_t x1, int32_t x2, int32_t x3) {
  return c.row == 0 ? x0 : (c.row == 1 ? x1 : (c.row == 2 ? x2 : x3));
}

struct SlJac {
  int32_t x, y, z;
};

// 2P, dbl-2009-l in the three product levels of jac_dbl_wave ({A, B, YZ}, {C, T = (X + B)^2, A^2},
// {P = A w}, w = 2T - 2A - 2C - 3A^2; X3 = 9A^2 - 4T + 4A + 4C, Y3 = 9P - 8C, Z3 = 2YZ), one row per
// product. X, Y below 2.02p and Z below 2.1p in magnitude, in and out: then A, B, C, A^2, YZ are
// below 1.03p, T (of |X + B| < 3.1p) below 1.06p, |w| < 7.1p, P below 1.05p, the raw X3, Y3 below
// 17.2p (folded), and 2YZ below 2.06p (only normalized). Z = 0 mod p stays so (the identity
// doubles to itself).
template <class F>
__device__ __forceinline__ SlJac sl_dbl(const SlLane& c, const SlJac& q) {
  const int32_t r1 = sl_mul<F>(c, c.row == 0 ? q.x : q.y, c.row == 0 ? q.x : (c.row == 1 ? q.y : q.z));
  int32_t a, b, yz;
  sl_rows(r1, a, b, yz);
  const int32_t xb = sl_carry(c, q.x + b);
  const int32_t o = c.row == 0 ? b : (c.row == 1 ? xb : a);
  const int32_t r2 = sl_mul<F>(c, o, o);
  int32_t cc, t, a2;
  sl_rows(r2, cc, t, a2);
  const int64_t w64 = 2 * (int64_t)t - 2 * (int64_t)a - 2 * (int64_t)cc - 3 * (int64_t)a2;
  const int32_t w = sl_carry(c, sl_carry64(c, w64));
  const int32_t pa = sl_mul<F>(c, a, w);
  SlJac r;
  r.x = sl_fold<F>(c, 9 * (int64_t)a2 - 4 * (int64_t)t + 4 * (int64_t)a + 4 * (int64_t)cc);
  r.y = sl_fold<F>(c, 9 * (int64_t)pa - 8 * (int64_t)cc);
  r.z = sl_carry(c, 2 * yz);
  return r;
}

// A coordinate (limbs below 2^29) into the sliced layout: lane k takes limb k.
__device__ __forceinline__ int32_t sl_from_fe(const Fe& x) {
  const int k = (int)(threadIdx.x & 15);
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) v = k == i ? x.l[i] : v;
  return (int32_t)v;
}

// A sliced value (|v| < 2.1p) back to a coordinate below 2p with limbs below 2^29, in every lane:
// v + 3p is positive and below 2^256, its limbs gathered and reduced as lc_finish does.
template <class F>
__device__ __forceinline__ Fe sl_to_fe(const SlLane& c, int32_t v) {
  const int32_t u = v + c.p3;
  Acc64 a;
  sl_for<0, NL>([&](auto i) { a.l[decltype(i)::value] = (int64_t)sl_limb<decltype(i)::value>(u); });
  return lc_finish(F{}, a);
}

__device__ __forceinline__ SlJac sl_from_jac(const Jac& p) {
  SlJac r;
  r.x = sl_from_fe(p.x);
  r.y = sl_from_fe(p.y);
  r.z = sl_from_fe(p.z);
  return r;
}

// Coordinates below 2p, not canonical: a zero coordinate comes out as p (the identity as Z = p), so
// a result handed to code that tests Z == 0 exactly goes through jac_canonical first.
template <class F>
__device__ __forceinline__ Jac sl_to_jac(const SlLane& c, const SlJac& s) {
  Jac r;
  r.x = sl_to_fe<F>(c, s.x);
  r.y = sl_to_fe<F>(c, s.y);
  r.z = sl_to_fe<F>(c, s.z);
  return r;
}

// Whether v (|v| < 2.1p) is 0 mod p, in every lane (all rows hold the same value).
template <class F>
__device__ __forceinline__ bool sl_is_zero(const SlLane& c, int32_t v) {
  return fe_is_zero_lazy(F{}, sl_to_fe<F>(c, v));
}

// P + Q, add-2007-bl (U1 = X1 Z2^2, U2 = X2 Z1^2, S1 = Y1 Z2^3, S2 = Y2 Z1^3, H = U2 - U1,
// I = (2H)^2, J = H I, r = 2(S2 - S1), V = U1 I, X3 = r^2 - J - 2V, Y3 = r (V - X3) - 2 S1 J,
// Z3 = 2 Z1 Z2 H) in five levels of at most four products, one per row:
//   {Z1Z1, Z2Z2, Y1 Z2, Y2 Z1}, {U1, U2, S1, S2}, {I, r^2, Z1 Z2}, {J, V, Z1 Z2 H},
//   {r (3V + J - r^2), S1 J}.
// Inputs X, Y below 2.02p and Z below 2.1p in magnitude: the products of the first two levels lie
// in (-0.03p, 1.03p), |2H|, |r| < 2.06p, the products after them in (-0.05p, 1.05p), V - X3 =
// 3V + J - r^2 below 4.1p; outputs below 2.02p. P = Q doubles, P = -Q gives the identity (Z = 0),
// as zk_curve.hpp does.
template <class F>
__device__ __forceinline__ SlJac sl_add(const SlLane& c, const SlJac& p, const SlJac& q) {
  if (sl_is_zero<F>(c, p.z)) return q;
  if (sl_is_zero<F>(c, q.z)) return p;
  const int32_t r1 = sl_mul<F>(c, sl_pick(c, p.z, q.z, p.y, q.y), sl_pick(c, p.z, q.z, q.z, p.z));
  int32_t z1z1, z2z2, y1z2, y2z1;
  sl_rows(r1, z1z1, z2z2, y1z2, y2z1);
  const int32_t r2 = sl_mul<F>(c, sl_pick(c, p.x, q.x, y1z2, y2z1), sl_pick(c, z2z2, z1z1, z2z2, z1z1));
  int32_t u1, u2, s1, s2;
  sl_rows(r2, u1, u2, s1, s2);
  const int32_t h = sl_carry(c, u2 - u1);
  const int32_t rr = sl_carry(c, 2 * (s2 - s1));
  if (sl_is_zero<F>(c, h)) {
    if (sl_is_zero<F>(c, rr)) return sl_dbl<F>(c, p);
    SlJac id = p;
    id.z = 0;
    return id;
  }
  const int32_t h2 = sl_carry(c, 2 * h);
  const int32_t r3 = sl_mul<F>(c, sl_pick(c, h2, rr, p.z, p.z), sl_pick(c, h2, rr, q.z, q.z));
  int32_t i, rsq, z12;
  sl_rows(r3, i, rsq, z12);
  const int32_t r4 = sl_mul<F>(c, sl_pick(c, h, u1, z12, z12), sl_pick(c, i, i, h, h));
  int32_t j, v, zh;
  sl_rows(r4, j, v, zh);
  const int32_t e = sl_carry(c, sl_carry64(c, 3 * (int64_t)v + (int64_t)j - (int64_t)rsq));
  const int32_t r5 = sl_mul<F>(c, c.row == 0 ? rr : s1, c.row == 0 ? e : j);
  int32_t m1, m2, m3;  // rows 1..3 hold the same product
  sl_rows(r5, m1, m2, m3);
  SlJac r;
  r.x = sl_fold<F>(c, (int64_t)rsq - (int64_t)j - 2 * (int64_t)v);
  r.y = sl_fold<F>(c, (int64_t)m1 - 2 * (int64_t)m2);
  r.z = sl_fold<F>(c, 2 * (int64_t)zh);
  return r;
}

// 2^n P for P below 2p (the whole wavefront).
template <class F>
__device__ __forceinline__ Jac jac_dbl_n_sliced(const Jac& p, int n) {
  const SlLane c = sl_lane<F>();
  SlJac s = sl_from_jac(p);
  for (int e = 0; e < n; ++e) s = sl_dbl<F>(c, s);
  return sl_to_jac<F>(c, s);
}

}  // namespace zk
