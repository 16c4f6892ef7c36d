// Wave-parallel Jacobian formulas for latency-bound chains (one point per wavefront).
//
// A long chain of dependent doublings and additions (an MSM's window
// combination, a bucket reduction's Horner over bit rows) runs on one point,
// so one lane would do every field product in turn. Here every lane holds the
// point and the independent products of each level of a formula run on
// different lanes at once: lane k of each 16-lane row computes product k of
// the level from operands it selects by its index, and a DPP row broadcast
// (row_newbcast) then hands product k to every lane of the row, one vector move
// per limb, so that values stay in vector registers (no scalar copies, no
// wait states). The four rows of the wavefront compute the same.
// dbl-2009-l needs 3 such levels instead of 7 products in a row, add-2007-bl 5
// instead of 16. The formulas and their special cases are those of zk_curve.hpp,
// so every coordinate is the same residue as theirs.
#pragma once

#include "zk_curve.hpp"

namespace zk {

// The lane's index in its 16-lane row.
__device__ __forceinline__ int wave_lane() { return (int)(threadIdx.x & 15); }

// The value lane `src` of the row holds, in every lane of the row.
template <int src>
__device__ __forceinline__ Fe lane_bcast(const Fe& x) {
  static_assert(src >= 0 && src < 16, "a row has 16 lanes");
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) r.l[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)x.l[i], 0x150 + src, 0xf, 0xf, true);
  return r;
}

// Operand k for lane k, limb by limb as values: a conditional over whole structs, or over limbs
// as lvalues, selects addresses and keeps the operands in scratch memory.
__device__ __forceinline__ Fe lane_pick(int lane, const Fe& a0, const Fe& a1) {
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const uint32_t x0 = a0.l[i], x1 = a1.l[i];
    r.l[i] = lane == 0 ? x0 : x1;
  }
  return r;
}

__device__ __forceinline__ Fe lane_pick(int lane, const Fe& a0, const Fe& a1, const Fe& a2) {
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const uint32_t x0 = a0.l[i], x1 = a1.l[i], x2 = a2.l[i];
    r.l[i] = lane == 0 ? x0 : (lane == 1 ? x1 : x2);
  }
  return r;
}

__device__ __forceinline__ Fe lane_pick(int lane, const Fe& a0, const Fe& a1, const Fe& a2, const Fe& a3) {
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const uint32_t x0 = a0.l[i], x1 = a1.l[i], x2 = a2.l[i], x3 = a3.l[i];
    r.l[i] = lane == 0 ? x0 : lane == 1 ? x1 : lane == 2 ? x2 : x3;
  }
  return r;
}

__device__ __forceinline__ Fe lane_pick(int lane, const Fe& a0, const Fe& a1, const Fe& a2, const Fe& a3, const Fe& a4) {
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const uint32_t x0 = a0.l[i], x1 = a1.l[i], x2 = a2.l[i], x3 = a3.l[i], x4 = a4.l[i];
    r.l[i] = lane == 0 ? x0 : lane == 1 ? x1 : lane == 2 ? x2 : lane == 3 ? x3 : x4;
  }
  return r;
}

// ---- lane-parallel linear combinations ---------------------------------------------------
// Between the product levels, lane k forms the combination its next operand (or output) needs,
// sum_j c_j(k) V_j + K(k) p over values V_j that every lane holds, with small per-lane integer
// coefficients (Acc64 and lc_* in zk_field.hpp). Values stay lazily reduced: product operands only
// need normalized limbs and a bound (a product of operands below 13p stays below 2p), and results
// are brought below 2p by subtracting a quotient estimated from the top limb. Every coordinate is
// the same residue as in zk_curve.hpp's formulas; only the representative differs, below 2p, until
// a caller canonicalizes it.

__device__ __forceinline__ int lane_coef(int lane, int c0, int c1) { return lane == 0 ? c0 : c1; }
__device__ __forceinline__ int lane_coef(int lane, int c0, int c1, int c2) {
  return lane == 0 ? c0 : (lane == 1 ? c1 : c2);
}

// x + y limb by limb, no carries: for x, y with limbs below 2^29 the sum's limbs stay below 2^30.
__device__ __forceinline__ Fe fe_add_nc(const Fe& x, const Fe& y) {
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) r.l[i] = x.l[i] + y.l[i];
  return r;
}

// 2x limb by limb (limbs below 2^30).
__device__ __forceinline__ Fe fe_dbl_nc(const Fe& x) { return fe_add_nc(x, x); }

// x == 0 mod p for x below 2p with normalized limbs.
template <class F>
__device__ __forceinline__ bool fe_is_zero_lazy(const F& f, const Fe& x) {
  return fe_is_zero(x) || fe_eq(x, fe_const<F>(f.p));
}

// 2P, dbl-2009-l (A = X^2, B = Y^2, C = B^2, D = 2((X + B)^2 - A - C), E = 3A, F = E^2,
// X3 = F - 2D, Y3 = E (D - X3) - 8C, Z3 = 2 Y Z) in three product levels:
//   {A, B, Y Z}, {C, T = (X + B)^2, A^2}, {P = A w} with w = 2T - 2A - 2C - 3A^2,
// since D - X3 = 3w and E (D - X3) = 9P; then X3 = 9A^2 - 4T + 4A + 4C, Y3 = 9P - 8C.
// Inputs below 2p; outputs below 2p. P of order 2 does not occur in a group of odd order.
template <class F>
__device__ __forceinline__ Jac jac_dbl_wave(const F& f, const Jac& p) {
  if (fe_is_zero_lazy(f, p.z) || fe_is_zero_lazy(f, p.y)) return jac_identity(f);
  const int lane = wave_lane();
  const Fe r1 = mont_raw(f, lane_pick(lane, p.x, p.y, p.y), lane_pick(lane, p.x, p.y, p.z));
  const Fe a = lane_bcast<0>(r1), b = lane_bcast<1>(r1), yz = lane_bcast<2>(r1);
  const Fe o = lane_pick(lane, b, fe_add_nc(p.x, b), a);  // X + B below 4p
  const Fe r2 = mont_raw(f, o, o);
  const Fe c = lane_bcast<0>(r2), t = lane_bcast<1>(r2), a2 = lane_bcast<2>(r2);
  Acc64 w = lc_start(f, 14);  // w + 14p in (0, 18p)
  lc_add(w, 2, t);
  lc_add(w, -2, a);
  lc_add(w, -2, c);
  lc_add(w, -3, a2);
  const Fe pa = mont_raw(f, a, lc_carry(w));  // the same in every lane
  Acc64 x = lc_start(f, lane_coef(lane, 8, 16, 0));  // lanes 0, 1, 2: X3, Y3, Z3
  lc_add(x, lane_coef(lane, 9, 0, 0), a2);
  lc_add(x, lane_coef(lane, -4, 0, 0), t);
  lc_add(x, lane_coef(lane, 4, 0, 0), a);
  lc_add(x, lane_coef(lane, 4, -8, 0), c);
  lc_add(x, lane_coef(lane, 0, 9, 0), pa);
  lc_add(x, lane_coef(lane, 0, 0, 2), yz);
  const Fe r3 = lc_finish(f, x);
  Jac r;
  r.x = lane_bcast<0>(r3);
  r.y = lane_bcast<1>(r3);
  r.z = lane_bcast<2>(r3);
  return r;
}

// P + Q, add-2007-bl (U1 = X1 Z2^2, U2 = X2 Z1^2, S1 = Y1 Z2^3, S2 = Y2 Z1^3, H = U2 - U1,
// I = (2H)^2, J = H I, r = 2(S2 - S1), V = U1 I, X3 = r^2 - J - 2V, Y3 = r (V - X3) - 2 S1 J,
// Z3 = ((Z1 + Z2)^2 - Z1^2 - Z2^2) H) in five product levels: {Z1Z1, Z2Z2, Y1 Z2, Y2 Z1,
// (Z1 + Z2)^2}, {U1, U2, S1, S2}, {I, r^2, Z3}, {J, V}, {r (3V + J - r^2), S1 J}.
// Inputs below 2p; outputs below 2p.
template <class F>
__device__ __forceinline__ Jac jac_add_wave(const F& f, const Jac& p, const Jac& q) {
  if (fe_is_zero_lazy(f, p.z)) return q;
  if (fe_is_zero_lazy(f, q.z)) return p;
  const int lane = wave_lane();
  const Fe zs = fe_add_nc(p.z, q.z);
  const Fe r1 = mont_raw(f, lane_pick(lane, p.z, q.z, p.y, q.y, zs), lane_pick(lane, p.z, q.z, q.z, p.z, zs));
  const Fe z1z1 = lane_bcast<0>(r1), z2z2 = lane_bcast<1>(r1), y1z2 = lane_bcast<2>(r1), y2z1 = lane_bcast<3>(r1),
           zz = lane_bcast<4>(r1);
  const Fe r2 = mont_raw(f, lane_pick(lane, p.x, q.x, y1z2, y2z1), lane_pick(lane, z2z2, z1z1, z2z2, z1z1));
  const Fe u1 = lane_bcast<0>(r2), u2 = lane_bcast<1>(r2), s1 = lane_bcast<2>(r2), s2 = lane_bcast<3>(r2);
  // lanes 0, 1, 2: H = U2 - U1, S2 - S1, (Z1 + Z2)^2 - Z1Z1 - Z2Z2, each below 2p
  Acc64 d = lc_start(f, lane_coef(lane, 2, 2, 4));
  lc_add(d, lane_coef(lane, 1, 0, 0), u2);
  lc_add(d, lane_coef(lane, -1, 0, 0), u1);
  lc_add(d, lane_coef(lane, 0, 1, 0), s2);
  lc_add(d, lane_coef(lane, 0, -1, 0), s1);
  lc_add(d, lane_coef(lane, 0, 0, 1), zz);
  lc_add(d, lane_coef(lane, 0, 0, -1), z1z1);
  lc_add(d, lane_coef(lane, 0, 0, -1), z2z2);
  const Fe rd = lc_finish(f, d);
  const Fe h = lane_bcast<0>(rd), sd = lane_bcast<1>(rd), zd = lane_bcast<2>(rd);
  if (fe_is_zero_lazy(f, h)) return fe_is_zero_lazy(f, sd) ? jac_dbl_wave(f, p) : jac_identity(f);
  const Fe h2 = fe_dbl_nc(h), rr = fe_dbl_nc(sd);  // 2H and r, below 4p
  const Fe r3 = mont_raw(f, lane_pick(lane, h2, rr, zd), lane_pick(lane, h2, rr, h));
  const Fe i = lane_bcast<0>(r3), rsq = lane_bcast<1>(r3);
  Jac r;
  r.z = lane_bcast<2>(r3);
  const Fe r4 = mont_raw(f, lane_pick(lane, h, u1), i);
  const Fe j = lane_bcast<0>(r4), v = lane_bcast<1>(r4);
  Acc64 e = lc_start(f, lane_coef(lane, 2, 0));  // lane 0: V - X3 = 3V + J - r^2, below 10p; lane 1: J
  lc_add(e, lane_coef(lane, 3, 0), v);
  lc_add(e, 1, j);
  lc_add(e, lane_coef(lane, -1, 0), rsq);
  const Fe r5 = mont_raw(f, lane_pick(lane, rr, s1), lc_carry(e));
  const Fe m1 = lane_bcast<0>(r5), m2 = lane_bcast<1>(r5);
  Acc64 o = lc_start(f, lane_coef(lane, 6, 4));  // lanes 0, 1: X3 = r^2 - J - 2V, Y3 = m1 - 2 m2
  lc_add(o, lane_coef(lane, 1, 0), rsq);
  lc_add(o, lane_coef(lane, -1, 0), j);
  lc_add(o, lane_coef(lane, -2, 0), v);
  lc_add(o, lane_coef(lane, 0, 1), m1);
  lc_add(o, lane_coef(lane, 0, -2), m2);
  const Fe ro = lc_finish(f, o);
  r.x = lane_bcast<0>(ro);
  r.y = lane_bcast<1>(ro);
  return r;
}

// The same point with canonical coordinates (for chains of the lazily reduced formulas above).
template <class F>
__device__ __forceinline__ Jac jac_canonical(const F& f, const Jac& p) {
  Jac r;
  r.x = canonical(f, p.x);
  r.y = canonical(f, p.y);
  r.z = canonical(f, p.z);
  return r;
}

}  // namespace zk
