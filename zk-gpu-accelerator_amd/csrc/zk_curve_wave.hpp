// Wave-parallel Jacobian formulas for latency-bound chains (one point per wavefront).
//
// A long chain of dependent doublings and additions (an MSM's window
// combination, a bucket reduction's Horner over bit rows) runs on one point,
// so one lane would do every field product in turn. Here the whole wavefront
// holds the point (the same value in every lane) and the independent products
// of each level of a formula run on different lanes at once: lane k computes
// product k of the level from operands it selects by its lane index, and every
// lane then reads the level's results back from lanes 0..k with readlane.
// dbl-2009-l needs 3 such levels instead of 7 products in a row, add-2007-bl 5
// instead of 16. The formulas, their special cases and the canonical products
// are those of zk_curve.hpp, so every coordinate is identical to theirs.
//
// The operand selections depend on the lane index, so each level's products
// stay vector multiply-adds (uniform operands would invite scalar code).
#pragma once

#include "zk_curve.hpp"

namespace zk {

__device__ __forceinline__ int wave_lane() { return (int)(threadIdx.x & 63); }

// The value lane `src` holds, in every lane.
__device__ __forceinline__ Fe lane_bcast(const Fe& x, int src) {
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) r.l[i] = (uint32_t)__builtin_amdgcn_readlane((int)x.l[i], src);
  return r;
}

__device__ __forceinline__ Fe lane_pick(int lane, const Fe& a0, const Fe& a1) { return lane == 0 ? a0 : a1; }

__device__ __forceinline__ Fe lane_pick(int lane, const Fe& a0, const Fe& a1, const Fe& a2) {
  return lane == 0 ? a0 : (lane == 1 ? a1 : a2);
}

__device__ __forceinline__ Fe lane_pick(int lane, const Fe& a0, const Fe& a1, const Fe& a2, const Fe& a3, const Fe& a4) {
  return lane == 0 ? a0 : lane == 1 ? a1 : lane == 2 ? a2 : lane == 3 ? a3 : a4;
}

// 2P, dbl-2009-l in three product levels: {A = X^2, B = Y^2}, {C = B^2, T = (X + B)^2,
// F = (3A)^2}, {E (D - X3), Y Z}.
template <class F>
__device__ __forceinline__ Jac jac_dbl_wave(const F& f, const Jac& p) {
  if (fe_is_zero(p.z) || fe_is_zero(p.y)) return jac_identity(f);
  const int lane = wave_lane();
  Fe o = lane_pick(lane, p.x, p.y);
  Fe r1 = mont(f, o, o);
  const Fe a = lane_bcast(r1, 0), b = lane_bcast(r1, 1);
  const Fe e = fadd(f, fdbl(f, a), a);
  o = lane_pick(lane, b, fadd(f, p.x, b), e);
  Fe r2 = mont(f, o, o);
  const Fe c = lane_bcast(r2, 0), t = lane_bcast(r2, 1), ff = lane_bcast(r2, 2);
  const Fe d = fdbl(f, fsub(f, fsub(f, t, a), c));
  Jac r;
  r.x = fsub(f, ff, fdbl(f, d));
  Fe u = lane_pick(lane, e, p.y);
  Fe v = lane_pick(lane, fsub(f, d, r.x), p.z);
  Fe r3 = mont(f, u, v);
  const Fe c8 = fdbl(f, fdbl(f, fdbl(f, c)));
  r.y = fsub(f, lane_bcast(r3, 0), c8);
  r.z = fdbl(f, lane_bcast(r3, 1));
  return r;
}

// P + Q, add-2007-bl in five product levels: {Z1Z1, Z2Z2, Y1 Z2, Y2 Z1, (Z1 + Z2)^2},
// {U1, U2, S1, S2}, {I = (2H)^2, r^2, Z3}, {J = H I, V = U1 I}, {r (V - X3), S1 J}.
template <class F>
__device__ __forceinline__ Jac jac_add_wave(const F& f, const Jac& p, const Jac& q) {
  if (fe_is_zero(p.z)) return q;
  if (fe_is_zero(q.z)) return p;
  const int lane = wave_lane();
  Fe u = lane_pick(lane, p.z, q.z, p.y, q.y, fadd(f, p.z, q.z));
  Fe v = lane_pick(lane, p.z, q.z, q.z, p.z, fadd(f, p.z, q.z));
  Fe r1 = mont(f, u, v);
  const Fe z1z1 = lane_bcast(r1, 0), z2z2 = lane_bcast(r1, 1), y1z2 = lane_bcast(r1, 2), y2z1 = lane_bcast(r1, 3),
           zz = lane_bcast(r1, 4);
  u = lane_pick(lane, p.x, q.x, y1z2, y2z1, y2z1);
  v = lane_pick(lane, z2z2, z1z1, z2z2, z1z1, z1z1);
  Fe r2 = mont(f, u, v);
  const Fe u1 = lane_bcast(r2, 0), u2 = lane_bcast(r2, 1), s1 = lane_bcast(r2, 2), s2 = lane_bcast(r2, 3);
  const Fe h = fsub(f, u2, u1);
  const Fe rr = fdbl(f, fsub(f, s2, s1));
  if (fe_is_zero(h)) return fe_is_zero(rr) ? jac_dbl_wave(f, p) : jac_identity(f);
  const Fe h2 = fdbl(f, h);
  u = lane_pick(lane, h2, rr, fsub(f, fsub(f, zz, z1z1), z2z2));
  v = lane_pick(lane, h2, rr, h);
  Fe r3 = mont(f, u, v);
  const Fe i = lane_bcast(r3, 0), r2sq = lane_bcast(r3, 1);
  Jac r;
  r.z = lane_bcast(r3, 2);
  u = lane_pick(lane, h, u1);
  Fe r4 = mont(f, u, i);
  const Fe j = lane_bcast(r4, 0), vv = lane_bcast(r4, 1);
  r.x = fsub(f, fsub(f, r2sq, j), fdbl(f, vv));
  u = lane_pick(lane, rr, s1);
  v = lane_pick(lane, fsub(f, vv, r.x), j);
  Fe r5 = mont(f, u, v);
  r.y = fsub(f, lane_bcast(r5, 0), fdbl(f, lane_bcast(r5, 1)));
  return r;
}

}  // namespace zk
