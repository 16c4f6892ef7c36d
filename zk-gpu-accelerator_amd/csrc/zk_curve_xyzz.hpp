// Extended Jacobian ("XYZZ") coordinates for the MSM's bucket accumulation.
//
// A point (X, Y, ZZ, ZZZ) with ZZ^3 = ZZZ^2 stands for the affine (X / ZZ, Y / ZZZ); ZZ = 0
// is the identity. Adding an affine point costs 8 products and 2 squarings (madd-2008-s in
// the Explicit-Formulas Database) instead of the 7 + 4 of madd-2007-bl in Jacobian
// coordinates, with half the additions. Buckets leave the accumulation as Jacobian points:
// with ZZ = Z^2 and ZZZ = Z^3, (X ZZ^2, Y ZZZ^2, ZZZ) is (x Z^6, y Z^9, Z^3), the Jacobian
// point with Z' = Z^3. Coordinates are in the 261-form; the addition keeps them below 2p (lazily
// reduced), and the conversion to Jacobian coordinates returns canonical ones.
#pragma once

#include "zk_curve_wave.hpp"

namespace zk {

struct Xyzz {
  Fe x, y, zz, zzz;
};

template <class F>
__device__ __forceinline__ Xyzz xyzz_identity(const F& f) {
  Xyzz r;
  r.x = fe_const<F>(f.one);
  r.y = r.x;
  r.zz = fe_zero();
  r.zzz = fe_zero();
  return r;
}

// 2(x2, y2) for a finite affine point, mdbl-2008-s: U = 2 y2, V = U^2, W = U V, S = x2 V,
// M = 3 x2^2, X3 = M^2 - 2S, Y3 = M (S - X3) - W y2, ZZ3 = V, ZZZ3 = W. y2 = 0 gives the identity.
template <class F>
__device__ __forceinline__ Xyzz xyzz_dbl_affine(const F& f, const Aff& q) {
  if (fe_is_zero(q.y)) return xyzz_identity(f);
  const Fe u = fdbl(f, q.y);
  const Fe v = mont_sqr(f, u);
  const Fe w = mont(f, u, v);
  const Fe s = mont(f, q.x, v);
  const Fe x2 = mont_sqr(f, q.x);
  const Fe m = fadd(f, fdbl(f, x2), x2);
  Xyzz r;
  r.x = fsub(f, mont_sqr(f, m), fdbl(f, s));
  r.y = fsub(f, mont(f, m, fsub(f, s, r.x)), mont(f, w, q.y));
  r.zz = v;
  r.zzz = w;
  return r;
}

// P + (x2, y2) for a finite affine point, madd-2008-s: U2 = x2 ZZ1, S2 = y2 ZZZ1, P = U2 - X1,
// R = S2 - Y1, PP = P^2, PPP = P PP, Q = X1 PP, X3 = R^2 - PPP - 2Q, Y3 = R (Q - X3) - Y1 PPP,
// ZZ3 = ZZ1 PP, ZZZ3 = ZZZ1 PPP. P = 0 doubles (R = 0) or gives the identity.
// Lazily reduced: P's coordinates below 2p (the identity has ZZ exactly 0, and a finite point's ZZ
// is never 0 mod p), products without their final subtraction, R and Q - X3 as unreduced operands,
// X3 from one signed combination; the result's coordinates are below 2p. Independent products
// run in pairs (mont_raw2, sqr_raw2): U2 with S2, PP with R^2, PPP with Q, ZZ3 with ZZZ3.
// With neg the point added is -(x2, y2): the sign goes into R = -S2 - Y1 (+ 4p), and an identity
// P becomes (x2, y2, 1, -1), the same point as (x2, -y2, 1, 1) since ZZ^3 = ZZZ^2 still holds,
// so no coordinate is negated.
template <class F>
__device__ __forceinline__ Xyzz xyzz_madd(const F& f, const Xyzz& p, const Aff& q, bool neg = false) {
  if (fe_is_zero(p.zz)) {
    Xyzz r;
    r.x = q.x;
    r.y = q.y;
    r.zz = fe_const<F>(f.one);
    r.zzz = neg ? fe_const<F>(f.mone) : r.zz;
    return r;
  }
  Fe u2, s2;
  mont_raw2(f, q.x, p.zz, q.y, p.zzz, u2, s2);
  const Fe pp0 = fsub_lazy(f, u2, p.x);  // [0, 2p)
  if (fe_is_zero(pp0) || fe_eq(pp0, fe_const<F>(f.p))) {
    Aff qs = q;
    if (neg) qs.y = fneg(f, q.y);
    return fe_eq(canonical(f, neg ? fneg(f, canonical(f, s2)) : s2), canonical(f, p.y)) ? xyzz_dbl_affine(f, qs)
                                                                                        : xyzz_identity(f);
  }
  const Fe rr = fsub_norm_signed(f, s2, p.y, neg);  // +-S2 - Y1 + 2p or 4p, in (0, 4p]
  Fe pp, rr2, ppp, qq;
  sqr_raw2(f, pp0, rr, pp, rr2);
  mont_raw2(f, pp0, pp, p.x, pp, ppp, qq);
  Acc64 xa = lc_start(f, 6);  // R^2 - PPP - 2Q + 6p in (0, 8p)
  lc_add(xa, 1, rr2);
  lc_add(xa, -1, ppp);
  lc_add(xa, -2, qq);
  Xyzz r;
  r.x = lc_finish(f, xa);
  if constexpr (UsesIlp<F>::value) {
    r.y = fsub_lazy(f, mont_raw(f, rr, fsub_for_mul(f, qq, r.x)), mont_raw(f, p.y, ppp));
  } else {
    // R (Q - X3) + (2p - Y1) PPP with one reduction
    r.y = mont_raw_sum2(f, fsub_for_mul(f, qq, r.x), rr, fsub_norm(f, fe_zero(), p.y), ppp);
  }
  mont_raw2(f, p.zz, pp, p.zzz, ppp, r.zz, r.zzz);
  return r;
}

// P + Q, add-2007-bl (jac_add in zk_curve.hpp), lazily reduced for the MSM's internal sums:
// coordinates below 2p with normalized limbs in and out, products without their final
// subtraction, H and r = 2(S2 - S1) brought below 2p by signed combinations (for their zero tests),
// the outputs likewise. The same residues as jac_add; P = Q doubles, P = -Q gives the identity.
template <class F>
__device__ __forceinline__ Jac jac_add_lazy(const F& f, const Jac& p, const Jac& q) {
  if (fe_is_zero_lazy(f, p.z)) return q;
  if (fe_is_zero_lazy(f, q.z)) return p;
  const Fe z1z1 = sqr_raw(f, p.z), z2z2 = sqr_raw(f, q.z);
  const Fe u1 = mont_raw(f, p.x, z2z2), u2 = mont_raw(f, q.x, z1z1);
  const Fe s1 = mont_raw(f, mont_raw(f, p.y, q.z), z2z2);
  const Fe s2 = mont_raw(f, mont_raw(f, q.y, p.z), z1z1);
  Acc64 hd = lc_start(f, 2);  // H + 2p in (0, 4p)
  lc_add(hd, 1, u2);
  lc_add(hd, -1, u1);
  const Fe h = lc_finish(f, hd);
  Acc64 rd = lc_start(f, 4);  // r + 4p in (0, 8p)
  lc_add(rd, 2, s2);
  lc_add(rd, -2, s1);
  const Fe rr = lc_finish(f, rd);
  if (fe_is_zero_lazy(f, h)) return fe_is_zero_lazy(f, rr) ? jac_dbl(f, jac_canonical(f, p)) : jac_identity(f);
  const Fe i = sqr_raw(f, fe_dbl_nc(h));  // (2H)^2: 2H below 4p with limbs below 2^30
  const Fe j = mont_raw(f, h, i), v = mont_raw(f, u1, i);
  Acc64 xa = lc_start(f, 6);  // r^2 - J - 2V + 6p in (0, 8p)
  lc_add(xa, 1, sqr_raw(f, rr));
  lc_add(xa, -1, j);
  lc_add(xa, -2, v);
  Jac r;
  r.x = lc_finish(f, xa);
  const Fe m1 = mont_raw(f, rr, fsub_for_mul(f, v, r.x)), m2 = mont_raw(f, s1, j);
  Acc64 ya = lc_start(f, 4);  // r (V - X3) - 2 S1 J + 4p in (0, 6p)
  lc_add(ya, 1, m1);
  lc_add(ya, -2, m2);
  r.y = lc_finish(f, ya);
  Acc64 za = lc_start(f, 4);  // (Z1 + Z2)^2 - Z1Z1 - Z2Z2 + 4p in (0, 6p), a product operand
  lc_add(za, 1, sqr_raw(f, fe_add_nc(p.z, q.z)));
  lc_add(za, -1, z1z1);
  lc_add(za, -1, z2z2);
  r.z = mont_raw(f, lc_carry(za), h);
  return r;
}

// 2P, dbl-2008-s-1 (a = 0): U = 2Y, V = U^2, W = U V, S = X V, M = 3X^2, X3 = M^2 - 2S,
// Y3 = M (S - X3) - W Y, ZZ3 = V ZZ, ZZZ3 = W ZZZ. For coordinates below 2p; canonical results.
// (Only the rare equal-operand case of xyzz_add_lazy takes it.)
template <class F>
__device__ __forceinline__ Xyzz xyzz_dbl(const F& f, const Xyzz& p0) {
  Xyzz p;
  p.x = canonical(f, p0.x);
  p.y = canonical(f, p0.y);
  p.zz = canonical(f, p0.zz);
  p.zzz = canonical(f, p0.zzz);
  if (fe_is_zero(p.zz) || fe_is_zero(p.y)) return xyzz_identity(f);
  const Fe u = fdbl(f, p.y);
  const Fe v = mont_sqr(f, u);
  const Fe w = mont(f, u, v);
  const Fe s = mont(f, p.x, v);
  const Fe x2 = mont_sqr(f, p.x);
  const Fe m = fadd(f, fdbl(f, x2), x2);
  Xyzz r;
  r.x = fsub(f, mont_sqr(f, m), fdbl(f, s));
  r.y = fsub(f, mont(f, m, fsub(f, s, r.x)), mont(f, w, p.y));
  r.zz = mont(f, v, p.zz);
  r.zzz = mont(f, w, p.zzz);
  return r;
}

// P + Q, add-2008-s: U1 = X1 ZZ2, U2 = X2 ZZ1, S1 = Y1 ZZZ2, S2 = Y2 ZZZ1, P = U2 - U1,
// R = S2 - S1, PP = P^2, PPP = P PP, Q = U1 PP, X3 = R^2 - PPP - 2Q, Y3 = R (Q - X3) - S1 PPP,
// ZZ3 = ZZ1 ZZ2 PP, ZZZ3 = ZZZ1 ZZZ2 PPP (14 products against add-2007-bl's 16). Lazily reduced
// like xyzz_madd: coordinates below 2p with normalized limbs in and out. P = Q doubles, P = -Q
// gives the identity.
template <class F>
__device__ __forceinline__ Xyzz xyzz_add_lazy(const F& f, const Xyzz& p, const Xyzz& q) {
  if (fe_is_zero_lazy(f, p.zz)) return q;
  if (fe_is_zero_lazy(f, q.zz)) return p;
  const Fe u1 = mont_raw(f, p.x, q.zz), u2 = mont_raw(f, q.x, p.zz);
  const Fe s1 = mont_raw(f, p.y, q.zzz), s2 = mont_raw(f, q.y, p.zzz);
  Acc64 pd = lc_start(f, 2);  // P + 2p in (0, 4p)
  lc_add(pd, 1, u2);
  lc_add(pd, -1, u1);
  const Fe pp0 = lc_finish(f, pd);
  const Fe rr = fsub_norm(f, s2, s1);  // R + 2p, below 4p
  if (fe_is_zero_lazy(f, pp0)) {
    Acc64 rd = lc_start(f, 2);
    lc_add(rd, 1, s2);
    lc_add(rd, -1, s1);
    return fe_is_zero_lazy(f, lc_finish(f, rd)) ? xyzz_dbl(f, p) : xyzz_identity(f);
  }
  const Fe pp = sqr_raw(f, pp0);
  const Fe ppp = mont_raw(f, pp0, pp);
  const Fe qq = mont_raw(f, u1, pp);
  Acc64 xa = lc_start(f, 6);  // R^2 - PPP - 2Q + 6p in (0, 8p)
  lc_add(xa, 1, sqr_raw(f, rr));
  lc_add(xa, -1, ppp);
  lc_add(xa, -2, qq);
  Xyzz r;
  r.x = lc_finish(f, xa);
  Acc64 ya = lc_start(f, 2);  // R (Q - X3) - S1 PPP + 2p in (0, 4p)
  lc_add(ya, 1, mont_raw(f, rr, fsub_for_mul(f, qq, r.x)));
  lc_add(ya, -1, mont_raw(f, s1, ppp));
  r.y = lc_finish(f, ya);
  r.zz = mont_raw(f, mont_raw(f, p.zz, q.zz), pp);
  r.zzz = mont_raw(f, mont_raw(f, p.zzz, q.zzz), ppp);
  return r;
}

// A Jacobian point (coordinates below 2p) in XYZZ coordinates: (X, Y, Z^2, Z^3).
template <class F>
__device__ __forceinline__ Xyzz jac_to_xyzz(const F& f, const Jac& p) {
  Xyzz r;
  r.x = p.x;
  r.y = p.y;
  r.zz = sqr_raw(f, p.z);
  r.zzz = mont_raw(f, r.zz, p.z);
  return r;
}

// The same point in Jacobian coordinates: (X ZZ^2, Y ZZZ^2, ZZZ); the identity maps to (1, 1, 0).
template <class F>
__device__ __forceinline__ Jac xyzz_to_jac(const F& f, const Xyzz& p) {
  if (fe_is_zero_lazy(f, p.zz)) return jac_identity(f);
  Jac r;
  r.x = mont(f, p.x, mont_sqr(f, p.zz));
  r.y = mont(f, p.y, mont_sqr(f, p.zzz));
  r.z = canonical(f, p.zzz);
  return r;
}

}  // namespace zk
