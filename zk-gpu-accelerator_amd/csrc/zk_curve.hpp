// BN254 G1 (y^2 = x^3 + 3 over F_p) in Jacobian coordinates on the device.
//
// The formulas are the Explicit-Formulas Database's a = 0 ones that the
// reference uses (src/zkgpu/curve.py:236-303; curve_gpu.py:90-232):
// dbl-2009-l, add-2007-bl and madd-2007-bl, with its special cases: a Z = 0
// operand returns the other one, H = 0 doubles or gives the identity, and a
// doubling with Y = 0 gives the identity (1, 1, 0). Every coordinate is kept
// canonical, so (X, Y, Z) equals the reference's exactly. Coordinates here
// are in the 261-form (zk_field.hpp); kernels convert at their boundaries.
#pragma once

#include "zk_field.hpp"

namespace zk {

struct Jac {
  Fe x, y, z;
};

struct Aff {
  Fe x, y;
};

template <class F>
__device__ __forceinline__ Jac jac_identity(const F& f) {
  Jac r;
  r.x = fe_const<F>(f.one);
  r.y = r.x;
  r.z = fe_zero();
  return r;
}

// 2P, dbl-2009-l: A = X^2, B = Y^2, C = B^2, D = 2((X + B)^2 - A - C), E = 3A,
// F = E^2, X3 = F - 2D, Y3 = E(D - X3) - 8C, Z3 = 2YZ.
template <class F>
__device__ __forceinline__ Jac jac_dbl(const F& f, const Jac& p) {
  if (fe_is_zero(p.z) || fe_is_zero(p.y)) return jac_identity(f);
  Fe a = mont_sqr(f, p.x);
  Fe b = mont_sqr(f, p.y);
  Fe c = mont_sqr(f, b);
  Fe t = mont_sqr(f, fadd(f, p.x, b));
  t = fsub(f, fsub(f, t, a), c);
  Fe d = fdbl(f, t);
  Fe e = fadd(f, fdbl(f, a), a);
  Fe ff = mont_sqr(f, e);
  Jac r;
  r.x = fsub(f, ff, fdbl(f, d));
  Fe c8 = fdbl(f, fdbl(f, fdbl(f, c)));
  r.y = fsub(f, mont(f, e, fsub(f, d, r.x)), c8);
  r.z = fdbl(f, mont(f, p.y, p.z));
  return r;
}

// P + Q, add-2007-bl: Z1Z1 = Z1^2, Z2Z2 = Z2^2, U1 = X1 Z2Z2, U2 = X2 Z1Z1,
// S1 = Y1 Z2 Z2Z2, S2 = Y2 Z1 Z1Z1, H = U2 - U1, I = (2H)^2, J = H I,
// r = 2(S2 - S1), V = U1 I, X3 = r^2 - J - 2V, Y3 = r(V - X3) - 2 S1 J,
// Z3 = ((Z1 + Z2)^2 - Z1Z1 - Z2Z2) H.
template <class F>
__device__ __forceinline__ Jac jac_add(const F& f, const Jac& p, const Jac& q) {
  if (fe_is_zero(p.z)) return q;
  if (fe_is_zero(q.z)) return p;
  Fe z1z1 = mont_sqr(f, p.z), z2z2 = mont_sqr(f, q.z);
  Fe u1 = mont(f, p.x, z2z2), u2 = mont(f, q.x, z1z1);
  Fe s1 = mont(f, mont(f, p.y, q.z), z2z2);
  Fe s2 = mont(f, mont(f, q.y, p.z), z1z1);
  Fe h = fsub(f, u2, u1);
  Fe rr = fdbl(f, fsub(f, s2, s1));
  if (fe_is_zero(h)) return fe_is_zero(rr) ? jac_dbl(f, p) : jac_identity(f);
  Fe i = mont_sqr(f, fdbl(f, h));
  Fe j = mont(f, h, i);
  Fe v = mont(f, u1, i);
  Jac r;
  r.x = fsub(f, fsub(f, mont_sqr(f, rr), j), fdbl(f, v));
  r.y = fsub(f, mont(f, rr, fsub(f, v, r.x)), fdbl(f, mont(f, s1, j)));
  Fe zz = mont_sqr(f, fadd(f, p.z, q.z));
  r.z = mont(f, fsub(f, fsub(f, zz, z1z1), z2z2), h);
  return r;
}

// P + (x2, y2) for a finite affine point, madd-2007-bl: Z1Z1 = Z1^2,
// U2 = x2 Z1Z1, S2 = y2 Z1 Z1Z1, H = U2 - X1, HH = H^2, I = 4HH, J = H I,
// r = 2(S2 - Y1), V = X1 I, X3 = r^2 - J - 2V, Y3 = r(V - X3) - 2 Y1 J,
// Z3 = (Z1 + H)^2 - Z1Z1 - HH.
template <class F>
__device__ __forceinline__ Jac jac_madd(const F& f, const Jac& p, const Aff& q) {
  if (fe_is_zero(p.z)) {
    Jac r;
    r.x = q.x;
    r.y = q.y;
    r.z = fe_const<F>(f.one);
    return r;
  }
  Fe z1z1 = mont_sqr(f, p.z);
  Fe u2 = mont(f, q.x, z1z1);
  Fe s2 = mont(f, mont(f, q.y, p.z), z1z1);
  Fe h = fsub(f, u2, p.x);
  Fe rr = fdbl(f, fsub(f, s2, p.y));
  if (fe_is_zero(h)) return fe_is_zero(rr) ? jac_dbl(f, p) : jac_identity(f);
  Fe hh = mont_sqr(f, h);
  Fe i = fdbl(f, fdbl(f, hh));
  Fe j = mont(f, h, i);
  Fe v = mont(f, p.x, i);
  Jac r;
  r.x = fsub(f, fsub(f, mont_sqr(f, rr), j), fdbl(f, v));
  r.y = fsub(f, mont(f, rr, fsub(f, v, r.x)), fdbl(f, mont(f, p.y, j)));
  r.z = fsub(f, fsub(f, mont_sqr(f, fadd(f, p.z, h)), z1z1), hh);
  return r;
}

template <class F>
__device__ __forceinline__ Jac jac_neg(const F& f, const Jac& p) {
  Jac r = p;
  r.y = fneg(f, p.y);
  return r;
}

// Row conversions: 256-form rows in HBM <-> 261-form coordinates.
template <class F>
__device__ __forceinline__ Fe load_coord(const F& f, const uint64_t* row) {
  return form256_to_261(f, load_row(row));
}

template <class F>
__device__ __forceinline__ void store_coord(const F& f, uint64_t* row, const Fe& x) {
  store_row(row, form261_to_256(f, x));
}

template <class F>
__device__ __forceinline__ Jac load_jac(const F& f, const uint64_t* p) {
  Jac r;
  r.x = load_coord(f, p);
  r.y = load_coord(f, p + 4);
  r.z = load_coord(f, p + 8);
  return r;
}

template <class F>
__device__ __forceinline__ void store_jac(const F& f, uint64_t* p, const Jac& a) {
  store_coord(f, p, a.x);
  store_coord(f, p + 4, a.y);
  store_coord(f, p + 8, a.z);
}

template <class F>
__device__ __forceinline__ Aff load_aff(const F& f, const uint64_t* p) {
  Aff r;
  r.x = load_coord(f, p);
  r.y = load_coord(f, p + 4);
  return r;
}

// Affine (x, y) of a finite Jacobian point: x = X/Z^2, y = Y/Z^3 (one inversion).
template <class F>
__device__ __forceinline__ Aff jac_to_aff(const F& f, const Jac& p) {
  Fe zi = finv(f, p.z);
  Fe zi2 = mont_sqr(f, zi);
  Aff r;
  r.x = mont(f, p.x, zi2);
  r.y = mont(f, mont(f, p.y, zi2), zi);
  return r;
}

}  // namespace zk
