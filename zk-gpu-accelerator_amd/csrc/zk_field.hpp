// Prime-field arithmetic for gfx950 on nine 29-bit limbs.
//
// Representation. An element is nine 29-bit limbs (261 bits) in VGPRs. With
// limbs below 2^29, each partial product is below 2^58, so a whole column of
// a 9x9 product-scanning Montgomery multiplication (at most 18 products plus
// the carry in) stays below 2^63 in one 64-bit accumulator: every partial
// product is a single v_mad_u64_u32 with no carry handling. Montgomery
// reduction is interleaved per column (finely integrated product scanning)
// with R' = 2^261.
//
// Memory layout. Rows in HBM keep the reference's layout: four little-endian
// 64-bit limbs of a * 2^256 mod p (the "256-form"). load_row/store_row repack
// between eight 32-bit words and nine 29-bit limbs with funnel shifts.
//
// Forms. mont(x, y) = x * y * 2^-261 mod p. Values are tagged by the power of
// two they carry: the 256-form a*2^256, the 261-form a*2^261 and plain a.
// mont keeps a form when one operand is in the 261-form, so twiddles,
// coset powers and curve coordinates inside kernels are kept in the 261-form,
// and data that is only ever multiplied by such constants (NTT values) stays
// in the 256-form with no conversion at all. The fixed multipliers c256,
// c266, c517 and c522 (zk_constants.hpp) move values between forms.
//
// Reductions. mont returns a value below 2p for inputs below 2p (below
// 2^257 for the BN254 moduli); canonical() subtracts p once. The arithmetic
// below returns canonical values (in [0, p)) from canonical inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zk_constants.hpp"

namespace zk {

constexpr int NL = 9;
constexpr int LB = 29;
constexpr uint32_t LMASK = (1u << LB) - 1;

struct Fe {
  uint32_t l[NL];
};

// A modulus chosen at run time (any odd p in [3, 2^255)), passed by value to
// kernels; it has the same member names as FrConst / FqConst, so templates
// read f.p[i] whether the field is built in or registered.
struct FieldRT {
  uint32_t p[NL];
  uint32_t p2[NL];  // 2p
  uint32_t c256[NL];
  uint32_t c266[NL];
  uint32_t c517[NL];
  uint32_t c522[NL];
  uint32_t one[NL];
  uint32_t p32[8];
  uint32_t ninv;
};

template <class F>
__device__ __forceinline__ Fe fe_const(const uint32_t (&c)[NL]) {
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) r.l[i] = c[i];
  return r;
}

__device__ __forceinline__ Fe fe_zero() {
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) r.l[i] = 0;
  return r;
}

__device__ __forceinline__ bool fe_is_zero(const Fe& a) {
  uint32_t x = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) x |= a.l[i];
  return x == 0;
}

__device__ __forceinline__ bool fe_eq(const Fe& a, const Fe& b) {
  uint32_t x = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) x |= a.l[i] ^ b.l[i];
  return x == 0;
}

// a - p if a >= p, for a < 2p.
template <class F>
__device__ __forceinline__ Fe canonical(const F& f, const Fe& a) {
  Fe d;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    uint32_t t = a.l[i] - f.p[i] - borrow;
    borrow = t >> 31;  // limbs are below 2^29, so a wrap sets the top bit
    d.l[i] = t & LMASK;
  }
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) r.l[i] = borrow ? a.l[i] : d.l[i];
  return r;
}

// Quotient digits. m_k = t_k (-p^-1) only has to be right mod 2^29 to clear the low 29 bits of
// column k. Field tags may set kFreeQuotient to keep m_0..m_7 at their full 32 bits (one
// v_and_b32 a column less). The m p part of a column then reaches (2^32 - 1) sum_j p_j
// instead of 2^29 sum_j p_j, so a tag may set it only for a modulus whose limbs sum to little
// (free_quotient_fits: the m p part, a carry below 2^35 and an a b part of 9 * 2^60 stay
// below 2^64) and only where every product's operands fit that a b part: limbs below 2^30 in
// both operands, or below 3 * 2^29 against 2^29, or the two-product sum of mont_raw_sum2. The
// last quotient stays masked, so sum_k m_k 2^(29k) < 2^261 (1 + 2^-26) and a product still
// returns at most a b / 2^261 + p (1 + 2^-26).
template <class F, class = void>
struct FreeQuotient {
  static constexpr bool value = false;
};
template <class F>
struct FreeQuotient<F, decltype((void)F::kFreeQuotient)> {
  static constexpr bool value = F::kFreeQuotient;
};

template <class F>
constexpr bool free_quotient_fits() {
  unsigned __int128 sum = 0;
  for (int i = 0; i < NL; ++i) sum += F::p[i];
  const unsigned __int128 col = (sum << 32) + ((unsigned __int128)9 << 60) + ((unsigned __int128)1 << 35);
  return col < ((unsigned __int128)1 << 64);
}

template <class F>
__device__ __forceinline__ uint32_t mont_quotient(const F& f, uint64_t t, int k) {
  const uint32_t m = (uint32_t)t * f.ninv;
  if constexpr (FreeQuotient<F>::value) {
    static_assert(free_quotient_fits<F>(), "the modulus' limbs sum too high for unmasked quotients");
    return k < NL - 1 ? m : m & LMASK;
  }
  return m & LMASK;
}

// x * y * 2^-261 mod p, below 2p: one accumulator per column (fewest instructions).
template <class F>
__device__ __forceinline__ Fe mont_raw_serial(const F& f, const Fe& a, const Fe& b) {
  uint32_t m[NL];
  Fe r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < NL; ++k) {
#pragma unroll
    for (int i = 0; i <= k; ++i) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
    for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * f.p[k - i];
    m[k] = mont_quotient(f, acc, k);
    acc += (uint64_t)m[k] * f.p[0];  // clears the low 29 bits
    acc >>= LB;
  }
#pragma unroll
  for (int k = NL; k < 2 * NL - 1; ++k) {
#pragma unroll
    for (int i = k - (NL - 1); i < NL; ++i) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
    for (int i = k - (NL - 1); i < NL; ++i) acc += (uint64_t)m[i] * f.p[k - i];
    r.l[k - NL] = (uint32_t)acc & LMASK;
    acc >>= LB;
  }
  r.l[NL - 1] = (uint32_t)acc;
  return r;
}

// Keeps v where it is in a chain of multiply-adds: the compiler may not reassociate across it.
__device__ __forceinline__ void in_order(uint64_t& v) { asm("" : "+v"(v)); }

// Two independent products a b and c d (each as mont_raw_serial) in lockstep. Left alone, the
// compiler splits each column into two partial sums joined by a 64-bit addition, since a
// multiply-add that reads the one just before it needs a wait state; here each column is one
// in-order chain per product and the other product's multiply-add sits between two dependent
// ones: one instruction a column less per product, at the cost of a few s_nop a product (which
// the other wavefronts on the SIMD fill). Same limb bounds and results as two mont_raw_serial.
template <class F>
__device__ __forceinline__ void mont_raw_pair(const F& f, const Fe& a, const Fe& b, const Fe& c, const Fe& d, Fe& r,
                                              Fe& s) {
  uint32_t m[NL], n[NL];
  uint64_t x = 0, y = 0;
#pragma unroll
  for (int k = 0; k < 2 * NL - 1; ++k) {
    const int lo = k < NL ? 0 : k - (NL - 1), hi = k < NL ? k : NL - 1;
#pragma unroll
    for (int i = lo; i <= hi; ++i) {
      x += (uint64_t)a.l[i] * b.l[k - i];
      in_order(x);
      y += (uint64_t)c.l[i] * d.l[k - i];
      in_order(y);
    }
#pragma unroll
    for (int i = lo; i < (k < NL ? k : NL); ++i) {
      x += (uint64_t)m[i] * f.p[k - i];
      in_order(x);
      y += (uint64_t)n[i] * f.p[k - i];
      in_order(y);
    }
    if (k < NL) {
      m[k] = mont_quotient(f, x, k);
      n[k] = mont_quotient(f, y, k);
      x += (uint64_t)m[k] * f.p[0];  // clears the low 29 bits
      y += (uint64_t)n[k] * f.p[0];
    } else {
      r.l[k - NL] = (uint32_t)x & LMASK;
      s.l[k - NL] = (uint32_t)y & LMASK;
    }
    x >>= LB;
    y >>= LB;
  }
  r.l[NL - 1] = (uint32_t)x;
  s.l[NL - 1] = (uint32_t)y;
}

// a^2 and c^2 (each as mont_sqr_raw) in lockstep, as mont_raw_pair.
template <class F>
__device__ __forceinline__ void mont_sqr_raw_pair(const F& f, const Fe& a, const Fe& c, Fe& r, Fe& s) {
  uint32_t m[NL], n[NL], da[NL], dc[NL];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    da[i] = a.l[i] << 1;
    dc[i] = c.l[i] << 1;
  }
  uint64_t x = 0, y = 0;
#pragma unroll
  for (int k = 0; k < 2 * NL - 1; ++k) {
    const int lo = k < NL ? 0 : k - (NL - 1);
#pragma unroll
    for (int i = lo; 2 * i < k; ++i) {
      x += (uint64_t)a.l[i] * da[k - i];
      in_order(x);
      y += (uint64_t)c.l[i] * dc[k - i];
      in_order(y);
    }
    if ((k & 1) == 0) {
      x += (uint64_t)a.l[k / 2] * a.l[k / 2];
      in_order(x);
      y += (uint64_t)c.l[k / 2] * c.l[k / 2];
      in_order(y);
    }
#pragma unroll
    for (int i = lo; i < (k < NL ? k : NL); ++i) {
      x += (uint64_t)m[i] * f.p[k - i];
      in_order(x);
      y += (uint64_t)n[i] * f.p[k - i];
      in_order(y);
    }
    if (k < NL) {
      m[k] = mont_quotient(f, x, k);
      n[k] = mont_quotient(f, y, k);
      x += (uint64_t)m[k] * f.p[0];
      y += (uint64_t)n[k] * f.p[0];
    } else {
      r.l[k - NL] = (uint32_t)x & LMASK;
      s.l[k - NL] = (uint32_t)y & LMASK;
    }
    x >>= LB;
    y >>= LB;
  }
  r.l[NL - 1] = (uint32_t)x;
  s.l[NL - 1] = (uint32_t)y;
}

// (a b + c d) 2^-261 mod p below 2p with one reduction, serial columns: for a < 6p with limbs below
// 3 * 2^29 (a product operand from fsub_for_mul), b < 4p and c, d < 2p with normalized limbs, a
// column sums at most 9 * 3 * 2^58 + 9 * 2^58 + 9 * 2^58 + 2^35 < 2^64, and (ab + cd) / 2^261 <
// 28p^2 / 2^261 < 0.2p, so the result stays below 2p. Saves one Montgomery reduction against two
// products.
template <class F>
__device__ __forceinline__ Fe mont_raw_sum2(const F& f, const Fe& a, const Fe& b, const Fe& c, const Fe& d) {
  uint32_t m[NL];
  Fe r;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < NL; ++k) {
#pragma unroll
    for (int i = 0; i <= k; ++i) acc += (uint64_t)a.l[i] * b.l[k - i] + (uint64_t)c.l[i] * d.l[k - i];
#pragma unroll
    for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * f.p[k - i];
    m[k] = mont_quotient(f, acc, k);
    acc += (uint64_t)m[k] * f.p[0];  // clears the low 29 bits
    acc >>= LB;
  }
#pragma unroll
  for (int k = NL; k < 2 * NL - 1; ++k) {
#pragma unroll
    for (int i = k - (NL - 1); i < NL; ++i) acc += (uint64_t)a.l[i] * b.l[k - i] + (uint64_t)c.l[i] * d.l[k - i];
#pragma unroll
    for (int i = k - (NL - 1); i < NL; ++i) acc += (uint64_t)m[i] * f.p[k - i];
    r.l[k - NL] = (uint32_t)acc & LMASK;
    acc >>= LB;
  }
  r.l[NL - 1] = (uint32_t)acc;
  return r;
}

// The same product with independent partial sums per column: the a*b column sums depend on
// nothing but the inputs, and the m*p sums take the newest m last, so the chain from one
// column's m to the next is a few instructions instead of a whole column. About 15% more
// instructions, several times less latency on one lane (for the serial parts of the MSM).
template <class F>
__device__ __forceinline__ Fe mont_raw_ilp(const F& f, const Fe& a, const Fe& b) {
  uint32_t m[NL];
  Fe r;
  uint64_t pa[2 * NL - 1];
#pragma unroll
  for (int k = 0; k < 2 * NL - 1; ++k) {
    uint64_t s = 0;
#pragma unroll
    for (int i = (k < NL ? 0 : k - (NL - 1)); i <= (k < NL ? k : NL - 1); ++i) s += (uint64_t)a.l[i] * b.l[k - i];
    pa[k] = s;
  }
  uint64_t carry = 0;
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    uint64_t pm = 0;
#pragma unroll
    for (int i = 0; i < k; ++i) pm += (uint64_t)m[i] * f.p[k - i];
    uint64_t t = pa[k] + carry + pm;
    m[k] = mont_quotient(f, t, k);
    t += (uint64_t)m[k] * f.p[0];
    carry = t >> LB;
  }
#pragma unroll
  for (int k = NL; k < 2 * NL - 1; ++k) {
    uint64_t pm = 0;
#pragma unroll
    for (int i = k - (NL - 1); i < NL; ++i) pm += (uint64_t)m[i] * f.p[k - i];
    uint64_t t = pa[k] + pm + carry;
    r.l[k - NL] = (uint32_t)t & LMASK;
    carry = t >> LB;
  }
  r.l[NL - 1] = (uint32_t)carry;
  return r;
}

// Field tags may set kIlp to pick the low-latency product.
template <class F, class = void>
struct UsesIlp {
  static constexpr bool value = false;
};
template <class F>
struct UsesIlp<F, decltype((void)F::kIlp)> {
  static constexpr bool value = F::kIlp;
};

template <class F>
__device__ __forceinline__ Fe mont_raw(const F& f, const Fe& a, const Fe& b) {
  if constexpr (UsesIlp<F>::value) return mont_raw_ilp(f, a, b);
  else return mont_raw_serial(f, a, b);
}

// x^2 * 2^-261 mod p, below 2p: the 36 cross products are formed once against
// a doubled operand (30-bit limbs, still far from overflowing a column).
template <class F>
__device__ __forceinline__ Fe mont_sqr_raw(const F& f, const Fe& a) {
  uint32_t m[NL], d[NL];
  Fe r;
#pragma unroll
  for (int i = 0; i < NL; ++i) d[i] = a.l[i] << 1;
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < NL; ++k) {
#pragma unroll
    for (int i = 0; 2 * i < k; ++i) acc += (uint64_t)a.l[i] * d[k - i];
    if ((k & 1) == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
    for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * f.p[k - i];
    m[k] = mont_quotient(f, acc, k);
    acc += (uint64_t)m[k] * f.p[0];
    acc >>= LB;
  }
#pragma unroll
  for (int k = NL; k < 2 * NL - 1; ++k) {
#pragma unroll
    for (int i = k - (NL - 1); 2 * i < k; ++i) acc += (uint64_t)a.l[i] * d[k - i];
    if ((k & 1) == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
    for (int i = k - (NL - 1); i < NL; ++i) acc += (uint64_t)m[i] * f.p[k - i];
    r.l[k - NL] = (uint32_t)acc & LMASK;
    acc >>= LB;
  }
  r.l[NL - 1] = (uint32_t)acc;
  return r;
}

template <class F>
__device__ __forceinline__ Fe mont(const F& f, const Fe& a, const Fe& b) {
  return canonical(f, mont_raw(f, a, b));
}

template <class F>
__device__ __forceinline__ Fe mont_sqr(const F& f, const Fe& a) {
  if constexpr (UsesIlp<F>::value) return canonical(f, mont_raw_ilp(f, a, a));
  return canonical(f, mont_sqr_raw(f, a));
}

// (a + b) mod p for canonical a, b.
template <class F>
__device__ __forceinline__ Fe fadd(const F& f, const Fe& a, const Fe& b) {
  Fe s;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    uint32_t t = a.l[i] + b.l[i] + c;
    s.l[i] = t & LMASK;
    c = t >> LB;
  }
  return canonical(f, s);  // s < 2p < 2^261, so the carry out of the top limb is 0
}

// (a - b) mod p for canonical a, b.
template <class F>
__device__ __forceinline__ Fe fsub(const F& f, const Fe& a, const Fe& b) {
  Fe d;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    uint32_t t = a.l[i] - b.l[i] - borrow;
    borrow = t >> 31;
    d.l[i] = t & LMASK;
  }
  // a < b: the difference wrapped modulo 2^261, so add p back (the carry out cancels the wrap)
  uint32_t mask = 0u - borrow, c = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    uint32_t t = d.l[i] + (f.p[i] & mask) + c;
    d.l[i] = t & LMASK;
    c = t >> LB;
  }
  return d;
}

// ---- lazily reduced arithmetic (for long chains of additions and products, e.g. NTT passes) ----
// Values are kept in [0, 2p). mont_raw of operands below 6p and p stays below 2p (6p * p
// < p * 2^261), so a product needs no final subtraction.

// (a + b) reduced into [0, 2p), for a, b < 2p.
template <class F>
__device__ __forceinline__ Fe fadd_lazy(const F& f, const Fe& a, const Fe& b) {
  Fe s, d;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    uint32_t t = a.l[i] + b.l[i] + c;
    s.l[i] = t & LMASK;
    c = t >> LB;
  }
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    uint32_t t = s.l[i] - f.p2[i] - borrow;
    borrow = t >> 31;
    d.l[i] = t & LMASK;
  }
#pragma unroll
  for (int i = 0; i < NL; ++i) s.l[i] = borrow ? s.l[i] : d.l[i];
  return s;
}

// Built-in fields carry p4r: 4p with limbs 1..7 raised by 2^29 - 1 and limb 0 by 2^29, the
// borrows taken from the limb above (the top limb is 1 lower).
template <class F, class = void>
struct HasP4r {
  static constexpr bool value = false;
};
template <class F>
struct HasP4r<F, decltype((void)F::p4r)> {
  static constexpr bool value = true;
};

// The operand of a product: a - b + 4p in (2p, 6p) with limbs below 2^31 and no carries
// (fields with p4r), else a - b + 2p in (0, 4p) with normalized limbs; for a, b < 2p.
// Against p4r each limb a_i + p4r_i - b_i lies in [0, 3 * 2^29). mont_raw takes such an
// operand: a column then sums at most 9 * 2^60 + 9 * 2^58 + 2^35 < 2^64, and the result
// (a b + m p) / 2^261 < 6p^2 / 2^261 + p stays below 2p.
template <class F>
__device__ __forceinline__ Fe fsub_for_mul(const F& f, const Fe& a, const Fe& b) {
  Fe d;
  if constexpr (HasP4r<F>::value) {
#pragma unroll
    for (int i = 0; i < NL; ++i) d.l[i] = a.l[i] + F::p4r[i] - b.l[i];
    return d;
  }
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const int32_t t = (int32_t)(a.l[i] + f.p2[i]) - (int32_t)b.l[i] + c;  // |t| < 2^31
    d.l[i] = (uint32_t)t & LMASK;
    c = t >> LB;  // arithmetic: the signed carry
  }
  return d;
}

// a - b + 2p in (0, 4p) with normalized limbs (signed carries), for a, b < 2p: an operand that
// may be squared (both operands of a product then have limbs below 2^29).
template <class F>
__device__ __forceinline__ Fe fsub_norm(const F& f, const Fe& a, const Fe& b) {
  Fe d;
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const int32_t t = (int32_t)(a.l[i] + f.p2[i]) - (int32_t)b.l[i] + c;  // |t| < 2^31
    d.l[i] = (uint32_t)t & LMASK;
    c = t >> LB;
  }
  return d;
}

// (neg ? -a : a) - b + (neg ? 4p : 2p), in (0, 4p] with normalized limbs (signed carries), for
// a, b < 2p with normalized limbs: each limb sum lies in (-2^30, 2^30 + 2^29 + 4), inside 32 bits.
template <class F>
__device__ __forceinline__ Fe fsub_norm_signed(const F& f, const Fe& a, const Fe& b, bool neg) {
  const int32_t sgn = neg ? -1 : 0;  // (x ^ sgn) - sgn is -x when neg, else x
  const int32_t k = neg ? 2 : 1;
  Fe d;
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const int32_t t = (((int32_t)a.l[i] ^ sgn) - sgn) - (int32_t)b.l[i] + k * (int32_t)f.p2[i] + c;
    d.l[i] = (uint32_t)t & LMASK;
    c = t >> LB;
  }
  return d;
}

// x^2 * 2^-261 below 2p with the field's product schedule (the squaring shortcut for serial columns).
template <class F>
__device__ __forceinline__ Fe sqr_raw(const F& f, const Fe& a) {
  if constexpr (UsesIlp<F>::value) return mont_raw_ilp(f, a, a);
  else return mont_sqr_raw(f, a);
}

// r = a b, s = c d (as mont_raw), in lockstep with the serial schedule.
template <class F>
__device__ __forceinline__ void mont_raw2(const F& f, const Fe& a, const Fe& b, const Fe& c, const Fe& d, Fe& r, Fe& s) {
  if constexpr (UsesIlp<F>::value) {
    r = mont_raw_ilp(f, a, b);
    s = mont_raw_ilp(f, c, d);
  } else {
    mont_raw_pair(f, a, b, c, d, r, s);
  }
}

// r = a^2, s = c^2 (as sqr_raw), in lockstep with the serial schedule.
template <class F>
__device__ __forceinline__ void sqr_raw2(const F& f, const Fe& a, const Fe& c, Fe& r, Fe& s) {
  if constexpr (UsesIlp<F>::value) {
    r = mont_raw_ilp(f, a, a);
    s = mont_raw_ilp(f, c, c);
  } else {
    mont_sqr_raw_pair(f, a, c, r, s);
  }
}

// a - b in [0, 2p) for a, b < 2p.
template <class F>
__device__ __forceinline__ Fe fsub_lazy(const F& f, const Fe& a, const Fe& b) {
  Fe d;
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    uint32_t t = a.l[i] - b.l[i] - borrow;
    borrow = t >> 31;
    d.l[i] = t & LMASK;
  }
  uint32_t mask = 0u - borrow, c = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    uint32_t t = d.l[i] + (f.p2[i] & mask) + c;
    d.l[i] = t & LMASK;
    c = t >> LB;
  }
  return d;
}

// ---- small signed combinations (chains of additions and products with lazy reduction) ----
// sum_j c_j V_j + K p with small integer coefficients, in nine signed 64-bit limb accumulators (one
// v_mad_i64_i32 per term and limb), then either normalized (lc_carry: a product operand, only a
// bound needed) or brought below 2p (lc_finish: a quotient estimated from the top limb, for the
// BN254 moduli).

struct Acc64 {
  int64_t l[NL];
};

// K p, the offset that keeps a combination non-negative.
template <class F>
__device__ __forceinline__ Acc64 lc_start(const F&, int k) {
  Acc64 a;
#pragma unroll
  for (int i = 0; i < NL; ++i) a.l[i] = (int64_t)k * (int64_t)F::p[i];
  return a;
}

// a += c v for v with limbs below 2^31.
__device__ __forceinline__ void lc_add(Acc64& a, int c, const Fe& v) {
#pragma unroll
  for (int i = 0; i < NL; ++i) a.l[i] += (int64_t)c * (int64_t)(int32_t)v.l[i];
}

// Limbs 0..7 into [0, 2^29) with arithmetic carries; the top limb takes the rest (non-negative
// for a non-negative value).
__device__ __forceinline__ Fe lc_carry(Acc64 a) {
  Fe r;
#pragma unroll
  for (int i = 0; i < NL - 1; ++i) {
    a.l[i + 1] += a.l[i] >> LB;
    r.l[i] = (uint32_t)a.l[i] & LMASK;
  }
  r.l[NL - 1] = (uint32_t)a.l[NL - 1];
  return r;
}

// The value V (0 <= V < 2^261, limbs below 2^36 in magnitude) less q p, q = floor(t / (p_8 + 1))
// for t = the top limb less a margin for the uncarried lower limbs (they move V / 2^232 by less
// than 2^8): q p <= V, and V - q p < 2^232 (p_8 + 4096 + 2^8 + V / (2^232 (p_8 + 1)) + 1)
// < p (1 + 2^13 / p_8) < 2p for the BN254 moduli (p_8 > 2^21).
template <class F>
__device__ __forceinline__ Fe lc_finish(const F&, Acc64 a) {
  static_assert(F::p[NL - 1] > (1u << 21), "the quotient estimate needs a full-size top limb");
  const int64_t t = a.l[NL - 1] - 4096;
  const int32_t q = t > 0 ? (int32_t)((uint32_t)t / (F::p[NL - 1] + 1u)) : 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) a.l[i] -= (int64_t)q * (int64_t)F::p[i];
  return lc_carry(a);
}

// Sums kept wide between two butterfly stages (built-in fields, which carry p8r: 8p with limbs 1..7
// raised by 2^30 - 2 and limb 0 by 2^30, the borrows taken from the limb above).
// a + b without carries: for a, b < 2p with normalized limbs, a value below 4p with limbs
// 0..7 at most 2^30 - 2.
__device__ __forceinline__ Fe fadd_wide(const Fe& a, const Fe& b) {
  Fe s;
#pragma unroll
  for (int i = 0; i < NL; ++i) s.l[i] = a.l[i] + b.l[i];
  return s;
}

// The operand of a product: a - b + 8p in (4p, 12p) with no carries, for a, b from fadd_wide.
// Against p8r each limb a_i + p8r_i - b_i lies in [0, 5 * 2^29), so against a normalized operand
// below 2p a column of mont_raw sums at most 9 * 5 * 2^58 + 9 * 2^58 + 2^35 < 2^64, and the result
// (a b + m p) / 2^261 < 24 p^2 / 2^261 + p stays below 2p.
template <class F>
__device__ __forceinline__ Fe fsub_wide(const F&, const Fe& a, const Fe& b) {
  Fe d;
#pragma unroll
  for (int i = 0; i < NL; ++i) d.l[i] = a.l[i] + F::p8r[i] - b.l[i];
  return d;
}

// a + b below 2p with normalized limbs, for a, b from fadd_wide (a sum below 8p, limbs below 2^31).
template <class F>
__device__ __forceinline__ Fe fadd_wide_finish(const F& f, const Fe& a, const Fe& b) {
  Acc64 t;
#pragma unroll
  for (int i = 0; i < NL; ++i) t.l[i] = (int64_t)(a.l[i] + b.l[i]);
  return lc_finish(f, t);
}

// a - b in [0, 2p) with normalized limbs, for a, b from fadd_wide (a - b + 4p lies in (0, 8p)).
template <class F>
__device__ __forceinline__ Fe fsub_wide_finish(const F& f, const Fe& a, const Fe& b) {
  Acc64 t = lc_start(f, 4);
#pragma unroll
  for (int i = 0; i < NL; ++i) t.l[i] += (int64_t)a.l[i] - (int64_t)b.l[i];
  return lc_finish(f, t);
}

template <class F>
__device__ __forceinline__ Fe fdbl(const F& f, const Fe& a) {
  return fadd(f, a, a);
}

template <class F>
__device__ __forceinline__ Fe fneg(const F& f, const Fe& a) {
  return fsub(f, fe_zero(), a);
}

// ---- conversion between 8x32-bit words (memory) and 9x29-bit limbs ----

__device__ __forceinline__ Fe from_words(const uint32_t w[8]) {
  Fe r;
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int bit = LB * k, word = bit >> 5, sh = bit & 31;
    uint32_t lo = w[word];
    uint32_t hi = (word + 1 < 8) ? w[word + 1] : 0u;
    uint32_t v = sh ? ((lo >> sh) | (hi << (32 - sh))) : lo;
    r.l[k] = v & LMASK;
  }
  return r;
}

__device__ __forceinline__ void to_words(const Fe& a, uint32_t w[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int bit = 32 * j;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int lo = LB * k;  // limb k covers bits [lo, lo + 29)
      if (lo + LB <= bit || lo >= bit + 32) continue;
      if (lo >= bit) v |= a.l[k] << (lo - bit);
      else v |= a.l[k] >> (bit - lo);
    }
    w[j] = v;
  }
}

// Rows live in device memory. Pointers that reach a kernel through a descriptor in memory are
// generic to the compiler, which would then issue flat accesses (and make every LDS wait also
// wait for them); the row accessors address the global space explicitly.
template <class T>
using GlobalPtr = __attribute__((address_space(1))) T*;
using U32x4 = uint32_t __attribute__((ext_vector_type(4)));  // a 16-byte access

template <class T>
__device__ __forceinline__ GlobalPtr<const T> global_ptr(const void* p) {
  return (GlobalPtr<const T>)p;
}

template <class T>
__device__ __forceinline__ GlobalPtr<T> global_ptr(void* p) {
  return (GlobalPtr<T>)p;
}

// Loads one 32-byte row (two 16-byte loads).
__device__ __forceinline__ Fe load_row(const uint64_t* row) {
  const GlobalPtr<const U32x4> p = global_ptr<U32x4>(row);
  const U32x4 x = p[0], y = p[1];
  uint32_t w[8] = {x[0], x[1], x[2], x[3], y[0], y[1], y[2], y[3]};
  return from_words(w);
}

__device__ __forceinline__ void store_row(uint64_t* row, const Fe& a) {
  uint32_t w[8];
  to_words(a, w);
  const GlobalPtr<U32x4> p = global_ptr<U32x4>(row);
  p[0] = U32x4{w[0], w[1], w[2], w[3]};
  p[1] = U32x4{w[4], w[5], w[6], w[7]};
}

// Canonical plain value < 2^256 from raw 32-bit words, reduced mod p, in the 256-form.
template <class F>
__device__ __forceinline__ Fe plain_to_256(const F& f, const Fe& x) {
  return mont(f, x, fe_const<F>(f.c517));  // x * 2^517 * 2^-261 = x * 2^256, reduced
}

template <class F>
__device__ __forceinline__ Fe form256_to_plain(const F& f, const Fe& x) {
  Fe c = fe_zero();
  c.l[0] = 32;  // x * 2^5 * 2^-261 = a * 2^256 * 2^-256
  return mont(f, x, c);
}

template <class F>
__device__ __forceinline__ Fe form256_to_261(const F& f, const Fe& x) {
  if constexpr (HasP4r<F>::value) {
    // x 2^5 mod p without a product (built-in moduli, x < 2^256 with normalized limbs): the limbs
    // shifted up 5 bits, then less a quotient estimated from the top limb (lc_finish: exact for
    // values below 2^261, the result below 2p), then one conditional subtraction.
    Acc64 a;
    a.l[0] = (int64_t)((x.l[0] << 5) & LMASK);
#pragma unroll
    for (int i = 1; i < NL - 1; ++i) a.l[i] = (int64_t)(((x.l[i] << 5) | (x.l[i - 1] >> (LB - 5))) & LMASK);
    a.l[NL - 1] = (int64_t)((x.l[NL - 1] << 5) | (x.l[NL - 2] >> (LB - 5)));
    return canonical(f, lc_finish(f, a));
  } else {
    return mont(f, x, fe_const<F>(f.c266));
  }
}

template <class F>
__device__ __forceinline__ Fe form261_to_256(const F& f, const Fe& x) {
  return mont(f, x, fe_const<F>(f.c256));
}

// a * b in the 256-form for two 256-form operands: (a * b * 2^-261) * 2^5, the second factor as
// form256_to_261's shift and quotient for the built-in moduli (one product instead of two), else
// a * (b * 2^5) * 2^-261.
template <class F>
__device__ __forceinline__ Fe mul256(const F& f, const Fe& a, const Fe& b) {
  if constexpr (HasP4r<F>::value) return form256_to_261(f, mont_raw(f, a, b));  // below 2p < 2^256
  else return mont(f, a, mont(f, b, fe_const<F>(f.c266)));
}

// x^e for a 261-form x and a 256-bit exponent (plain words, little-endian), in the 261-form.
template <class F>
__device__ __forceinline__ Fe fpow(const F& f, const Fe& x, const uint32_t e[8]) {
  Fe acc = fe_const<F>(f.one);
  for (int i = 255; i >= 0; --i) {
    acc = mont_sqr(f, acc);
    if ((e[i >> 5] >> (i & 31)) & 1u) acc = mont(f, acc, x);
  }
  return acc;
}

// x^-1 for a 261-form x != 0 (Fermat), in the 261-form.
template <class F>
__device__ __forceinline__ Fe finv(const F& f, const Fe& x) {
  uint32_t e[8];
  uint32_t borrow = 2;  // e = p - 2
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    uint32_t w = f.p32[i];
    e[i] = w - borrow;
    borrow = w < borrow ? 1u : 0u;
  }
  return fpow(f, x, e);
}

}  // namespace zk
