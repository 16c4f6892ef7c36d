// Scalar recoding for the MSM: reduction mod r, the GLV split and signed c-bit digits.
//
// Plain integer arithmetic on little-endian 32-bit words, usable on the device
// (the MSM's digit kernel) and on the host (the test harness compiles this
// header with a host compiler and checks it against Python integers).
#pragma once

#include <stdint.h>

#include "zk_constants.hpp"

#if defined(__HIPCC__)
#define ZK_HD __host__ __device__ __forceinline__
#else
#define ZK_HD inline
#endif

namespace zk {

// r, the BN254 group order, in little-endian 32-bit words.
struct ScalarOrder {
  static constexpr uint32_t r[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                                    0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
};

// w mod r in place: any 256-bit value is below 6r, so at most five subtractions.
ZK_HD void reduce_mod_r(uint32_t (&w)[8]) {
  for (int rep = 0; rep < 6; ++rep) {
    uint32_t d[8], borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint64_t t = (uint64_t)w[i] - ScalarOrder::r[i] - borrow;
      d[i] = (uint32_t)t;
      borrow = (uint32_t)(t >> 63);
    }
    if (borrow) break;  // w < r
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = d[i];
  }
}

// c bits of w starting at bit `bit` (zero past the top word); c <= 20.
ZK_HD uint32_t raw_window(const uint32_t (&w)[8], int bit, int c) {
  const int word = bit >> 5, sh = bit & 31;
  if (word >= 8) return 0;
  uint64_t v = w[word];
  if (word + 1 < 8) v |= (uint64_t)w[word + 1] << 32;
  return (uint32_t)(v >> sh) & ((1u << c) - 1u);
}

// Signed digit of window `win` given the carry in, in [-2^(c-1), 2^(c-1)]; updates the carry.
// Over windows 0..floor(B/c) the digits of any value below 2^B satisfy sum d_w 2^(cw) = value.
// A window of exactly 2^(c-1) gives +2^(c-1), or with ties_negative -2^(c-1) and a carry (either
// choice in any window below the top one leaves the sum unchanged), so that one window's digits
// lie in [-(2^(c-1) - 1), 2^(c-1)], the other's in [-2^(c-1), 2^(c-1) - 1].
ZK_HD int signed_digit(const uint32_t (&w)[8], int win, int c, uint32_t& carry, bool ties_negative = false) {
  const uint32_t raw = raw_window(w, win * c, c) + carry;
  const uint32_t half = 1u << (c - 1);
  if (raw > half || (ties_negative && raw == half)) {
    carry = 1;
    return (int)raw - (int)(1u << c);
  }
  carry = 0;
  return (int)raw;
}

// out[0 .. NA + NB) += a * b (schoolbook on 32-bit words).
template <int NA, int NB>
ZK_HD void words_mul(const uint32_t (&a)[NA], const uint32_t* b, uint32_t* out) {
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    uint64_t carry = 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const uint64_t t = (uint64_t)a[i] * b[j] + out[i + j] + carry;  // < 2^64
      out[i + j] = (uint32_t)t;
      carry = t >> 32;
    }
#pragma unroll
    for (int j = i + NB; j < NA + NB; ++j) {
      const uint64_t t = (uint64_t)out[j] + carry;
      out[j] = (uint32_t)t;
      carry = t >> 32;
    }
  }
}

// x -= y (ny words, zero-extended) over 8 words, mod 2^256.
ZK_HD void words_sub8(uint32_t (&x)[8], const uint32_t* y, int ny) {
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint64_t t = (uint64_t)x[i] - (i < ny ? y[i] : 0u) - borrow;
    x[i] = (uint32_t)t;
    borrow = (uint32_t)(t >> 63);
  }
}

// |x| of an 8-word two's-complement value; returns whether x was negative.
ZK_HD bool words_abs8(uint32_t (&x)[8]) {
  const bool neg = x[7] >> 31;
  if (neg) {
    uint32_t carry = 1;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint64_t t = (uint64_t)(~x[i]) + carry;
      x[i] = (uint32_t)t;
      carry = (uint32_t)(t >> 32);
    }
  }
  return neg;
}

// The GLV split of k (< r): k = k1 + k2 lambda (mod r), |k1|, |k2| < 2^127, with
// c1 = floor(k g1 / 2^256), c2 = floor(k g2 / 2^256), k1 = k - c1 a1 - c2 a2, k2 = c1 b1n - c2 b2
// (constants in Glv, zk_constants.hpp). m1, m2 receive |k1|, |k2| (high words zero).
ZK_HD void glv_split(const uint32_t (&k)[8], uint32_t (&m1)[8], bool& neg1, uint32_t (&m2)[8], bool& neg2) {
  uint32_t t1[11] = {0}, t2[13] = {0};
  words_mul<8, 3>(k, Glv::g1, t1);
  words_mul<8, 5>(k, Glv::g2, t2);
  const uint32_t c1[3] = {t1[8], t1[9], t1[10]};
  const uint32_t c2[5] = {t2[8], t2[9], t2[10], t2[11], t2[12]};
  uint32_t p1[5] = {0}, p2[9] = {0};
  words_mul<3, 2>(c1, Glv::a1, p1);
  words_mul<5, 4>(c2, Glv::a2, p2);
#pragma unroll
  for (int i = 0; i < 8; ++i) m1[i] = k[i];
  words_sub8(m1, p1, 5);
  words_sub8(m1, p2, 8);  // c2 a2 < 2^255: its ninth word is zero
  uint32_t q1[7] = {0}, q2[7] = {0};
  words_mul<3, 4>(c1, Glv::b1n, q1);
  words_mul<5, 2>(c2, Glv::b2, q2);
#pragma unroll
  for (int i = 0; i < 8; ++i) m2[i] = i < 7 ? q1[i] : 0u;
  words_sub8(m2, q2, 7);
  neg1 = words_abs8(m1);
  neg2 = words_abs8(m2);
}

}  // namespace zk
