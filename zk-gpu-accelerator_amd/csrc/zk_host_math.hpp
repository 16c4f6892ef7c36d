// Host-side modular arithmetic on 4x64-bit limbs (Montgomery, R = 2^256),
// used to derive table seeds (roots of unity, coset powers, inverses) before
// the device expands them. Not on any hot path.
#pragma once

#include <stdint.h>
#include <string.h>

namespace zk {

struct U256 {
  uint64_t w[4] = {0, 0, 0, 0};
  static U256 of(uint64_t x) {
    U256 r;
    r.w[0] = x;
    return r;
  }
  static U256 from(const uint64_t* p) {
    U256 r;
    memcpy(r.w, p, 32);
    return r;
  }
  bool is_zero() const { return (w[0] | w[1] | w[2] | w[3]) == 0; }
  bool bit(int i) const { return (w[i >> 6] >> (i & 63)) & 1; }
};

class HostField {
 public:
  explicit HostField(const U256& p) : p_(p) {
    uint64_t inv = 1;
    for (int i = 0; i < 7; ++i) inv *= 2 - p.w[0] * inv;
    ninv_ = 0 - inv;
    U256 x = U256::of(1);
    for (int i = 0; i < 512; ++i) {
      if (i == 256) one_ = x;
      x = add(x, x);
    }
    r2_ = x;
  }
  const U256& p() const { return p_; }
  U256 add(const U256& a, const U256& b) const {
    U256 r;
    unsigned __int128 c = 0;
    for (int i = 0; i < 4; ++i) {
      c += (unsigned __int128)a.w[i] + b.w[i];
      r.w[i] = (uint64_t)c;
      c >>= 64;
    }
    if (c || geq(r, p_)) r = sub_raw(r, p_);
    return r;
  }
  U256 sub(const U256& a, const U256& b) const {
    if (geq(a, b)) return sub_raw(a, b);
    return sub_raw(add_raw(a, p_), b);
  }
  U256 mul(const U256& a, const U256& b) const {  // Montgomery product a*b*2^-256
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
      unsigned __int128 c = 0;
      for (int j = 0; j < 4; ++j) {
        c += (unsigned __int128)a.w[j] * b.w[i] + t[j];
        t[j] = (uint64_t)c;
        c >>= 64;
      }
      c += t[4];
      t[4] = (uint64_t)c;
      t[5] = (uint64_t)(c >> 64);
      uint64_t m = t[0] * ninv_;
      c = (unsigned __int128)m * p_.w[0] + t[0];
      c >>= 64;
      for (int j = 1; j < 4; ++j) {
        c += (unsigned __int128)m * p_.w[j] + t[j];
        t[j - 1] = (uint64_t)c;
        c >>= 64;
      }
      c += t[4];
      t[3] = (uint64_t)c;
      t[4] = t[5] + (uint64_t)(c >> 64);
    }
    U256 r;
    memcpy(r.w, t, 32);
    if (t[4] || geq(r, p_)) r = sub_raw(r, p_);
    return r;
  }
  U256 to_mont(const U256& plain) const { return mul(reduce(plain), r2_); }
  U256 from_mont(const U256& m) const { return mul(m, U256::of(1)); }
  U256 one() const { return one_; }
  U256 pow(const U256& base_m, const U256& e) const {
    U256 acc = one_;
    for (int i = 255; i >= 0; --i) {
      acc = mul(acc, acc);
      if (e.bit(i)) acc = mul(acc, base_m);
    }
    return acc;
  }
  U256 inv(const U256& a_m) const {
    U256 e = sub_raw(p_, U256::of(2));
    return pow(a_m, e);
  }
  // Plain helpers.
  U256 mulp(const U256& a, const U256& b) const { return from_mont(mul(to_mont(a), to_mont(b))); }
  U256 powp(const U256& a, const U256& e) const { return from_mont(pow(to_mont(a), e)); }
  U256 invp(const U256& a) const { return from_mont(inv(to_mont(a))); }
  U256 reduce(U256 a) const {
    while (geq(a, p_)) a = sub_raw(a, p_);
    return a;
  }

  static bool geq(const U256& a, const U256& b) {
    for (int i = 3; i >= 0; --i)
      if (a.w[i] != b.w[i]) return a.w[i] > b.w[i];
    return true;
  }
  static U256 sub_raw(const U256& a, const U256& b) {
    U256 r;
    unsigned __int128 br = 0;
    for (int i = 0; i < 4; ++i) {
      unsigned __int128 d = (unsigned __int128)a.w[i] - b.w[i] - br;
      r.w[i] = (uint64_t)d;
      br = (d >> 64) ? 1 : 0;
    }
    return r;
  }
  static U256 add_raw(const U256& a, const U256& b) {
    U256 r;
    unsigned __int128 c = 0;
    for (int i = 0; i < 4; ++i) {
      c += (unsigned __int128)a.w[i] + b.w[i];
      r.w[i] = (uint64_t)c;
      c >>= 64;
    }
    return r;
  }

 private:
  U256 p_, one_, r2_;
  uint64_t ninv_;
};

}  // namespace zk
