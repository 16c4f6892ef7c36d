// Host build of the MSM's scalar recoding (zk_scalar.hpp) for tests/test_glv.py: the same
// functions the digit kernel runs, called on the CPU and checked against Python integers.
#include <stdint.h>

#include "../../zk-gpu-accelerator_amd/csrc/zk_scalar.hpp"

extern "C" {

void zkt_reduce(const uint32_t* in, uint32_t* out) {
  uint32_t w[8];
  for (int i = 0; i < 8; ++i) w[i] = in[i];
  zk::reduce_mod_r(w);
  for (int i = 0; i < 8; ++i) out[i] = w[i];
}

void zkt_glv(const uint32_t* k, uint32_t* m1, int* neg1, uint32_t* m2, int* neg2) {
  uint32_t kk[8], a[8], b[8];
  for (int i = 0; i < 8; ++i) kk[i] = k[i];
  bool n1 = false, n2 = false;
  zk::glv_split(kk, a, n1, b, n2);
  for (int i = 0; i < 8; ++i) {
    m1[i] = a[i];
    m2[i] = b[i];
  }
  *neg1 = n1;
  *neg2 = n2;
}

// The signed digits of v over `windows` windows of c bits; returns the carry out of the last.
int zkt_digits(const uint32_t* v, int c, int windows, int32_t* out) {
  uint32_t w[8];
  for (int i = 0; i < 8; ++i) w[i] = v[i];
  uint32_t carry = 0;
  for (int win = 0; win < windows; ++win) out[win] = zk::signed_digit(w, win, c, carry);
  return (int)carry;
}

// The digits of a GLV half as the digit kernel stores them: with `neg`, ties go negative below the
// top window and every digit is negated (the half is -v).
int zkt_half_digits(const uint32_t* v, int c, int windows, int neg, int32_t* out) {
  uint32_t w[8];
  for (int i = 0; i < 8; ++i) w[i] = v[i];
  uint32_t carry = 0;
  for (int win = 0; win < windows; ++win) {
    const int d = zk::signed_digit(w, win, c, carry, neg && win + 1 < windows);
    out[win] = neg ? -d : d;
  }
  return (int)carry;
}

}  // extern "C"
