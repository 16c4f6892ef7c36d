// Field registry shared by the kernels' host code: the two built-in BN254
// fields (compile-time constants) and moduli registered at run time.
#pragma once

#include <stdint.h>

#include "zk_field.hpp"
#include "zk_runtime.hpp"

namespace zk {

// Runtime parameters of a registered field id (>= 2); false if unknown.
bool lookup_field(int id, FieldRT* out);

// Host-side constants of an odd modulus in [3, 2^255) given as four u64 limbs.
bool make_field_rt(const uint64_t modulus[4], FieldRT* out, std::string* err);

// Calls fn(FrConst{}), fn(FqConst{}) or fn(FieldRT) for a field id, so one
// generic lambda instantiates the kernels for every field type.
template <class Fn>
int with_field(int id, Fn&& fn) {
  if (id == 0) return fn(FrConst{});
  if (id == 1) return fn(FqConst{});
  FieldRT f;
  if (!lookup_field(id, &f)) return fail("unknown field id " + std::to_string(id));
  return fn(f);
}

inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace zk
