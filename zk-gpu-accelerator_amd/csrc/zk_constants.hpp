// Generated by gen_constants.py from the BN254 moduli; do not edit.
// Each field lists its modulus p in 29-bit limbs, -p^-1 mod 2^29 and the
// multipliers between representations (see zk_field.hpp).
#pragma once
#include <stdint.h>

namespace zk {

struct FrConst {
  // modulus 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001
  static constexpr uint32_t p[9] = {0x10000001u, 0x1f0fac9fu, 0x0e5c2450u, 0x07d090f3u, 0x1585d283u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr uint32_t p2[9] = {0x00000002u, 0x1e1f593fu, 0x1cb848a1u, 0x0fa121e6u, 0x0b0ba506u, 0x05b68181u, 0x014dc282u, 0x1cb84c68u, 0x0060c89cu};
  static constexpr uint32_t p4r[9] = {0x20000004u, 0x3c3eb27du, 0x39709142u, 0x3f4243ccu, 0x36174a0bu, 0x2b6d0301u, 0x229b8503u, 0x397098cfu, 0x00c19138u};
  static constexpr uint32_t p8r[9] = {0x40000008u, 0x587d64fau, 0x52e12285u, 0x5e848799u, 0x4c2e9417u, 0x56da0603u, 0x45370a06u, 0x52e1319eu, 0x01832271u};
  static constexpr uint32_t c256[9] = {0x0ffffffbu, 0x04b1a0e2u, 0x18334a6bu, 0x18ed2b3eu, 0x1462e36fu, 0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};
  static constexpr uint32_t c266[9] = {0x0fffead7u, 0x1d5444f4u, 0x04438aa5u, 0x03b4d096u, 0x134c84dau, 0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};
  static constexpr uint32_t c517[9] = {0x142db4dfu, 0x19d6990eu, 0x1472f48cu, 0x06dbe7e3u, 0x0b84d579u, 0x10f9faf7u, 0x121f4380u, 0x17a112deu, 0x001275c7u};
  static constexpr uint32_t c522[9] = {0x05b69bd4u, 0x06170a5au, 0x020cddceu, 0x1db6310bu, 0x0e54d0ffu, 0x1cf855e3u, 0x1c15e103u, 0x07d09161u, 0x000a054au};
  static constexpr uint32_t one[9] = {0x0fffff57u, 0x1ea70ab4u, 0x052c068bu, 0x17504f49u, 0x0aa8075bu, 0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};
  static constexpr uint32_t mone[9] = {0x000000aau, 0x0068a1ebu, 0x09301dc5u, 0x108041aau, 0x0addcb27u, 0x0598fff2u, 0x0ed19539u, 0x09315e8bu, 0x00229c18u};
  static constexpr uint32_t p32[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u, 0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr uint32_t ninv = 0x0fffffffu;
};

struct FqConst {
  // modulus 0x30644e72e131a029b85045b68181585d97816a916871ca8d3c208c16d87cfd47
  static constexpr uint32_t p[9] = {0x187cfd47u, 0x010460b6u, 0x1c72a34fu, 0x02d522d0u, 0x1585d978u, 0x02db40c0u, 0x00a6e141u, 0x0e5c2634u, 0x0030644eu};
  static constexpr uint32_t p2[9] = {0x10f9fa8eu, 0x0208c16du, 0x18e5469eu, 0x05aa45a1u, 0x0b0bb2f0u, 0x05b68181u, 0x014dc282u, 0x1cb84c68u, 0x0060c89cu};
  static constexpr uint32_t p4r[9] = {0x21f3f51cu, 0x241182dau, 0x31ca8d3bu, 0x2b548b42u, 0x361765dfu, 0x2b6d0301u, 0x229b8503u, 0x397098cfu, 0x00c19138u};
  static constexpr uint32_t p8r[9] = {0x43e7ea38u, 0x482305b4u, 0x43951a76u, 0x56a91685u, 0x4c2ecbbeu, 0x56da0603u, 0x45370a06u, 0x52e1319eu, 0x01832271u};
  static constexpr uint32_t c256[9] = {0x058f0d9du, 0x1aea1c6eu, 0x11c2cf74u, 0x11d651ebu, 0x1462c0a7u, 0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u};
  static constexpr uint32_t c266[9] = {0x13349ca1u, 0x1a5d84a8u, 0x0a3e5cacu, 0x100249e0u, 0x12b951e8u, 0x0e92d304u, 0x14cb95b3u, 0x041b9d3du, 0x00058003u};
  static constexpr uint32_t c517[9] = {0x0f6b5c04u, 0x08ead878u, 0x1645525du, 0x1aefe9cdu, 0x09d605edu, 0x0483a115u, 0x0d08508bu, 0x0dba4804u, 0x001982b4u};
  static constexpr uint32_t c522[9] = {0x059bac10u, 0x0d1503a3u, 0x018016b8u, 0x10ab0ca8u, 0x02632639u, 0x02c0169fu, 0x169bfd53u, 0x11869d4cu, 0x002a11a6u};
  static constexpr uint32_t one[9] = {0x157ccc21u, 0x141c2758u, 0x185230d3u, 0x014c0419u, 0x0aa36fb9u, 0x1d4240ceu, 0x11d54c07u, 0x052ac7a8u, 0x000dc836u};
  static constexpr uint32_t mone[9] = {0x03003126u, 0x0ce8395eu, 0x0420727bu, 0x01891eb7u, 0x0ae269bfu, 0x0598fff2u, 0x0ed19539u, 0x09315e8bu, 0x00229c18u};
  static constexpr uint32_t p32[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u, 0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
  static constexpr uint32_t ninv = 0x04866389u;
};

// The same fields, tagged to use the low-latency Montgomery product (zk_field.hpp).
struct FrIlp : FrConst {
  static constexpr bool kIlp = true;
};
struct FqIlp : FqConst {
  static constexpr bool kIlp = true;
};

// GLV split of a scalar k (< r) into k1 + k2 lambda with |k1|, |k2| < 2^127:
// c1 = floor(k g1 / 2^256), c2 = floor(k g2 / 2^256), k1 = k - c1 a1 - c2 a2,
// k2 = c1 b1n - c2 b2, over the lattice basis (a1, -b1n), (a2, b2) of
// {(a, b): a + b lambda = 0 mod r}. (beta x, y) = lambda (x, y) on G1.
// lambda = 0xb3c4d79d41a917585bfc41088d8daaa78b17ea66b99c90dd
struct Glv {
  static constexpr uint32_t beta261[9] = {0x0a337995u, 0x158d1d23u, 0x189c9b98u, 0x12fa4e45u, 0x185faadcu, 0x0176f16du, 0x0eed93bau, 0x14291140u, 0x000c0afeu};
  static constexpr uint32_t a1[2] = {0x94d213e3u, 0x89d32568u};
  static constexpr uint32_t b1n[4] = {0x7d4f1128u, 0x8211bbebu, 0xeeb859fcu, 0x6f4d8248u};
  static constexpr uint32_t a2[4] = {0x1221250bu, 0x0be4e154u, 0xeeb859fdu, 0x6f4d8248u};
  static constexpr uint32_t b2[2] = {0x94d213e3u, 0x89d32568u};
  static constexpr uint32_t g1[3] = {0xc7e0b3d7u, 0xd91d232eu, 0x00000002u};
  static constexpr uint32_t g2[5] = {0x391eb18eu, 0x7a7bd9d4u, 0xa773d2cfu, 0x4ccef014u, 0x00000002u};
  static constexpr int kBits = 127;  // |k1|, |k2| < 2^kBits
};

}  // namespace zk
