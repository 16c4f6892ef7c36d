"""
Writes zk_constants.hpp: BN254 Fr/Fq parameters in the 29-bit limb form the kernels use.

Run from this directory: python gen_constants.py. The numbers are derived
here from the two moduli, so the header can be regenerated and checked
(tests/test_constants.py recomputes every entry).
"""

R_SCALAR = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P_BASE = 21888242871839275222246405745257275088696311157297823662689037894645226208583
LIMB_BITS = 29
NLIMBS = 9
RBITS = LIMB_BITS * NLIMBS  # 261


def limbs29(v: int) -> list[int]:
    return [(v >> (LIMB_BITS * i)) & ((1 << LIMB_BITS) - 1) for i in range(NLIMBS)]


def field_constants(m: int) -> dict[str, list[int] | int]:
    return {
        "p": limbs29(m),
        "ninv": (-pow(m, -1, 1 << LIMB_BITS)) % (1 << LIMB_BITS),
        "c256": limbs29(pow(2, 256, m)),   # multiplier taking a 2^261 form to the 2^256 form
        "c266": limbs29(pow(2, 266, m)),   # multiplier taking the 2^256 form to the 2^261 form
        "c517": limbs29(pow(2, 517, m)),   # multiplier taking a plain value to the 2^256 form
        "c522": limbs29(pow(2, 522, m)),   # multiplier taking a plain value to the 2^261 form
        "one": limbs29(pow(2, RBITS, m)),  # 1 in the 2^261 form
        "p32": [(m >> (32 * i)) & 0xFFFFFFFF for i in range(8)],
    }


def emit(name: str, m: int) -> str:
    c = field_constants(m)
    out = [f"struct {name} {{", f"  // modulus {m:#x}"]
    for key in ("p", "c256", "c266", "c517", "c522", "one"):
        vals = ", ".join(f"0x{v:08x}u" for v in c[key])
        out.append(f"  static constexpr uint32_t {key}[9] = {{{vals}}};")
    out.append(f"  static constexpr uint32_t p32[8] = {{{', '.join(f'0x{v:08x}u' for v in c['p32'])}}};")
    out.append(f"  static constexpr uint32_t ninv = 0x{c['ninv']:08x}u;")
    out.append("};")
    return "\n".join(out)


HEADER = """// Generated by gen_constants.py from the BN254 moduli; do not edit.
// Each field lists its modulus p in 29-bit limbs, -p^-1 mod 2^29 and the
// multipliers between representations (see zk_field.hpp).
#pragma once
#include <stdint.h>

namespace zk {
"""

if __name__ == "__main__":
    tail = (
        "\n\n// The same fields, tagged to use the low-latency Montgomery product (zk_field.hpp).\n"
        "struct FrIlp : FrConst {\n  static constexpr bool kIlp = true;\n};\n"
        "struct FqIlp : FqConst {\n  static constexpr bool kIlp = true;\n};\n"
    )
    text = HEADER + "\n" + emit("FrConst", R_SCALAR) + "\n\n" + emit("FqConst", P_BASE) + tail + "\n}  // namespace zk\n"
    with open("zk_constants.hpp", "w") as f:
        f.write(text)
