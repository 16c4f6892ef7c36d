"""
Writes zk_constants.hpp: BN254 Fr/Fq parameters in the 29-bit limb form the kernels use.

Run from this directory: python gen_constants.py. The numbers are derived
here from the two moduli, so the header can be regenerated and checked
(tests/test_abi.py and tests/test_glv.py recompute every entry).
"""

R_SCALAR = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P_BASE = 21888242871839275222246405745257275088696311157297823662689037894645226208583
LIMB_BITS = 29
NLIMBS = 9
RBITS = LIMB_BITS * NLIMBS  # 261


def limbs29(v: int) -> list[int]:
    return [(v >> (LIMB_BITS * i)) & ((1 << LIMB_BITS) - 1) for i in range(NLIMBS)]


def redundant29(v: int, lift: int = 1) -> list[int]:
    """v in 29-bit limbs, each of limbs 0..7 raised by lift * 2^29 (borrowed from the next limb), so
    that a + v - b stays non-negative limb by limb for a b whose limbs 0..7 are at most
    lift * 2^29 - lift (normalized b for lift 1, a sum of two normalized values for lift 2)."""
    l = limbs29(v)
    up = lift << LIMB_BITS
    r = [l[0] + up] + [l[i] + up - lift for i in range(1, NLIMBS - 1)] + [l[NLIMBS - 1] - lift]
    assert sum(x << (LIMB_BITS * i) for i, x in enumerate(r)) == v and r[-1] >= 0
    return r


def field_constants(m: int) -> dict[str, list[int] | int]:
    return {
        "p": limbs29(m),
        "p2": limbs29(2 * m),              # 2p, the bound of lazily reduced values
        "p4r": redundant29(4 * m),         # 4p with every limb below the top at least 2^29 - 1
        "p8r": redundant29(8 * m, 2),      # 8p with every limb below the top at least 2^30 - 2
        "ninv": (-pow(m, -1, 1 << LIMB_BITS)) % (1 << LIMB_BITS),
        "c256": limbs29(pow(2, 256, m)),   # multiplier taking a 2^261 form to the 2^256 form
        "c266": limbs29(pow(2, 266, m)),   # multiplier taking the 2^256 form to the 2^261 form
        "c517": limbs29(pow(2, 517, m)),   # multiplier taking a plain value to the 2^256 form
        "c522": limbs29(pow(2, 522, m)),   # multiplier taking a plain value to the 2^261 form
        "one": limbs29(pow(2, RBITS, m)),  # 1 in the 2^261 form
        "mone": limbs29(-pow(2, RBITS, m) % m),  # -1 in the 2^261 form
        "p32": [(m >> (32 * i)) & 0xFFFFFFFF for i in range(8)],
    }


def emit(name: str, m: int) -> str:
    c = field_constants(m)
    out = [f"struct {name} {{", f"  // modulus {m:#x}"]
    for key in ("p", "p2", "p4r", "p8r", "c256", "c266", "c517", "c522", "one", "mone"):
        vals = ", ".join(f"0x{v:08x}u" for v in c[key])
        out.append(f"  static constexpr uint32_t {key}[9] = {{{vals}}};")
    out.append(f"  static constexpr uint32_t p32[8] = {{{', '.join(f'0x{v:08x}u' for v in c['p32'])}}};")
    out.append(f"  static constexpr uint32_t ninv = 0x{c['ninv']:08x}u;")
    out.append("};")
    return "\n".join(out)


# ---- GLV: the endomorphism (x, y) -> (beta x, y) = lambda (x, y) of BN254 G1 -----------------


def _sqrt_mod(a: int, p: int) -> int:
    """Tonelli-Shanks square root modulo an odd prime."""
    q, s = p - 1, 0
    while q % 2 == 0:
        q, s = q // 2, s + 1
    z = 2
    while pow(z, (p - 1) // 2, p) != p - 1:
        z += 1
    m, c, t, r = s, pow(z, q, p), pow(a, q, p), pow(a, (q + 1) // 2, p)
    while t != 1:
        i, t2 = 0, t
        while t2 != 1:
            t2, i = t2 * t2 % p, i + 1
        b = pow(c, 1 << (m - i - 1), p)
        m, c, t, r = i, b * b % p, t * b * b % p, r * b % p
    return r


def _cube_roots(p: int) -> list[int]:
    """The two primitive cube roots of unity mod p: (-1 +- sqrt(-3)) / 2."""
    s = _sqrt_mod(p - 3, p)
    half = pow(2, -1, p)
    return sorted({(-1 + s) * half % p, (-1 - s) * half % p})


def _affine_mul(k: int, pt: tuple[int, int], p: int) -> tuple[int, int] | None:
    """k * pt on y^2 = x^3 + 3 over F_p by double-and-add in affine coordinates."""
    acc, cur = None, pt
    while k:
        if k & 1:
            acc = _affine_add(acc, cur, p)
        cur, k = _affine_add(cur, cur, p), k >> 1
    return acc


def _affine_add(a, b, p):
    if a is None:
        return b
    if b is None:
        return a
    if a[0] == b[0]:
        if (a[1] + b[1]) % p == 0:
            return None
        m = 3 * a[0] * a[0] * pow(2 * a[1], -1, p) % p
    else:
        m = (b[1] - a[1]) * pow(b[0] - a[0], -1, p) % p
    x = (m * m - a[0] - b[0]) % p
    return x, (m * (a[0] - x) - a[1]) % p


def glv_constants() -> dict[str, int]:
    """lambda (the smaller cube root of unity mod r), the beta with (beta x, y) = lambda (x, y),
    and a reduced lattice basis (a1, b1), (a2, b2) of {(a, b): a + b lambda = 0 mod r}
    (extended Euclid on (r, lambda), Gallant-Lambert-Vanstone 2001, section 4)."""
    from math import isqrt

    lam = _cube_roots(R_SCALAR)[0]
    g = (1, 2)
    lg = _affine_mul(lam, g, P_BASE)
    beta = next(b for b in _cube_roots(P_BASE) if lg == (b * g[0] % P_BASE, g[1]))
    rs, ts = [R_SCALAR, lam], [0, 1]
    while rs[-1]:
        q = rs[-2] // rs[-1]
        rs.append(rs[-2] - q * rs[-1])
        ts.append(ts[-2] - q * ts[-1])
    m = max(i for i in range(len(rs)) if rs[i] >= isqrt(R_SCALAR))
    a1, b1 = rs[m + 1], -ts[m + 1]
    a2, b2 = min((rs[m], -ts[m]), (rs[m + 2], -ts[m + 2]), key=lambda v: v[0] ** 2 + v[1] ** 2)
    assert (a1 + b1 * lam) % R_SCALAR == 0 and (a2 + b2 * lam) % R_SCALAR == 0
    assert a1 * b2 - a2 * b1 == R_SCALAR
    # the decomposition below relies on these signs
    assert a1 > 0 and b1 < 0 and a2 > 0 and b2 > 0
    g1 = (b2 * (1 << 256) + R_SCALAR // 2) // R_SCALAR   # round(b2 2^256 / r)
    g2 = (-b1 * (1 << 256) + R_SCALAR // 2) // R_SCALAR  # round(-b1 2^256 / r)
    # |k1|, |k2| < 1.125 (|a1| + |a2|) and 1.125 (|b1| + |b2|): both below 2^127
    assert 9 * (a1 + a2) < 8 << 127 and 9 * (-b1 + b2) < 8 << 127
    return {"lambda": lam, "beta": beta, "a1": a1, "b1n": -b1, "a2": a2, "b2": b2, "g1": g1, "g2": g2}


def words32(v: int, n: int) -> list[int]:
    assert v < 1 << (32 * n)
    return [(v >> (32 * i)) & 0xFFFFFFFF for i in range(n)]


def emit_glv() -> str:
    c = glv_constants()
    sizes = {"a1": 2, "b1n": 4, "a2": 4, "b2": 2, "g1": 3, "g2": 5}
    out = [
        "// GLV split of a scalar k (< r) into k1 + k2 lambda with |k1|, |k2| < 2^127:",
        "// c1 = floor(k g1 / 2^256), c2 = floor(k g2 / 2^256), k1 = k - c1 a1 - c2 a2,",
        "// k2 = c1 b1n - c2 b2, over the lattice basis (a1, -b1n), (a2, b2) of",
        "// {(a, b): a + b lambda = 0 mod r}. (beta x, y) = lambda (x, y) on G1.",
        f"// lambda = {c['lambda']:#x}",
        "struct Glv {",
        f"  static constexpr uint32_t beta261[9] = {{{', '.join(f'0x{v:08x}u' for v in limbs29(c['beta'] * (1 << RBITS) % P_BASE))}}};",
    ]
    for key, n in sizes.items():
        out.append(f"  static constexpr uint32_t {key}[{n}] = {{{', '.join(f'0x{v:08x}u' for v in words32(c[key], n))}}};")
    out.append("  static constexpr int kBits = 127;  // |k1|, |k2| < 2^kBits")
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
    text = (HEADER + "\n" + emit("FrConst", R_SCALAR) + "\n\n" + emit("FqConst", P_BASE) + tail + "\n"
            + emit_glv() + "\n\n}  // namespace zk\n")
    with open("zk_constants.hpp", "w") as f:
        f.write(text)
