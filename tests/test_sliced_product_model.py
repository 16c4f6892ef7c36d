"""CPU model of the limb-sliced Montgomery product (csrc/zk_curve_sliced.hpp: sl_mul, sl_carry64,
sl_carry), lane by lane in the device's integer widths.

The product runs as nine CIOS steps across a 16-lane row: lane k holds a 64-bit column; a step
adds a_i b, adds m p for a quotient m taken from lane 0, and shifts the row down one lane, keeping
each column's bits above 29 in place as the next column's carry. Only the last quotient is masked
to 29 bits. These tests replay that arithmetic with the inputs at the ends of the stated ranges and
check the three facts the kernels rely on: no column reaches 2^62 in magnitude (the device adds in
64 bits), the result is a b 2^-261 mod p, and its magnitude is below |a b| / 2^261 + p (1 + 2^-26).
"""
import random

import pytest

P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
LB = 29
MASK = (1 << LB) - 1
NL = 9
NINV = (-pow(P, -1, 1 << LB)) % (1 << LB)  # the field's 29-bit Montgomery constant
R_INV = pow(1 << (LB * NL), -1, P)
P_LIMBS = [(P >> (LB * k)) & MASK for k in range(NL)]


def wrap64(v):
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >> 63 else v


def wrap32(v):
    v &= (1 << 32) - 1
    return v - (1 << 32) if v >> 31 else v


def value(limbs):
    return sum(l << (LB * k) for k, l in enumerate(limbs[:NL]))


def sl_carry64(v):
    """Lanes 0..7 keep 29 bits and pass (int32)(v >> 29) up one lane; lane 8 keeps its low 32 bits."""
    h = [wrap32(x >> LB) if k < NL - 1 else 0 for k, x in enumerate(v)]
    out = []
    for k, x in enumerate(v):
        low = (x & MASK) if k < NL - 1 else (wrap32(x) if k == NL - 1 else 0)
        out.append(wrap32(low + (h[k - 1] if k > 0 else 0)))
    return out


def sl_carry(v):
    h = [(x >> LB) if k < NL - 1 else 0 for k, x in enumerate(v)]
    return [wrap32(((x & MASK) if k < NL - 1 else (x if k == NL - 1 else 0)) + (h[k - 1] if k > 0 else 0))
            for k, x in enumerate(v)]


def sl_mul(a, b, mask_every_step=False, quotients=None):
    """Returns (result lanes, largest column magnitude seen); appends the quotients to `quotients`."""
    p = P_LIMBS + [0] * (16 - NL)
    t = [0] * 16
    peak = 0
    for i in range(NL):
        t = [x + a[i] * b[k] for k, x in enumerate(t)]
        m = ((t[0] & 0xFFFFFFFF) * NINV) & 0xFFFFFFFF
        if mask_every_step or i == NL - 1:
            m &= MASK
        if quotients is not None:
            quotients.append(m)
        t = [x + m * p[k] for k, x in enumerate(t)]
        peak = max(peak, max(abs(x) for x in t))
        assert all(wrap64(x) == x for x in t), "a column left the signed 64-bit range"
        assert t[0] & MASK == 0
        low = [x & MASK for x in t]
        t = [(x >> LB) + (low[k + 1] if k + 1 < 16 else 0) for k, x in enumerate(t)]
    return sl_carry(sl_carry64(t)), peak


def random_operand(rng, extreme):
    """Limbs 0..7 in [-2^6, 2^29 + 2^6], a signed top limb, |value| below 2^258."""
    limbs = []
    for _ in range(NL - 1):
        if extreme:
            limbs.append(rng.choice([-(1 << 6), 0, (1 << LB) + (1 << 6), MASK, rng.randrange(-(1 << 6), (1 << LB) + (1 << 6) + 1)]))
        else:
            limbs.append(rng.randrange(0, 1 << LB))
    top = rng.choice([(1 << 26) - 2, -((1 << 26) - 2), rng.randrange(-(1 << 26) + 2, (1 << 26) - 1)]) if extreme else rng.randrange(0, 1 << 23)
    limbs.append(top)
    return limbs + [0] * (16 - NL)


@pytest.mark.parametrize("extreme", [False, True])
def test_sliced_product_bounds_and_residue(extreme):
    rng = random.Random(0x51CE + extreme)
    worst = 0
    for _ in range(1500):
        a, b = random_operand(rng, extreme), random_operand(rng, extreme)
        va, vb = value(a), value(b)
        assert abs(va) < 1 << 258 and abs(vb) < 1 << 258
        r, peak = sl_mul(a, b)
        worst = max(worst, peak)
        assert peak < (1 << 33) + (1 << 58) * 1.08 + (1 << 61)
        vr = value(r)
        assert (vr - va * vb * R_INV) % P == 0
        bound = abs(va * vb) / (1 << (LB * NL)) + P * (1 + 2 ** -26)
        assert abs(vr) < bound
        # normalized: limbs 0..7 in [-2^6, 2^29 + 2^6], as the next product's inputs need
        assert all(-(1 << 6) <= x <= (1 << LB) + (1 << 6) for x in r[:NL - 1])
        assert all(x == 0 for x in r[NL:])
    assert worst < 1 << 62


def test_unmasked_quotients_reach_32_bits():
    """The bounds above are not vacuous: the unmasked quotients do reach past 2^31, the last one
    never passes 2^29, and the product agrees with the one that masks every quotient, modulo p."""
    rng = random.Random(7)
    seen = []
    for _ in range(200):
        a, b = random_operand(rng, True), random_operand(rng, True)
        qs = []
        r1, _ = sl_mul(a, b, quotients=qs)
        r2, _ = sl_mul(a, b, mask_every_step=True)
        assert (value(r1) - value(r2)) % P == 0
        assert qs[-1] <= MASK
        seen.extend(qs[:-1])
    assert max(seen) >= 1 << 31


def test_model_constants_match_the_field_header():
    src = open(__file__.replace("tests/test_sliced_product_model.py", "zk-gpu-accelerator_amd/csrc/zk_constants.hpp")).read()
    assert f"ninv = 0x{NINV:08x}u" in src
