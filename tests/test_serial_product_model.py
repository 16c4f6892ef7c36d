"""CPU model of the thread-serial Montgomery products (csrc/zk_field.hpp: mont_raw_serial,
mont_sqr_raw, mont_raw_sum2, mont_raw_ilp, and mont_raw_pair / mont_sqr_raw_pair, which run two
of them in lockstep with the same columns) with unmasked quotient digits, as the MSM kernels'
field tag FqMsm runs them (csrc/zk_msm.hip).

Product scanning: column k sums a_i b_(k-i) and m_i p_(k-i) into one 64-bit accumulator with the
carry of column k - 1, takes m_k = (t mod 2^32) (-p^-1 mod 2^29) mod 2^32 (only m_8 masked to 29
bits), adds m_k p_0 and shifts by 29. Every term is non-negative, so the partial sums the compiler
splits a column into, and mont_raw_ilp's separate a b and m p sums, are bounded by the column's
total. These tests check, for every operand form the MSM passes to a product:
- the worst case of every column, from the limb bounds and m_i = 2^32 - 1, stays below 2^64;
- replays at random and extreme operands give a b 2^-261 mod p, below a b / 2^261 + p (1 + 2^-26)
  and below 2p, with every column inside 64 bits;
- the quotients the replays take do pass 2^29, so the bound is not vacuous.
"""
import random
import zlib

import pytest

P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
LB = 29
MASK = (1 << LB) - 1
NL = 9
NINV = (-pow(P, -1, 1 << LB)) % (1 << LB)
R = 1 << (LB * NL)
R_INV = pow(R, -1, P)
P_LIMBS = [(P >> (LB * k)) & MASK for k in range(NL)]
U64 = 1 << 64


def value(limbs):
    return sum(l << (LB * k) for k, l in enumerate(limbs))


def limbs_of(v):
    """Normalized limbs; the top limb takes what is left."""
    out = [(v >> (LB * k)) & MASK for k in range(NL - 1)]
    out.append(v >> (LB * (NL - 1)))
    return out


def product(pairs, free_quotient=True, quotients=None):
    """sum over (a, b) in pairs of a b 2^-261 mod p with one reduction, column by column as the
    device does. Returns (result limbs, largest column total)."""
    m = []
    acc = 0
    peak = 0
    r = []
    for k in range(2 * NL - 1):
        lo = max(0, k - (NL - 1))
        hi = min(k, NL - 1)
        for a, b in pairs:
            acc += sum(a[i] * b[k - i] for i in range(lo, hi + 1))
        acc += sum(m[i] * P_LIMBS[k - i] for i in range(lo, min(k, NL)) if i < len(m))
        if k < NL:
            q = ((acc & 0xFFFFFFFF) * NINV) & 0xFFFFFFFF
            if not free_quotient or k == NL - 1:
                q &= MASK
            m.append(q)
            if quotients is not None:
                quotients.append(q)
            acc += q * P_LIMBS[0]
            assert acc & MASK == 0
            peak = max(peak, acc)
            assert acc < U64, f"column {k} left 64 bits"
            acc >>= LB
        else:
            peak = max(peak, acc)
            assert acc < U64, f"column {k} left 64 bits"
            r.append(acc & MASK)
            acc >>= LB
    r.append(acc & 0xFFFFFFFF)
    assert acc < 1 << 32
    return r, peak


# Operand forms: (per-limb maxima, value bound). Normalized values below 2p or 4p; a p4r
# difference (fsub_for_mul: limbs below 3 * 2^29, value below 6p); a sum of two normalized values
# (fe_add_nc: limbs below 2^30, value below 4p); an lc_carry result of a combination below 18p.
NORM2 = ([MASK] * (NL - 1) + [(2 * P) >> (LB * (NL - 1))], 2 * P)
NORM4 = ([MASK] * (NL - 1) + [(4 * P) >> (LB * (NL - 1))], 4 * P)
P4R = ([3 * (1 << LB) - 1] * (NL - 1) + [(6 * P) >> (LB * (NL - 1))], 6 * P)
WIDE = ([(1 << 30) - 2] * (NL - 1) + [(4 * P) >> (LB * (NL - 1))], 4 * P)
CARRY18 = ([MASK] * (NL - 1) + [(18 * P) >> (LB * (NL - 1))], 18 * P)

# The products the MSM's formulas take: each entry is a list of (a form, b form) pairs summed
# under one reduction.
CASES = {
    "norm x norm (coordinates, products)": [(NORM2, NORM2)],
    "norm 4p squared (R + 2p in madd, add)": [(NORM4, NORM4)],
    "p4r x norm 4p (R (Q - X3))": [(P4R, NORM4)],
    "sum2 (R (Q - X3) + (2p - Y1) PPP)": [(P4R, NORM4), (NORM4, NORM2)],
    "wide x wide ((2H)^2, (Z1 + Z2)^2, (X + B)^2)": [(WIDE, WIDE)],
    "wide x lc_carry (r (3V + J - r^2))": [(WIDE, CARRY18)],
    "norm x lc_carry (A w, Z3 = Z' H)": [(NORM2, CARRY18)],
}


def worst_columns(pairs):
    """Largest total each column can reach: every limb at its maximum, every free quotient at
    2^32 - 1, m_8 at 2^29 - 1, and the largest carry the previous column can pass."""
    carry = 0
    worst = 0
    for k in range(2 * NL - 1):
        lo = max(0, k - (NL - 1))
        hi = min(k, NL - 1)
        ab = sum(A[0][i] * B[0][k - i] for A, B in pairs for i in range(lo, hi + 1))
        mp = sum(((1 << 32) - 1 if i < NL - 1 else MASK) * P_LIMBS[k - i] for i in range(lo, min(k, NL - 1) + 1))
        total = ab + mp + carry
        worst = max(worst, total)
        carry = total >> LB
    return worst


@pytest.mark.parametrize("case", sorted(CASES))
def test_worst_column_fits_64_bits(case):
    worst = worst_columns(CASES[case])
    assert worst < U64, f"{case}: 2^{worst.bit_length()}"


def test_modulus_limbs_sum_as_the_device_check_assumes():
    # free_quotient_fits in zk_field.hpp: (sum p_j) 2^32 + 9 * 2^60 + 2^35 < 2^64
    assert (sum(P_LIMBS) << 32) + (9 << 60) + (1 << 35) < U64
    # and the a b part of every form above stays within 9 * 2^60 per column
    for pairs in CASES.values():
        for k in range(2 * NL - 1):
            lo, hi = max(0, k - (NL - 1)), min(k, NL - 1)
            assert sum(A[0][i] * B[0][k - i] for A, B in pairs for i in range(lo, hi + 1)) <= 9 << 60


def random_operand(rng, form, extreme):
    maxima, bound = form
    while True:
        if extreme:
            limbs = [rng.choice([mx, mx, mx - rng.randrange(1 << 8), rng.randrange(mx + 1)]) for mx in maxima[:-1]]
        else:
            limbs = [rng.randrange(mx + 1) for mx in maxima[:-1]]
        top_room = (bound - 1 - value(limbs + [0])) >> (LB * (NL - 1))
        if top_room < 0:
            continue
        top = rng.choice([top_room, top_room, rng.randrange(top_room + 1)]) if extreme else rng.randrange(top_room + 1)
        limbs.append(min(top, maxima[-1]))
        if value(limbs) < bound:
            return limbs


@pytest.mark.parametrize("extreme", [False, True])
@pytest.mark.parametrize("case", sorted(CASES))
def test_replay_residue_and_bound(case, extreme):
    rng = random.Random(zlib.crc32(case.encode()) * 2 + extreme)
    pairs = CASES[case]
    for _ in range(150):
        ops = [(random_operand(rng, A, extreme), random_operand(rng, B, extreme)) for A, B in pairs]
        r, peak = product(ops)
        exact = sum(value(a) * value(b) for a, b in ops)
        vr = value(r)
        assert (vr - exact * R_INV) % P == 0
        assert vr < exact / R + P * (1 + 2 ** -26)
        assert vr < 2 * P
        assert all(x <= MASK for x in r[:NL - 1])
        assert peak < U64


def test_free_quotients_pass_29_bits_and_agree_with_masked_ones():
    rng = random.Random(11)
    seen = []
    for _ in range(200):
        a = random_operand(rng, NORM2, True)
        b = random_operand(rng, NORM2, True)
        qs = []
        r1, _ = product([(a, b)], quotients=qs)
        r2, _ = product([(a, b)], free_quotient=False)
        assert (value(r1) - value(r2)) % P == 0
        assert qs[-1] <= MASK
        seen.extend(qs[:-1])
    assert max(seen) >= 1 << 31


def test_squaring_schedule_sums_the_same_columns():
    """mont_sqr_raw adds a_i (2 a_j) for i < j and a_i^2: per column the same total as the general
    product of a with itself, so the general replay bounds it."""
    rng = random.Random(5)
    for _ in range(50):
        a = random_operand(rng, WIDE, True)
        for k in range(2 * NL - 1):
            lo, hi = max(0, k - (NL - 1)), min(k, NL - 1)
            general = sum(a[i] * a[k - i] for i in range(lo, hi + 1))
            d = [x << 1 for x in a]
            sqr = sum(a[i] * d[k - i] for i in range(lo, hi + 1) if 2 * i < k) + (a[k // 2] ** 2 if k % 2 == 0 else 0)
            assert general == sqr


def test_model_constants_match_the_headers():
    root = __file__.replace("tests/test_serial_product_model.py", "zk-gpu-accelerator_amd/csrc/")
    assert f"ninv = 0x{NINV:08x}u" in open(root + "zk_constants.hpp").read()
    msm = open(root + "zk_msm.hip").read()
    assert "struct FqMsm : FqConst {\n  static constexpr bool kFreeQuotient = true;" in msm
    # the tag stays out of the generic field kernels (the NTT's wide operands need masked digits)
    for name in ("zk_ntt.hip", "zk_field_ops.hip", "zk_pipeline.hip", "zk_constants.hpp"):
        assert "kFreeQuotient" not in open(root + name).read()
