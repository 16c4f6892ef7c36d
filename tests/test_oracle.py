"""
The oracle (oracle/zkref) pinned against known answers and independent mathematics.

Known answers come from the reference's tests (gnark/arkworks/snarkjs/py_ecc
constants, test_field256_gpu.py:60-85, test_domain.py:15-54,
test_curve.py:26-29, test_ntt.py:135-145, test_msm.py:170-171); the rest are
checks that do not reuse the oracle's own formulas: brute-force DFTs, Horner
evaluation, schoolbook convolution, the affine group law against the Jacobian
formulas, closed forms of Z3, and MSMs against the closed form (sum s_i k_i) G.
"""

import random

import pytest

from zkref import curve as oc
from zkref import domain as od
from zkref import msm as om
from zkref import ntt as on
from zkref.curve import JACOBIAN_IDENTITY, G, P, Point, from_jacobian, jacobian_add, jacobian_add_affine, jacobian_double, jacobian_mul
from zkref.field import P_BASE, R_SCALAR as R, montgomery_constants, mont_rows, rows_to_ints

SNARKJS_W28 = 19103219067921713944291392827692070036145651957329286315305642004821462161904
ARKWORKS_W28 = 1748695177688661943023146337482803886740723238769601073607632802312037301404
PY_ECC_2G = (
    1368015179489954701390400359078579693043519447331113978918064868415326638035,
    9918110051302171585080402603319702774565515993150576347155970296011118125764,
)


# ---- field ----------------------------------------------------------------


def test_montgomery_constants_match_gnark():
    assert montgomery_constants(R)[2] == 0xC2E1F593EFFFFFFF
    assert montgomery_constants(P_BASE)[2] == 0x87D20782E4866389
    assert P == P_BASE


def test_montgomery_one_matches_arkworks():
    rows = mont_rows([1, 0], R)
    r_mod = 0x0E0A77C19A07DF2F666EA36F7879462E36FC76959F60CD29AC96341C4FFFFFFB
    assert rows[0].tolist() == [(r_mod >> (64 * j)) & (2**64 - 1) for j in range(4)]
    assert rows[1].tolist() == [0, 0, 0, 0]
    assert rows_to_ints(mont_rows([5, R - 1, R + 3], R), R) == [5, R - 1, 3]


# ---- NTT ------------------------------------------------------------------


def test_known_non_residues():
    assert on.quadratic_non_residue(998_244_353) == 3
    assert on.quadratic_non_residue(12_289) == 11
    assert on.quadratic_non_residue(R) == 5


@pytest.mark.parametrize("n", [3, 5, 41, 12_289, 998_244_353, 2**61 - 1, 2**63 - 25, R])
def test_primes(n):
    assert on.is_probable_prime(n)


@pytest.mark.parametrize("n", [0, 1, 4, 561, 3_215_031_751, (2**61 - 1) * (2**31 - 1)])
def test_composites(n):
    assert not on.is_probable_prime(n)


def test_root_matches_snarkjs_and_arkworks():
    assert on.get_ntt_root(2**28, R) == SNARKJS_W28
    assert on.get_ntt_root(2**28, R, non_residue=7) == ARKWORKS_W28
    for k in (1, 2, 10, 20, 22, 27):
        assert on.get_ntt_root(2**k, R) == pow(SNARKJS_W28, 2 ** (28 - k), R)
    with pytest.raises(ValueError, match="not supported"):
        on.get_ntt_root(2**29, R)
    with pytest.raises(ValueError, match="non-residue"):
        on.get_ntt_root(16, R, non_residue=3)


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 64])
def test_ntt_is_the_dft(n):
    rng = random.Random(n)
    a = [rng.randrange(R) for _ in range(n)]
    assert on.ntt(a, R) == on.dft(a, R)
    assert on.ntt(a, R, invert=True) == on.dft(a, R, invert=True)


def test_dft_by_horner():
    """dft itself against direct polynomial evaluation at the roots."""
    rng = random.Random(3)
    a = [rng.randrange(R) for _ in range(16)]
    w = on.get_ntt_root(16, R)
    horner = []
    for k in range(16):
        x, acc = pow(w, k, R), 0
        for c in reversed(a):
            acc = (acc * x + c) % R
        horner.append(acc)
    assert on.dft(a, R) == horner


def test_ntt_of_x_lists_the_domain():
    for n in (8, 256):
        x = [0, 1] + [0] * (n - 2)
        assert on.ntt(x, R) == od.Radix2Domain(n).elements()


def test_ntt_convolution_matches_schoolbook():
    n = 32
    rng = random.Random(5)
    a, b = [rng.randrange(R) for _ in range(n)], [rng.randrange(R) for _ in range(n)]
    prod = on.ntt([x * y % R for x, y in zip(on.ntt(a, R), on.ntt(b, R))], R, invert=True)
    expected = [0] * n
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            expected[(i + j) % n] = (expected[(i + j) % n] + x * y) % R
    assert prod == expected


def test_ntt_validation_messages():
    with pytest.raises(ValueError, match="power of 2"):
        on.ntt([1, 2, 3], R)
    with pytest.raises(ValueError, match="not supported"):
        on.ntt(list(range(32)), 17)
    for m in (-7, 1, 2, 15, 21):
        with pytest.raises(ValueError, match="odd prime"):
            on.ntt([1, 2, 3, 4], m)
    assert on.ntt([], R) == []
    assert on.ntt([R + 42], R) == [42]


# ---- domain ---------------------------------------------------------------


def horner(coeffs, x, m=R):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % m
    return acc


@pytest.mark.parametrize("offset", [None, 7, R - 3])
def test_coset_fft_by_horner(offset):
    d = od.Radix2Domain(8)
    g = d.coset_offset if offset is None else offset
    rng = random.Random(2)
    coeffs = [rng.randrange(R) for _ in range(6)]
    assert d.coset_fft(coeffs, offset) == [horner(coeffs, g * x % R) for x in d.elements()]
    assert d.coset_ifft(d.coset_fft(coeffs, offset), offset) == coeffs + [0, 0]
    assert d.coset_offset == 5


def test_coset_offset_fallback_and_full_group():
    d = od.Radix2Domain(8, modulus=41)
    assert d.coset_offset == 4
    assert d.coset_fft([1, 2, 3]) == [horner([1, 2, 3], 4 * x % 41, 41) for x in d.elements()]
    full = od.Radix2Domain(16, modulus=17)
    assert full.coset_offset is None
    with pytest.raises(ValueError, match="no coset"):
        full.coset_fft([1, 2, 3])
    with pytest.raises(ValueError, match="outside the domain"):
        od.Radix2Domain(8).coset_fft([1], offset=od.Radix2Domain(8).omega)
    with pytest.raises(ValueError, match="do not fit"):
        od.Radix2Domain(8).fft(list(range(9)))


def test_groth16_quotient_identity():
    """With C = A*B on H, h * Z_H = A*B - C as polynomials (test_domain.py:124-150)."""
    n = 16
    d = od.Radix2Domain(n)
    rng = random.Random(4)
    a = [rng.randrange(R) for _ in range(n)]
    b = [rng.randrange(R) for _ in range(n)]
    c = [x * y % R for x, y in zip(a, b)]  # evaluations on H
    h = od.groth16_quotient(a, b, c, 7)
    assert h[-1] == 0
    ca, cb, cc = d.ifft(a), d.ifft(b), d.ifft(c)
    ab = [0] * (2 * n - 1)
    for i, x in enumerate(ca):
        for j, y in enumerate(cb):
            ab[i + j] = (ab[i + j] + x * y) % R
    num = [(ab[i] - (cc[i] if i < n else 0)) % R for i in range(2 * n - 1)]
    hz = [0] * (2 * n)
    for i, x in enumerate(h):
        hz[i + n] = (hz[i + n] + x) % R
        hz[i] = (hz[i] - x) % R
    assert num + [0] == hz


# ---- curve ----------------------------------------------------------------


def test_known_points():
    assert G.double() == Point(*PY_ECC_2G)
    assert G * 2 == Point(*PY_ECC_2G)
    assert (G * R).is_identity()
    assert G * (R - 1) == -G
    with pytest.raises(ValueError, match="not on the BN254 curve"):
        Point(1, 3)


def scaled(pt, lam):
    x, y = pt.affine()
    return (x * lam * lam % P, y * pow(lam, 3, P) % P, lam)


@pytest.fixture(scope="module")
def pts():
    rng = random.Random(2024)
    return [G * rng.randrange(1, R) for _ in range(4)], [rng.randrange(2, P) for _ in range(4)]


def test_jacobian_formulas_match_affine_law_and_z3_closed_forms(pts):
    """
    The affine result fixes (X/Z^2, Y/Z^3); the closed forms Z3 = 2 Y1 Z1 (dbl-2009-l),
    Z3 = 2 Z1 Z2 (X2 Z1^2 - X1 Z2^2) (add-2007-bl) and Z3 = 2 Z1 (x2 Z1^2 - X1) (madd-2007-bl)
    fix Z, so together they pin the exact triples.
    """
    (a, b, c, _), (l1, l2, l3, _) = pts
    ja, jb = scaled(a, l1), scaled(b, l2)
    d = jacobian_double(ja)
    assert from_jacobian(d) == a.double() and d[2] == 2 * ja[1] * ja[2] % P
    s = jacobian_add(ja, jb)
    assert from_jacobian(s) == a + b
    assert s[2] == 2 * ja[2] * jb[2] * (jb[0] * ja[2] ** 2 - ja[0] * jb[2] ** 2) % P
    m = jacobian_add_affine(ja, *c.affine())
    assert from_jacobian(m) == a + c
    assert m[2] == 2 * ja[2] * (c.x * ja[2] ** 2 - ja[0]) % P


def test_jacobian_special_cases(pts):
    (a, _, _, _), (l1, l2, _, _) = pts
    assert from_jacobian(jacobian_add(scaled(a, l1), scaled(a, l2))) == a.double()
    assert jacobian_add(scaled(a, l1), scaled(a, l2)) == jacobian_double(scaled(a, l1))
    assert jacobian_add(scaled(a, l1), scaled(-a, l2)) == JACOBIAN_IDENTITY
    assert jacobian_add_affine(scaled(a, l1), *(-a).affine()) == JACOBIAN_IDENTITY
    assert jacobian_add(JACOBIAN_IDENTITY, scaled(a, l1)) == scaled(a, l1)
    assert jacobian_add_affine(JACOBIAN_IDENTITY, *a.affine()) == (*a.affine(), 1)
    assert jacobian_double((5, 0, 1)) == JACOBIAN_IDENTITY


def test_jacobian_mul_matches_affine_ladder(pts):
    (a, _, _, _), _ = pts

    def ladder(pt, k):
        if k < 0:
            pt, k = -pt, -k
        res = Point.identity()
        while k:
            if k & 1:
                res = res + pt
            pt = pt.double()
            k >>= 1
        return res

    rng = random.Random(5)
    for k in (0, 1, 2, 3, -1, -7, R - 1, R, rng.randrange(R)):
        assert from_jacobian(jacobian_mul(a, k)) == ladder(a, k)


# ---- MSM --------------------------------------------------------------------


def test_window_model_known_answers():
    assert [om.pippenger_window_size(n) for n in (0, 100, 1_000, 10_000, 100_000)] == [1, 5, 7, 10, 13]
    assert om.pippenger_window_size(2**16) == 12 and om.pippenger_window_size(2**20) == 16
    with pytest.raises(ValueError, match="window_size"):
        om.check_window_size(17)


@pytest.mark.parametrize("w", [None, 1, 3, 8, 16])
def test_pippenger_closed_form(w):
    rng = random.Random(29)
    ks = [rng.randrange(1, R) for _ in range(20)]
    scalars = [rng.randrange(R) for _ in range(20)]
    expected = G * (sum(k * s for k, s in zip(ks, scalars)) % R)
    assert om.msm_pippenger([G * k for k in ks], scalars, window_size=w) == expected


def test_msm_implementations_agree_and_edge_cases():
    rng = random.Random(17)
    points = [G * rng.randint(1, 2**32) for _ in range(12)]
    scalars = [rng.randint(0, 2**32) for _ in range(12)]
    assert om.msm_naive(points, scalars) == om.msm_double_and_add(points, scalars) == om.msm_pippenger(points, scalars)
    inf = Point.identity()
    assert om.msm_pippenger([], []) == inf
    assert om.msm_pippenger([G, -G], [7, 7]) == inf
    assert om.msm_pippenger([G, G], [5, 5]) == G * 10
    assert om.msm_pippenger([inf, G, inf], [5, 3, -2]) == G * 3
    for s in (-5, R, R + 7, 2**300 + 1):
        assert om.msm_pippenger([G * 3], [s]) == G * 3 * (s % R)


# ---- planted mistakes (in the oracle) are caught by the checks above ---------


def test_planted_wrong_root_is_caught(monkeypatch):
    """A transform built on omega^2 stays invertible but fails the DFT / NTT(X) checks."""
    real = on.get_ntt_root
    monkeypatch.setattr(on, "get_ntt_root", lambda n, m, nr=None: pow(real(n, m), 3, m))
    x = [0, 1] + [0] * 6
    assert on.ntt(on.ntt(x, R), R, invert=True) == x  # round trip alone would pass
    monkeypatch.undo()
    bad = [pow(real(8, R), 3 * k, R) for k in range(8)]
    assert bad != od.Radix2Domain(8).elements()


def test_planted_missing_z_factor_is_caught(pts):
    """dbl-2009-l with Z3 = Y1 Z1 (a dropped factor 2) keeps no affine point and fails the closed form."""
    (a, _, _, _), (l1, _, _, _) = pts
    x3, y3, z3 = jacobian_double(scaled(a, l1))
    wrong = (x3, y3, z3 * pow(2, -1, P) % P)
    with pytest.raises(ValueError):
        assert from_jacobian(wrong) == a.double()
