"""
The C restatement of the oracle (oracle/c/zkref.c) against the Python oracle.

The C code serves the full-size GPU parity tests (2^20-point MSMs, 2^22-point
NTTs and the Groth16 pipeline), so it is itself checked here, on the CPU,
for every function those tests use.
"""

import random
import subprocess
from pathlib import Path

import numpy as np
import pytest

from zkref import fast
from zkref.curve import G, P, from_jacobian, jacobian_add, jacobian_add_affine, jacobian_double, jacobian_mul
from zkref.domain import groth16_quotient
from zkref.field import R_SCALAR as R, ints_to_limbs, limbs_to_ints
from zkref.msm import msm_pippenger
from zkref.ntt import ntt

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module", autouse=True)
def built():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


@pytest.mark.parametrize("k", [0, 1, 3, 6, 10])
def test_ntt(k):
    rng = random.Random(k)
    a = [rng.randrange(R) for _ in range(1 << k)]
    assert limbs_to_ints(fast.ntt_rows(ints_to_limbs(a))) == ntt(a, R)
    assert limbs_to_ints(fast.ntt_rows(ints_to_limbs(a), invert=True)) == ntt(a, R, invert=True)


def test_groth16_pipeline():
    rng = random.Random(8)
    n = 64
    a, b, c = ([rng.randrange(R) for _ in range(n)] for _ in range(3))
    got = fast.groth16_quotient(*(ints_to_limbs(v) for v in (a, b, c)), 7)
    assert limbs_to_ints(got) == groth16_quotient(a, b, c, 7)


def flat(js):
    return [c for j in js for c in j]


def test_point_formulas():
    rng = random.Random(9)
    js = [jacobian_mul(G * rng.randrange(1, R), rng.randrange(2, 2**64)) for _ in range(8)]
    js += [(1, 1, 0), js[0], (js[1][0], -js[1][1] % P, js[1][2])]
    other = js[::-1][:8] + [js[2], js[0], js[1]]
    a, b = ints_to_limbs(flat(js)), ints_to_limbs(flat(other))
    assert limbs_to_ints(fast.g1_op("add", a, b)) == flat([jacobian_add(x, y) for x, y in zip(js, other)])
    assert limbs_to_ints(fast.g1_op("double", a)) == flat([jacobian_double(x) for x in js])
    aff = [from_jacobian(j).affine() for j in other if j[2]]
    base = [j for j, o in zip(js, other) if o[2]]
    got = limbs_to_ints(fast.g1_op("add_affine", ints_to_limbs(flat(base)), ints_to_limbs(flat(aff))))
    assert got == flat([jacobian_add_affine(j, *q) for j, q in zip(base, aff)])


def test_fixed_base_and_msm():
    rng = random.Random(10)
    ks = [rng.randrange(1, R) for _ in range(40)]
    pts = fast.fixed_base_g1(ints_to_limbs(ks))
    pl = limbs_to_ints(pts.reshape(-1, 4))
    assert [(pl[2 * i], pl[2 * i + 1]) for i in range(40)] == [(G * k).affine() for k in ks]
    sc = [rng.randrange(R) for _ in range(40)]
    for w in (1, 4, 9):
        got = fast.msm_pippenger(pts, ints_to_limbs(sc), w)
        assert got == msm_pippenger([G * k for k in ks], sc, window_size=w).affine()
    assert got == (G * (sum(k * s for k, s in zip(ks, sc)) % R)).affine()
    assert fast.msm_pippenger(pts[:2], ints_to_limbs([0, 0]), 3) is None


def test_dot_mod_r_is_the_msm_multiple():
    """dot_mod_r is the scalar whose multiple of G equals the MSM over P_i = k_i G: pinned by the
    definition on Python ints (operands at and above r included, as msm_gpu.py:433 reduces
    scalars), by the oracle's Pippenger over those points, and by a planted swap of one pair."""
    rng = random.Random(11)
    ks = [rng.randrange(1, R) for _ in range(30)] + [R - 1, R, 2**256 - 1]
    sc = [rng.randrange(R) for _ in range(30)] + [2**256 - 1, 5, R + 3]
    got = fast.dot_mod_r(ints_to_limbs(ks), ints_to_limbs(sc))
    assert got == sum(k * s for k, s in zip(ks, sc)) % R
    pts = [G * (k % R) for k in ks[:30]]
    assert (G * fast.dot_mod_r(ints_to_limbs(ks[:30]), ints_to_limbs(sc[:30]))).affine() == \
        msm_pippenger(pts, sc[:30], window_size=5).affine()
    swapped = sc[:]
    swapped[0], swapped[1] = swapped[1], swapped[0]
    assert fast.dot_mod_r(ints_to_limbs(ks), ints_to_limbs(swapped)) != got
    assert fast.dot_mod_r(ints_to_limbs([]), ints_to_limbs([])) == 0
