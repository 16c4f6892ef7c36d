"""
The MSM's scalar recoding on the CPU: zk_scalar.hpp (the digit kernel's code) is compiled for
the host and checked against Python integers — reduction mod r, the GLV split
k = k1 + k2 lambda with |k1|, |k2| < 2^127, and signed c-bit digits — and the GLV constants
are rederived independently: lambda and beta from the oracle's group law (beta x, y) =
lambda (x, y), the basis from its defining congruence.
"""

import ctypes
import random
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from zkref.curve import G, P as FQ
from zkref.field import R_SCALAR as R

ROOT = Path(__file__).resolve().parents[1]
CONSTANTS = ROOT / "zk-gpu-accelerator_amd" / "csrc" / "zk_constants.hpp"


def words(v, n=8):
    return (ctypes.c_uint32 * n)(*[(v >> (32 * i)) & 0xFFFFFFFF for i in range(n)])


def value(arr, n=8):
    return sum(int(arr[i]) << (32 * i) for i in range(n))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    out = tmp_path_factory.mktemp("scalar") / "libscalar.so"
    src = ROOT / "tests" / "native" / "scalar_recode.cpp"
    subprocess.run([gxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-Wno-unknown-pragmas", str(src), "-o", str(out)],
                   check=True)
    return ctypes.CDLL(str(out))


def glv_struct() -> dict[str, int]:
    text = CONSTANTS.read_text()
    body = text[text.index("struct Glv {"):]
    body = body[: body.index("\n};")]
    out = {}
    for name, vals in re.findall(r"uint32_t (\w+)\[\d+\] = \{([^}]*)\}", body):
        limbs = [int(v.strip().rstrip("u"), 16) for v in vals.split(",")]
        bits = 29 if name == "beta261" else 32
        out[name] = sum(v << (bits * i) for i, v in enumerate(limbs))
    lam = int(re.search(r"lambda = (0x[0-9a-f]+)", text).group(1), 16)
    out["lambda"] = lam
    out["beta"] = out["beta261"] * pow(2, -261, FQ) % FQ
    return out


def split(lib, k):
    m1, m2 = (ctypes.c_uint32 * 8)(), (ctypes.c_uint32 * 8)()
    n1, n2 = ctypes.c_int(), ctypes.c_int()
    lib.zkt_glv(words(k), m1, ctypes.byref(n1), m2, ctypes.byref(n2))
    k1 = -value(m1) if n1.value else value(m1)
    k2 = -value(m2) if n2.value else value(m2)
    return k1, k2


def test_constants_are_the_endomorphism_and_a_lattice_basis():
    c = glv_struct()
    lam, beta = c["lambda"], c["beta"]
    assert (lam * lam + lam + 1) % R == 0 and lam != 1
    assert (beta * beta + beta + 1) % FQ == 0
    gx, gy = G.affine()
    assert (G * lam).affine() == (beta * gx % FQ, gy)
    a1, b1, a2, b2 = c["a1"], -c["b1n"], c["a2"], c["b2"]
    assert (a1 + b1 * lam) % R == 0 and (a2 + b2 * lam) % R == 0
    assert a1 * b2 - a2 * b1 == R  # a basis of the lattice (determinant r)
    assert c["g1"] == (b2 * 2**256 + R // 2) // R
    assert c["g2"] == (-b1 * 2**256 + R // 2) // R


def test_reduce_mod_r(lib):
    rng = random.Random(7)
    cases = [0, 1, R - 1, R, R + 1, 2 * R, 5 * R + 3, 2**256 - 1] + [rng.randrange(2**256) for _ in range(2000)]
    out = (ctypes.c_uint32 * 8)()
    for k in cases:
        lib.zkt_reduce(words(k), out)
        assert value(out) == k % R, k


def test_glv_split_identity_and_bounds(lib):
    lam = glv_struct()["lambda"]
    rng = random.Random(8)
    cases = [0, 1, 2, R - 1, R - 2, 2**32 - 1, lam, R - lam, (R + 1) // 2, 2**127, 2**127 - 1, 2**253]
    cases += [rng.randrange(R) for _ in range(20000)] + [rng.randrange(2**64) for _ in range(1000)]
    for k in cases:
        k1, k2 = split(lib, k)
        assert (k1 + k2 * lam - k) % R == 0, k
        assert abs(k1) < 2**127 and abs(k2) < 2**127, k
    # small scalars (the witness mix's booleans and 32-bit values) do not touch phi(P)
    for k in [0, 1, 12345, 2**32 - 1]:
        assert split(lib, k) == (k, 0)


@pytest.mark.parametrize("c", [1, 2, 5, 9, 13, 16, 20])
def test_signed_digits_reconstruct(lib, c):
    rng = random.Random(c)
    out = (ctypes.c_int32 * 300)()
    for bits in (127, 254):
        windows = bits // c + 1
        for v in [0, 1, 2**bits - 1, 2 ** (bits - 1)] + [rng.randrange(2**bits) for _ in range(300)]:
            carry = lib.zkt_digits(words(v), c, windows, out)
            d = [out[i] for i in range(windows)]
            assert carry == 0
            assert all(-(1 << (c - 1)) <= x <= (1 << (c - 1)) for x in d)
            assert sum(x << (c * i) for i, x in enumerate(d)) == v


@pytest.mark.parametrize("c", [5, 13, 15, 16])
def test_folded_half_digits_fit_16_bit_codes(lib, c):
    """A GLV half's digits as stored for the sort (sign folded in, a negative half's ties negative
    below the top window): they rebuild the half and lie in [-(2^(c-1) - 1), 2^(c-1)], so that
    (|d| - 1) with a sign bit never sets all 16 bits of a code for c <= 16."""
    rng = random.Random(100 + c)
    out = (ctypes.c_int32 * 300)()
    windows = 127 // c + 1
    half = 1 << (c - 1)
    # the halves the split gives (|k_i| < 2^126.97), and the window patterns with ties everywhere
    ties = sum(half << (c * i) for i in range(windows - 1))
    vals = [0, 1, ties, ties - 1, ties + 1, (1 << 126) + ties]
    vals += [abs(x) for k in [rng.randrange(R) for _ in range(400)] for x in split(lib, k)]
    for v in vals:
        for neg in (0, 1):
            carry = lib.zkt_half_digits(words(v), c, windows, neg, out)
            d = [out[i] for i in range(windows)]
            assert carry == 0, (v, neg)
            assert sum(x << (c * i) for i, x in enumerate(d)) == (-v if neg else v)
            assert all(-(half - 1) <= x <= half for x in d), (v, neg, d)


def test_top_16_bit_window_of_a_half_stays_below_its_tie():
    """The split's bound 1.125 (|a1| + |a2|) puts a half's top 16-bit window (bits 112..) at most
    at 32055, so with its carry in it never reaches 2^15, the tie the 16-bit codes cannot fold."""
    c = glv_struct()
    for a, b in ((c["a1"], c["a2"]), (c["b1n"], c["b2"])):
        assert ((9 * (a + b)) // 8 >> 112) + 1 < 1 << 15
