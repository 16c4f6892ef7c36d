"""
zkgpu.field256_gpu on the GPU against the oracle and Python integers.

Mirrors the reference's tests/test_field256_gpu.py (carry-edge operands, five
moduli, NTT sizes 1..4096, NTT(X) = domain elements, convolution, round trip,
broadcasting, empty batches, errors), with the oracle's ntt as the checker,
plus the full-size transforms bit for bit against the oracle's C restatement.
"""

import random

import numpy as np
import pytest

from zkgpu import field256_gpu, gpu_backend
from zkgpu.field256_gpu import BN254_FQ, BN254_FR, Field256
from zkref import fast
from zkref.curve import P
from zkref.domain import Radix2Domain
from zkref.field import R_SCALAR as R, ints_to_limbs, limbs_to_ints
from zkref.ntt import ntt

pytestmark = pytest.mark.gpu

FIELDS = [
    pytest.param(BN254_FR, id="bn254-r"),
    pytest.param(BN254_FQ, id="bn254-p"),
    pytest.param(Field256(2**255 - 19), id="2^255-19"),
    pytest.param(Field256(3), id="3"),
    pytest.param(Field256(2**64 + 13), id="2^64+13"),
]


def operands(p, seed):
    edge = sorted({e % p for e in (0, 1, 2, p - 1, p - 2, p // 2, 2**64 - 1, 2**128 - 1, 2**192 - 1)})
    rng = random.Random(seed)
    a = [x for x in edge for _ in edge] + [rng.randrange(p) for _ in range(500)]
    b = edge * len(edge) + [rng.randrange(p) for _ in range(500)]
    return a, b


def test_rows_hold_montgomery_form():
    rows = BN254_FR.from_ints([1, 0])
    assert rows.dtype == np.uint64 and rows.shape == (2, 4)
    r_mod = 0x0E0A77C19A07DF2F666EA36F7879462E36FC76959F60CD29AC96341C4FFFFFFB
    assert rows.tolist()[0] == [(r_mod >> (64 * j)) & (2**64 - 1) for j in range(4)]
    assert rows.tolist()[1] == [0, 0, 0, 0]


@pytest.mark.parametrize("field", FIELDS)
def test_round_trip(field):
    p = field.modulus
    values = [0, 1, p - 1, p, p + 1, -1, -p, 3 * p + 5, 2**256 + 9]
    assert field.to_ints(field.from_ints(values)) == [v % p for v in values]


@pytest.mark.parametrize("field", FIELDS)
@pytest.mark.parametrize("op", ["mul", "add", "sub"])
def test_ops_match_python(field, op):
    p = field.modulus
    a, b = operands(p, {"mul": 1, "add": 2, "sub": 3}[op])
    got = field.to_ints(getattr(field, op)(field.from_ints(a), field.from_ints(b)))
    py = {"mul": lambda x, y: x * y, "add": lambda x, y: x + y, "sub": lambda x, y: x - y}[op]
    assert got == [py(x, y) % p for x, y in zip(a, b)]


def test_plain_limbs_reduce_on_the_device():
    vals = [0, 1, R - 1, R, R + 5, 2**256 - 1, 5 * R + 3]
    assert BN254_FR.to_ints(BN254_FR.from_limbs(ints_to_limbs(vals))) == [v % R for v in vals]


def test_operations_compose():
    F, p = BN254_FQ, P
    a, b = operands(p, 4)
    A, B = F.from_ints(a), F.from_ints(b)
    lhs = F.to_ints(F.mul(F.add(A, B), F.sub(A, B)))
    assert lhs == F.to_ints(F.sub(F.mul(A, A), F.mul(B, B))) == [(x * x - y * y) % p for x, y in zip(a, b)]


@pytest.mark.parametrize("n", [1, 2, 4, 8, 64, 1024, 2048, 4096])
def test_ntt_matches_oracle(n):
    rng = random.Random(n)
    a = [rng.randrange(R) for _ in range(n)]
    rows = BN254_FR.from_ints(a)
    assert BN254_FR.to_ints(BN254_FR.ntt(rows)) == ntt(a, R)
    assert BN254_FR.to_ints(BN254_FR.ntt(rows, invert=True)) == ntt(a, R, invert=True)


@pytest.mark.parametrize("n", [8, 256, 1 << 11, 1 << 12])
def test_ntt_of_x_lists_the_domain(n):
    x = [0, 1] + [0] * (n - 2)
    assert BN254_FR.to_ints(BN254_FR.ntt(BN254_FR.from_ints(x))) == Radix2Domain(n).elements()


def test_ntt_round_trip_leaves_input_unchanged():
    rng = random.Random(9)
    a = [rng.randrange(R) for _ in range(512)]
    rows = BN254_FR.from_ints(a)
    before = rows.tolist()
    assert BN254_FR.to_ints(BN254_FR.ntt(BN254_FR.ntt(rows), invert=True)) == a
    assert rows.tolist() == before


def test_ntt_multiplies_polynomials():
    F, n = BN254_FR, 64
    rng = random.Random(5)
    a, b = [rng.randrange(R) for _ in range(n)], [rng.randrange(R) for _ in range(n)]
    product = F.ntt(F.mul(F.ntt(F.from_ints(a)), F.ntt(F.from_ints(b))), invert=True)
    expected = [0] * n
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            expected[(i + j) % n] = (expected[(i + j) % n] + x * y) % R
    assert F.to_ints(product) == expected


@pytest.mark.parametrize("offset", [5, 7, R - 3])
@pytest.mark.parametrize("n", [8, 1 << 12])
def test_coset_ntt_matches_domain(n, offset):
    d = Radix2Domain(n)
    rng = random.Random(n + offset)
    coeffs = [rng.randrange(R) for _ in range(n)]
    rows = BN254_FR.from_ints(coeffs)
    evals = BN254_FR.coset_ntt(rows, offset)
    assert BN254_FR.to_ints(evals) == d.coset_fft(coeffs, offset)
    assert BN254_FR.to_ints(BN254_FR.coset_ntt(evals, offset, invert=True)) == coeffs


def test_ntt_of_nothing():
    assert BN254_FR.ntt(BN254_FR.from_ints([])).shape == (0, 4)


@pytest.mark.parametrize(
    ("field", "n", "message"),
    [(BN254_FR, 6, "power of 2"), (BN254_FQ, 4, "not divisible"), (Field256(2**255 - 1), 2, "odd prime")],
)
def test_ntt_rejects_unsupported_sizes(field, n, message):
    with pytest.raises(ValueError, match=message):
        field.ntt(field.from_ints(range(n)))


def test_single_row_broadcasts_on_either_side():
    F, p = BN254_FR, R
    values = [3, 5, p - 1, 0]
    rows, c = F.from_ints(values), F.from_ints([p - 2])
    assert F.to_ints(F.mul(rows, c)) == [v * (p - 2) % p for v in values]
    assert F.to_ints(F.sub(c, rows)) == [(p - 2 - v) % p for v in values]
    assert F.to_ints(F.add(c, c)) == [(2 * (p - 2)) % p]


def test_empty_batches_and_mismatch():
    F = BN254_FR
    empty = F.from_ints([])
    assert empty.shape == (0, 4)
    assert F.to_ints(F.mul(empty, empty)) == []
    assert F.to_ints(F.add(empty, F.from_ints([7]))) == []
    with pytest.raises(ValueError, match="broadcast"):
        F.mul(F.from_ints([1, 2]), F.from_ints([1, 2, 3]))


def test_failed_self_test_keeps_nothing(monkeypatch):
    """A kernel that returns wrong values (as a miscompile would) fails the load-time self-test."""
    field = Field256(R)
    real = field256_gpu._limbs_to_ints
    monkeypatch.setattr(field256_gpu, "_limbs_to_ints", lambda x: [v ^ 1 for v in real(x)])
    with pytest.raises(RuntimeError, match="failed its self-test"):
        field.compile()
    assert field._kernels is None


def test_ntt_tables_cache_clear_and_layout():
    n = 16
    tw, n_inv = field256_gpu._ntt_tables(BN254_FR, n, False)
    w = Radix2Domain(n).omega
    assert BN254_FR.to_ints(tw) == [pow(w, k, R) for k in range(n // 2)]
    assert BN254_FR.to_ints(n_inv) == [pow(n, -1, R)]
    field256_gpu._ntt_tables.cache_clear()
    rng = random.Random(2)
    a = [rng.randrange(R) for _ in range(n)]
    assert BN254_FR.to_ints(BN254_FR.ntt(BN254_FR.from_ints(a))) == ntt(a, R)


def random_plain(n, seed):
    x = np.random.default_rng(seed).integers(0, 2**63, size=(n, 4), dtype=np.uint64)
    x[:, 3] &= np.uint64((1 << 60) - 1)  # below 2^252 < r
    return x


@pytest.mark.parametrize("log_n", [11, 13, 16, 20, 22])
def test_large_ntt_bit_exact(log_n):
    plain = random_plain(1 << log_n, log_n)
    rows = BN254_FR.from_limbs(plain)
    got = BN254_FR.to_limbs(BN254_FR.ntt(rows)).get()
    assert np.array_equal(got, fast.ntt_rows(plain))
    back = BN254_FR.to_limbs(BN254_FR.ntt(BN254_FR.ntt(rows), invert=True)).get()
    assert np.array_equal(back, plain)


def test_config1_ntt_2_12():
    """The reference's CPU-runnable case: Radix2Domain.fft at n = 2^12 on seeded coefficients."""
    rng = random.Random(0x5EED01)
    coeffs = [rng.randrange(R) for _ in range(1 << 12)]
    assert BN254_FR.to_ints(BN254_FR.ntt(BN254_FR.from_ints(coeffs))) == Radix2Domain(1 << 12).fft(coeffs)


def test_device_arrays_are_views():
    rows = BN254_FR.from_ints(list(range(10)))
    assert BN254_FR.to_ints(rows[2:5]) == [2, 3, 4]
    assert rows.reshape(5, 2, 4).shape == (5, 2, 4)
    assert BN254_FR.to_ints(rows.copy()) == list(range(10))
    assert gpu_backend.asarray(rows.get()).tolist() == rows.tolist()


def test_throughput_microbenchmarks_report_plausible_rates():
    """zk_valu_peak and zk_product_peak (the bench line's VALU and product ceilings) run and report
    rates an MI355X can reach: at most one v_mad_u64_u32 per lane per clock, and a 9-limb product
    costs 162 of them, so products/s stay below the multiply-add rate / 162."""
    import ctypes

    mads, prods = ctypes.c_double(0), ctypes.c_double(0)
    gpu_backend.call("zk_valu_peak", ctypes.byref(mads))
    gpu_backend.call("zk_product_peak", ctypes.byref(prods))
    assert 1.0 < mads.value < 60.0  # 10^12 multiply-adds per second
    assert 10.0 < prods.value < mads.value * 1e3 / 162  # 10^9 products per second
