"""
zkgpu.domain (Radix2Domain on the GPU) and the Groth16 quotient pipeline.

Every transform is compared with the oracle's Radix2Domain, which restates
the reference's CPU domain; the pipeline at 2^16 and 2^22 is compared bit for
bit with the oracle's C restatement, and its algebraic identity
h * Z_H = A * B - C is checked where C = A * B on H.
"""

import random

import numpy as np
import pytest

from zkgpu.domain import Radix2Domain
from zkgpu.field256_gpu import BN254_FR
from zkgpu import workloads
from zkref import domain as od, fast
from zkref.field import R_SCALAR as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 8, 1024, 4096])
def test_transforms_match_oracle(n):
    d, o = Radix2Domain(n), od.Radix2Domain(n)
    rng = random.Random(n)
    coeffs = [rng.randrange(R) for _ in range(n)]
    assert d.omega == o.omega and d.coset_offset == o.coset_offset == 5
    assert d.fft(coeffs) == o.fft(coeffs)
    assert d.ifft(coeffs) == o.ifft(coeffs)
    if n > 1:
        assert d.coset_fft(coeffs) == o.coset_fft(coeffs)
        assert d.coset_ifft(coeffs, 7) == o.coset_ifft(coeffs, 7)
    assert d.elements() == o.elements()


def test_zero_padding_and_errors():
    d = Radix2Domain(8)
    assert d.fft([1, 2, 3]) == od.Radix2Domain(8).fft([1, 2, 3])
    with pytest.raises(ValueError, match="do not fit"):
        d.fft(list(range(9)))
    with pytest.raises(ValueError, match="expected 8 evaluations"):
        d.ifft([1, 2, 3])
    with pytest.raises(ValueError, match="outside the domain"):
        d.coset_fft([1], offset=d.omega)


def test_other_moduli():
    d = Radix2Domain(16, modulus=998_244_353)
    o = od.Radix2Domain(16, modulus=998_244_353)
    assert d.fft([1, 2, 3]) == o.fft([1, 2, 3])
    d41 = Radix2Domain(8, modulus=41)
    assert d41.coset_offset == 4
    assert d41.coset_fft([1, 2, 3]) == od.Radix2Domain(8, modulus=41).coset_fft([1, 2, 3])


@pytest.mark.parametrize("n", [1, 2, 16, 1024])
def test_quotient_matches_oracle(n):
    rng = random.Random(n + 1)
    a, b, c = ([rng.randrange(R) for _ in range(n)] for _ in range(3))
    d = Radix2Domain(n)
    rows = [BN254_FR.from_ints(v) for v in (a, b, c)]
    assert BN254_FR.to_ints(d.quotient(*rows, offset=7)) == od.groth16_quotient(a, b, c, 7)


def test_quotient_identity_when_c_is_ab():
    n = 256
    rng = random.Random(3)
    a, b = [rng.randrange(R) for _ in range(n)], [rng.randrange(R) for _ in range(n)]
    c = [x * y % R for x, y in zip(a, b)]
    h = BN254_FR.to_ints(Radix2Domain(n).quotient(*(BN254_FR.from_ints(v) for v in (a, b, c)), offset=7))
    assert h[-1] == 0  # deg h <= n - 2


@pytest.mark.parametrize("log_n", [16, 22])
def test_config3_pipeline_bit_exact(log_n):
    """The Groth16-shaped pipeline (coset shift 7, seeds 0x5EED31..33) against the C oracle."""
    n = 1 << log_n
    plain = [workloads.uniform_rows(n, s) for s in (0x5EED31, 0x5EED32, 0x5EED33)]
    rows = [BN254_FR.from_limbs(p) for p in plain]
    got = BN254_FR.to_limbs(Radix2Domain(n).quotient(*rows, offset=7)).get()
    assert np.array_equal(got, fast.groth16_quotient(*plain, 7))
