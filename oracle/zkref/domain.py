"""
Radix-2 evaluation domain over a prime field (oracle; test infrastructure only).

Restates src/zkgpu/domain.py:33-166: H = <omega> of size n with omega from
get_ntt_root; the default coset shift is the smallest quadratic non-residue,
or the next candidate outside H when that one lies in H (domain.py:63-83);
fft zero-pads and rejects more than n coefficients; coset_fft scales
coefficient j by g^j before the fft; coset_ifft runs the ifft and then scales
by g^-j; vanishing_eval is x^n - 1. Error messages match the reference's.
"""

from __future__ import annotations

import operator
from collections.abc import Sequence
from dataclasses import dataclass, field

from zkref.field import R_SCALAR
from zkref.ntt import get_ntt_root, ntt, quadratic_non_residue


def scale_by_powers(values: Sequence[int], g: int, m: int) -> list[int]:
    """[v_j * g^j mod m] (domain.py:33-40)."""
    out, p = [], 1
    for v in values:
        out.append(v * p % m)
        p = p * g % m
    return out


@dataclass(frozen=True)
class Radix2Domain:
    size: int
    modulus: int = R_SCALAR
    omega: int = field(init=False)
    coset_offset: int | None = field(init=False)

    def __post_init__(self) -> None:
        n = operator.index(self.size)
        if n < 1 or n & (n - 1):
            raise ValueError(f"domain size must be a power of 2, got {n}")
        object.__setattr__(self, "size", n)
        object.__setattr__(self, "omega", get_ntt_root(n, self.modulus))
        g = quadratic_non_residue(self.modulus)
        offset = None
        while g < self.modulus:
            if pow(g, n, self.modulus) != 1:
                offset = g
                break
            g += 1
        object.__setattr__(self, "coset_offset", offset)

    @classmethod
    def at_least(cls, n: int, modulus: int = R_SCALAR) -> Radix2Domain:
        if n < 1:
            raise ValueError(f"a domain needs at least one point, got {n}")
        return cls(1 << (n - 1).bit_length(), modulus)

    def element(self, i: int) -> int:
        return pow(self.omega, i, self.modulus)

    def elements(self) -> list[int]:
        return scale_by_powers([1] * self.size, self.omega, self.modulus)

    def vanishing_eval(self, x: int) -> int:
        return (pow(x, self.size, self.modulus) - 1) % self.modulus

    def fft(self, coeffs: Sequence[int]) -> list[int]:
        return ntt(self._padded(coeffs), self.modulus)

    def ifft(self, evals: Sequence[int]) -> list[int]:
        if len(evals) != self.size:
            raise ValueError(f"expected {self.size} evaluations, got {len(evals)}")
        return ntt(list(evals), self.modulus, invert=True)

    def coset_fft(self, coeffs: Sequence[int], offset: int | None = None) -> list[int]:
        g = self._offset(offset)
        return self.fft(scale_by_powers(self._padded(coeffs), g, self.modulus))

    def coset_ifft(self, evals: Sequence[int], offset: int | None = None) -> list[int]:
        g = self._offset(offset)
        return scale_by_powers(self.ifft(evals), pow(g, -1, self.modulus), self.modulus)

    def _padded(self, coeffs: Sequence[int]) -> list[int]:
        if len(coeffs) > self.size:
            raise ValueError(
                f"{len(coeffs)} coefficients do not fit a domain of size {self.size}; "
                f"use Radix2Domain.at_least({len(coeffs)})"
            )
        return list(coeffs) + [0] * (self.size - len(coeffs))

    def _offset(self, offset: int | None) -> int:
        if offset is not None:
            g = operator.index(offset) % self.modulus
            if g == 0 or pow(g, self.size, self.modulus) == 1:
                raise ValueError(f"coset offset {offset} must be non-zero and outside the domain")
            return g
        if self.coset_offset is None:
            raise ValueError(
                f"the domain of size {self.size} is all of F_{self.modulus}*, so no coset outside it exists"
            )
        return self.coset_offset


def groth16_quotient(a_evals: Sequence[int], b_evals: Sequence[int], c_evals: Sequence[int], shift: int) -> list[int]:
    """
    h = coset_ifft((coset_fft(ifft A) * coset_fft(ifft B) - coset_fft(ifft C)) * (g^n - 1)^-1).

    The quotient step of demo/prover_simulation.py:114-120 with an explicit
    coset shift g (the pipeline workload uses g = 7).
    """
    n = len(a_evals)
    d = Radix2Domain(n)
    m = d.modulus
    z_inv = pow(d.vanishing_eval(shift), -1, m)
    a, b, c = (d.coset_fft(d.ifft(v), shift) for v in (a_evals, b_evals, c_evals))
    return d.coset_ifft([(x * y - z) * z_inv % m for x, y, z in zip(a, b, c)], shift)
