"""
zkgpu.domain: radix-2 evaluation domains with their transforms on the GPU.

The reference's Radix2Domain (src/zkgpu/domain.py:43-166) with the same
construction, validation, default coset shift and list-in / list-out
methods, but fft / ifft / coset_fft / coset_ifft run as fused GPU transforms
(the coset powers and n^-1 are tables applied inside the NTT's first and last
passes, cached per domain). The *_rows variants take and return (n, 4)
Montgomery rows that stay on the device, which is how a prover keeps its
polynomials in HBM; quotient() is the Groth16 step h = (A B - C) / Z_H on a
coset in one call (demo/prover_simulation.py:114-150).
"""

from __future__ import annotations

import operator
from collections.abc import Sequence
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from zkgpu import gpu_backend as _gb
from zkgpu.field import FIELD_MODULUS
from zkgpu.field256_gpu import BN254_FR, Field256, _row_count
from zkgpu.gpu_backend import DeviceArray
from zkgpu.ntt import get_ntt_root, quadratic_non_residue

_FIELDS: dict[int, Field256] = {FIELD_MODULUS: BN254_FR}


def _field(modulus: int) -> Field256:
    f = _FIELDS.get(modulus)
    if f is None:
        f = _FIELDS[modulus] = Field256(modulus)
    return f


@dataclass(frozen=True)
class Radix2Domain:
    """The subgroup of n-th roots of unity of F_modulus, n a power of two."""

    size: int
    modulus: int = FIELD_MODULUS
    omega: int = field(init=False)
    coset_offset: int | None = field(init=False)

    def __post_init__(self) -> None:
        size = operator.index(self.size)
        if size < 1 or size & (size - 1):
            raise ValueError(f"domain size must be a power of 2, got {size}")
        object.__setattr__(self, "size", size)
        object.__setattr__(self, "omega", get_ntt_root(size, self.modulus))
        start = quadratic_non_residue(self.modulus)
        offset = next((g for g in range(start, self.modulus) if pow(g, size, self.modulus) != 1), None)
        object.__setattr__(self, "coset_offset", offset)

    @classmethod
    def at_least(cls, n: int, modulus: int = FIELD_MODULUS) -> Radix2Domain:
        if n < 1:
            raise ValueError(f"a domain needs at least one point, got {n}")
        return cls(1 << (n - 1).bit_length(), modulus)

    # ---- elements ------------------------------------------------------------

    def element(self, i: int) -> int:
        return pow(self.omega, i, self.modulus)

    def elements(self) -> list[int]:
        """[w^0, ..., w^(n-1)]: the transform of the polynomial X, computed on the GPU."""
        x = [0] * self.size
        if self.size > 1:
            x[1] = 1
        else:
            x[0] = 1
        return self.fft(x) if self.size > 1 else [1]

    def vanishing_eval(self, x: int) -> int:
        return (pow(x, self.size, self.modulus) - 1) % self.modulus

    # ---- list in, list out (the reference's signatures) -----------------------

    def fft(self, coeffs: Sequence[int]) -> list[int]:
        f = _field(self.modulus)
        return f.to_ints(self.fft_rows(f.from_ints(self._padded(coeffs))))

    def ifft(self, evals: Sequence[int]) -> list[int]:
        if len(evals) != self.size:
            raise ValueError(f"expected {self.size} evaluations, got {len(evals)}")
        f = _field(self.modulus)
        return f.to_ints(self.ifft_rows(f.from_ints(evals)))

    def coset_fft(self, coeffs: Sequence[int], offset: int | None = None) -> list[int]:
        g = self._offset(offset)
        f = _field(self.modulus)
        return f.to_ints(f.coset_ntt(f.from_ints(self._padded(coeffs)), g))

    def coset_ifft(self, evals: Sequence[int], offset: int | None = None) -> list[int]:
        g = self._offset(offset)
        if len(evals) != self.size:
            raise ValueError(f"expected {self.size} evaluations, got {len(evals)}")
        f = _field(self.modulus)
        return f.to_ints(f.coset_ntt(f.from_ints(evals), g, invert=True))

    # ---- device rows in, device rows out -----------------------------------------

    def fft_rows(self, rows: Any) -> DeviceArray:
        self._check_rows(rows)
        return _field(self.modulus).ntt(rows)

    def ifft_rows(self, rows: Any) -> DeviceArray:
        self._check_rows(rows)
        return _field(self.modulus).ntt(rows, invert=True)

    def coset_fft_rows(self, rows: Any, offset: int | None = None) -> DeviceArray:
        self._check_rows(rows)
        return _field(self.modulus).coset_ntt(rows, self._offset(offset))

    def coset_ifft_rows(self, rows: Any, offset: int | None = None) -> DeviceArray:
        self._check_rows(rows)
        return _field(self.modulus).coset_ntt(rows, self._offset(offset), invert=True)

    def quotient(self, a: Any, b: Any, c: Any, offset: int | None = None) -> DeviceArray:
        """
        Coefficients of h = (A B - C) / Z_H from evaluations of A, B, C on H (Montgomery rows),
        through the coset g H: three inverse + coset NTTs, the pointwise quotient and one inverse
        coset NTT, on the device (BN254's r only).
        """
        if self.modulus != FIELD_MODULUS:
            raise ValueError("quotient runs over BN254's scalar field r only")
        g = self._offset(offset)
        for x in (a, b, c):
            self._check_rows(x)
        _gb.require_gpu("Radix2Domain.quotient")
        a, b, c = (_gb.asarray(x) for x in (a, b, c))
        if self.size == 1:
            # (a*b - c) / (g - 1) on the single point
            f = BN254_FR
            z = f.from_ints([pow((g - 1) % self.modulus, -1, self.modulus)])
            return f.mul(f.sub(f.mul(a, b), c), z)
        h = _gb.empty((self.size, 4))
        _gb.call("zk_groth16_quotient", a.ptr, b.ptr, c.ptr, h.ptr, self.size.bit_length() - 1, _gb.limbs_of(g), _gb.current_stream())
        return h

    # ---- validation ---------------------------------------------------------------

    def _check_rows(self, rows: Any) -> None:
        n = _row_count(rows, "rows")
        if n != self.size:
            raise ValueError(f"expected {self.size} rows, got {n}")

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


def _as_rows(x: Any) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.uint64)
