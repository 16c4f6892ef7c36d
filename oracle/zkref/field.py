"""
BN254 moduli and the (n, 4) uint64 Montgomery row format (oracle; test infrastructure only).

Follows src/zkgpu/field.py:17 (the scalar modulus r), src/zkgpu/curve.py:29
(the base modulus p) and src/zkgpu/field256_gpu.py:239-247,277-319 (limb
packing and Montgomery constants: a row holds a*2^256 mod m as four
little-endian 64-bit limbs).
"""

from __future__ import annotations

from collections.abc import Iterable

import numpy as np

R_SCALAR = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P_BASE = 21888242871839275222246405745257275088696311157297823662689037894645226208583
LIMBS = 4
R_BITS = 256
WORD = 1 << 64


def montgomery_constants(m: int) -> tuple[int, int, int]:
    """(R^2 mod m, R^-1 mod m, -m^-1 mod 2^64) for R = 2^256 (field256_gpu.py:285-287)."""
    return pow(2, 2 * R_BITS, m), pow(1 << R_BITS, -1, m), (-pow(m, -1, WORD)) % WORD


def to_mont(v: int, m: int) -> int:
    return (v % m) * (1 << R_BITS) % m


def from_mont(v: int, m: int) -> int:
    return v * pow(1 << R_BITS, -1, m) % m


def ints_to_limbs(values: Iterable[int]) -> np.ndarray:
    """(n, 4) little-endian uint64 limbs of values in [0, 2^256)."""
    vals = list(values)
    out = np.zeros((len(vals), LIMBS), dtype=np.uint64)
    for i, v in enumerate(vals):
        for j in range(LIMBS):
            out[i, j] = (v >> (64 * j)) & (WORD - 1)
    return out


def limbs_to_ints(limbs: np.ndarray) -> list[int]:
    arr = np.ascontiguousarray(limbs, dtype=np.uint64).reshape(-1, LIMBS)
    return [int(r[0]) | int(r[1]) << 64 | int(r[2]) << 128 | int(r[3]) << 192 for r in arr]


def mont_rows(values: Iterable[int], m: int) -> np.ndarray:
    """Montgomery-form rows of arbitrary ints, as Field256.from_ints produces them."""
    return ints_to_limbs(to_mont(v, m) for v in values)


def rows_to_ints(rows: np.ndarray, m: int) -> list[int]:
    """Plain values of Montgomery-form rows, as Field256.to_ints returns them."""
    r_inv = pow(1 << R_BITS, -1, m)
    return [v * r_inv % m for v in limbs_to_ints(rows)]
