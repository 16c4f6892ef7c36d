"""
Multi-scalar multiplication over BN254 G1 (oracle; test infrastructure only).

Restates src/zkgpu/msm.py:35-207: msm_naive (independent scalar
multiplications), msm_double_and_add (Straus/Shamir with one shared
accumulator), and msm_pippenger with the reference's window cost model
w = argmin_w ceil(254/w) (n + 2^(w+1)) over 1..16, unsigned w-bit digits,
2^w buckets filled by mixed additions, running sums for sum_b b*B_b, and
Horner's rule over the windows. The result is a unique group element, so
every correct MSM (including the GPU's signed-digit Pippenger) returns the
same affine Point.
"""

from __future__ import annotations

import operator

from zkref.curve import (
    JACOBIAN_IDENTITY,
    Jacobian,
    Point,
    from_jacobian,
    jacobian_add,
    jacobian_add_affine,
    jacobian_double,
)
from zkref.field import R_SCALAR

SCALAR_BITS = R_SCALAR.bit_length()  # 254
MAX_WINDOW_SIZE = 16


def msm_naive(points: list[Point], scalars: list[int]) -> Point:
    if len(points) != len(scalars):
        raise ValueError(f"Length mismatch: {len(points)} points vs {len(scalars)} scalars.")
    acc = Point.identity()
    for p, s in zip(points, scalars):
        acc = acc + p * s
    return acc


def msm_double_and_add(points: list[Point], scalars: list[int]) -> Point:
    if len(points) != len(scalars):
        raise ValueError(f"Length mismatch: {len(points)} points vs {len(scalars)} scalars.")
    terms = [((-p if s < 0 else p).affine(), abs(s)) for p, s in zip(points, scalars) if not p.is_identity()]
    bits = max((s.bit_length() for _, s in terms), default=0)
    acc = JACOBIAN_IDENTITY
    for bit in range(bits - 1, -1, -1):
        acc = jacobian_double(acc)
        for (x, y), s in terms:
            if (s >> bit) & 1:
                acc = jacobian_add_affine(acc, x, y)
    return from_jacobian(acc)


def pippenger_window_size(n: int) -> int:
    """msm.py:112-121."""
    n = operator.index(n)
    best, best_cost = 1, None
    for w in range(1, MAX_WINDOW_SIZE + 1):
        cost = -(-SCALAR_BITS // w) * (n + 2 ** (w + 1))
        if best_cost is None or cost < best_cost:
            best, best_cost = w, cost
    return best


def check_window_size(window_size: int) -> int:
    """msm.py:124-129."""
    w = operator.index(window_size)
    if not 1 <= w <= MAX_WINDOW_SIZE:
        raise ValueError(f"window_size must be in 1..{MAX_WINDOW_SIZE}, got {w}")
    return w


def window_sum(coords: list[tuple[int, int]], digits: list[int], buckets: int) -> Jacobian:
    """sum_i digit_i * P_i for one window through 2^w buckets and running sums (msm.py:138-153)."""
    b = [JACOBIAN_IDENTITY] * buckets
    for (x, y), d in zip(coords, digits):
        if d:
            b[d] = jacobian_add_affine(b[d], x, y)
    running = total = JACOBIAN_IDENTITY
    for k in range(buckets - 1, 0, -1):
        running = jacobian_add(running, b[k])
        total = jacobian_add(total, running)
    return total


def combine_windows(sums: list[Jacobian], w: int) -> Jacobian:
    """Horner from the top window down (msm.py:156-169)."""
    acc = sums[-1]
    for s in reversed(sums[:-1]):
        for _ in range(w):
            acc = jacobian_double(acc)
        acc = jacobian_add(acc, s)
    return acc


def pippenger_affine(coords: list[tuple[int, int]], scalars: list[int], w: int) -> Point:
    """msm.py:172-176, on validated coordinates and reduced scalars."""
    windows = -(-SCALAR_BITS // w)
    mask = (1 << w) - 1
    sums = [window_sum(coords, [(s >> (k * w)) & mask for s in scalars], 1 << w) for k in range(windows)]
    return from_jacobian(combine_windows(sums, w))


def msm_pippenger(points: list[Point], scalars: list[int], *, window_size: int | None = None) -> Point:
    """msm.py:179-207."""
    if len(points) != len(scalars):
        raise ValueError(f"Length mismatch: {len(points)} points vs {len(scalars)} scalars.")
    w = pippenger_window_size(len(points)) if window_size is None else check_window_size(window_size)
    coords, reduced = [], []
    for p, s in zip(points, scalars):
        if not p.is_identity():
            coords.append(p.affine())
            reduced.append(s % R_SCALAR)
    return pippenger_affine(coords, reduced, w) if coords else Point.identity()
