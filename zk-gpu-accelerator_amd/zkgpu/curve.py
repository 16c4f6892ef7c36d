"""
BN254 G1 host types: the affine Point the MSM returns, and Jacobian triples.

Keeps the reference's value type and conversions (src/zkgpu/curve.py:29-233):
Point validates y^2 = x^3 + 3 and canonicalises its coordinates on
construction (an off-curve input raises ValueError, which blocks
invalid-curve attacks), the identity is Point(None, None), Jacobian
(X, Y, Z) stands for (X/Z^2, Y/Z^3) with the identity (1, 1, 0), and
from_jacobian converts with one inverse.

Group operations on Points (+, double, scalar *) run on the GPU through
zkgpu.curve_gpu and zkgpu.msm_gpu; this module computes no point formulas
on the CPU.
"""

from __future__ import annotations

import operator
from dataclasses import dataclass

P: int = 21888242871839275222246405745257275088696311157297823662689037894645226208583
A: int = 0
B: int = 3
GX: int = 1
GY: int = 2

Jacobian = tuple[int, int, int]
JACOBIAN_IDENTITY: Jacobian = (1, 1, 0)


def _on_curve(x: int, y: int) -> bool:
    return (y * y - pow(x, 3, P) - B) % P == 0


@dataclass(frozen=True, slots=True)
class Point:
    """An affine BN254 G1 point; (None, None) is the point at infinity."""

    x: int | None
    y: int | None

    def __post_init__(self) -> None:
        if self.x is None and self.y is None:
            return
        if self.x is None or self.y is None:
            raise ValueError(f"only the identity has None coordinates, got ({self.x}, {self.y})")
        object.__setattr__(self, "x", operator.index(self.x) % P)
        object.__setattr__(self, "y", operator.index(self.y) % P)
        if not _on_curve(self.x, self.y):
            raise ValueError(f"({self.x}, {self.y}) is not on the BN254 curve y^2 = x^3 + 3")

    @staticmethod
    def identity() -> Point:
        return Point(None, None)

    def is_identity(self) -> bool:
        return self.x is None and self.y is None

    def affine(self) -> tuple[int, int]:
        if self.x is None or self.y is None:
            raise ValueError("the point at infinity has no affine coordinates")
        return self.x, self.y

    def is_on_curve(self) -> bool:
        return self.is_identity() or _on_curve(*self.affine())

    def __neg__(self) -> Point:
        if self.is_identity():
            return self
        x, y = self.affine()
        return Point(x, -y % P)

    def __add__(self, other: object) -> Point:
        """Point addition on the GPU (one mixed addition)."""
        if not isinstance(other, Point):
            return NotImplemented
        if self.is_identity():
            return other
        if other.is_identity():
            return self
        from zkgpu import curve_gpu

        rows = curve_gpu.add_affine(curve_gpu.to_device([self]), curve_gpu.affine_to_device([other]))
        return curve_gpu.from_device(rows)[0]

    def double(self) -> Point:
        from zkgpu import curve_gpu

        return curve_gpu.from_device(curve_gpu.double(curve_gpu.to_device([self])))[0]

    def __mul__(self, scalar: int) -> Point:
        """Scalar multiplication on the GPU (a one-point MSM; scalars act mod r)."""
        try:
            k = operator.index(scalar)
        except TypeError:
            return NotImplemented
        if self.is_identity():
            return self
        from zkgpu.msm_gpu import msm_pippenger_gpu

        return msm_pippenger_gpu([self.x], [self.y], [k])

    def __rmul__(self, scalar: int) -> Point:
        return self.__mul__(scalar)

    def __repr__(self) -> str:
        return "Point(∞)" if self.is_identity() else f"Point(x={self.x}, y={self.y})"


G: Point = Point(GX, GY)


def to_jacobian(point: Point) -> Jacobian:
    if point.is_identity():
        return JACOBIAN_IDENTITY
    x, y = point.affine()
    return (x, y, 1)


def from_jacobian(j: Jacobian) -> Point:
    """Affine Point of (X, Y, Z) with a single inverse; validated by Point."""
    x, y, z = j
    if z % P == 0:
        return Point.identity()
    zi = pow(z, -1, P)
    zi2 = zi * zi % P
    return Point(x * zi2 % P, y * zi2 * zi % P)
