"""
BN254 G1: affine group law and Jacobian formulas (oracle; test infrastructure only).

Restates src/zkgpu/curve.py:29-324. y^2 = x^3 + 3 over F_p, G = (1, 2).
Affine points are validated and canonicalised on construction (curve.py:69-79);
the identity is (None, None). Jacobian (X, Y, Z) means (X/Z^2, Y/Z^3) and the
identity is (1, 1, 0). The formulas are the Explicit-Formulas Database's
a = 0 ones the reference names: dbl-2009-l for doubling, add-2007-bl for
addition and madd-2007-bl for mixed addition, each with the reference's
special cases (Z = 0 returns the other operand, H = 0 doubles or gives the
identity, Y = 0 in a doubling gives the identity). Every coordinate is reduced
into [0, p), so the (X, Y, Z) these produce are exactly the reference's.
"""

from __future__ import annotations

import operator
from dataclasses import dataclass

P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
B = 3

Jacobian = tuple[int, int, int]
JACOBIAN_IDENTITY: Jacobian = (1, 1, 0)


@dataclass(frozen=True, slots=True)
class Point:
    x: int | None
    y: int | None

    def __post_init__(self) -> None:
        if self.x is None and self.y is None:
            return
        if self.x is None or self.y is None:
            raise ValueError(f"only the identity has None coordinates, got ({self.x}, {self.y})")
        object.__setattr__(self, "x", operator.index(self.x) % P)
        object.__setattr__(self, "y", operator.index(self.y) % P)
        if (self.y * self.y - self.x**3 - B) % P:
            raise ValueError(f"({self.x}, {self.y}) is not on the BN254 curve y^2 = x^3 + 3")

    @staticmethod
    def identity() -> Point:
        return Point(None, None)

    def is_identity(self) -> bool:
        return self.x is None

    def affine(self) -> tuple[int, int]:
        if self.x is None or self.y is None:
            raise ValueError("the point at infinity has no affine coordinates")
        return self.x, self.y

    def __add__(self, other: Point) -> Point:
        """Chord-and-tangent addition with one inverse (curve.py:116-137)."""
        if not isinstance(other, Point):
            return NotImplemented
        if self.is_identity():
            return other
        if other.is_identity():
            return self
        (x1, y1), (x2, y2) = self.affine(), other.affine()
        if x1 == x2:
            return self.double() if y1 == y2 else Point.identity()
        lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
        x3 = (lam * lam - x1 - x2) % P
        return Point(x3, (lam * (x1 - x3) - y1) % P)

    def double(self) -> Point:
        if self.is_identity():
            return self
        x, y = self.affine()
        if y == 0:
            return Point.identity()
        lam = 3 * x * x * pow(2 * y, -1, P) % P
        x3 = (lam * lam - 2 * x) % P
        return Point(x3, (lam * (x - x3) - y) % P)

    def __neg__(self) -> Point:
        return self if self.is_identity() else Point(self.x, -self.y % P)

    def __mul__(self, k: int) -> Point:
        try:
            k = operator.index(k)
        except TypeError:
            return NotImplemented
        return from_jacobian(jacobian_mul(self, k))

    __rmul__ = __mul__


G = Point(1, 2)


def to_jacobian(pt: Point) -> Jacobian:
    return JACOBIAN_IDENTITY if pt.is_identity() else (pt.x, pt.y, 1)  # type: ignore[return-value]


def from_jacobian(j: Jacobian) -> Point:
    """(X/Z^2, Y/Z^3) through Point's validation (curve.py:221-233)."""
    x, y, z = j
    if z == 0:
        return Point.identity()
    zi = pow(z, -1, P)
    zi2 = zi * zi % P
    return Point(x * zi2 % P, y * zi2 * zi % P)


def jacobian_double(j: Jacobian) -> Jacobian:
    """EFD dbl-2009-l (curve.py:236-252)."""
    x1, y1, z1 = j
    if z1 == 0 or y1 == 0:
        return JACOBIAN_IDENTITY
    a = x1 * x1 % P
    b = y1 * y1 % P
    c = b * b % P
    t = (x1 + b) % P
    d = 2 * (t * t - a - c) % P
    e = 3 * a % P
    f = e * e % P
    x3 = (f - 2 * d) % P
    y3 = (e * (d - x3) - 8 * c) % P
    z3 = 2 * y1 * z1 % P
    return (x3, y3, z3)


def jacobian_add(j1: Jacobian, j2: Jacobian) -> Jacobian:
    """EFD add-2007-bl (curve.py:255-281)."""
    x1, y1, z1 = j1
    x2, y2, z2 = j2
    if z1 == 0:
        return j2
    if z2 == 0:
        return j1
    z1z1, z2z2 = z1 * z1 % P, z2 * z2 % P
    u1, u2 = x1 * z2z2 % P, x2 * z1z1 % P
    s1 = y1 * z2 * z2z2 % P
    s2 = y2 * z1 * z1z1 % P
    h = (u2 - u1) % P
    rr = 2 * (s2 - s1) % P
    if h == 0:
        return jacobian_double(j1) if rr == 0 else JACOBIAN_IDENTITY
    i = (2 * h) ** 2 % P
    jj = h * i % P
    v = u1 * i % P
    x3 = (rr * rr - jj - 2 * v) % P
    y3 = (rr * (v - x3) - 2 * s1 * jj) % P
    z3 = ((z1 + z2) ** 2 - z1z1 - z2z2) * h % P
    return (x3, y3, z3)


def jacobian_add_affine(j1: Jacobian, x2: int, y2: int) -> Jacobian:
    """EFD madd-2007-bl for a finite affine second operand (curve.py:284-303)."""
    x1, y1, z1 = j1
    if z1 == 0:
        return (x2, y2, 1)
    z1z1 = z1 * z1 % P
    u2 = x2 * z1z1 % P
    s2 = y2 * z1 * z1z1 % P
    h = (u2 - x1) % P
    rr = 2 * (s2 - y1) % P
    if h == 0:
        return jacobian_double(j1) if rr == 0 else JACOBIAN_IDENTITY
    hh = h * h % P
    i = 4 * hh % P
    jj = h * i % P
    v = x1 * i % P
    x3 = (rr * rr - jj - 2 * v) % P
    y3 = (rr * (v - x3) - 2 * y1 * jj) % P
    z3 = ((z1 + h) ** 2 - z1z1 - hh) % P
    return (x3, y3, z3)


def jacobian_mul(pt: Point, k: int) -> Jacobian:
    """Left-to-right double-and-add with mixed additions (curve.py:306-324)."""
    if k < 0:
        pt, k = -pt, -k
    if k == 0 or pt.is_identity():
        return JACOBIAN_IDENTITY
    x, y = pt.affine()
    acc = JACOBIAN_IDENTITY
    for bit in bin(k)[2:]:
        acc = jacobian_double(acc)
        if bit == "1":
            acc = jacobian_add_affine(acc, x, y)
    return acc
