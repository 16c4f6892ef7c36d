"""
zkgpu.curve_gpu: BN254 G1 point arithmetic on the MI355X, in Jacobian coordinates.

The reference's element-wise point API (src/zkgpu/curve_gpu.py:354-475):
add, add_affine and double over (n, 3, 4) Jacobian and (n, 2, 4) affine
uint64 rows of Montgomery-form BN254_FQ coordinates, with a single point
broadcasting against a batch. The kernels (csrc/zk_curve.hpp) use the
reference's formulas (EFD add-2007-bl, madd-2007-bl, dbl-2009-l) and special
cases, so the GPU returns exactly zkgpu.curve's (X, Y, Z), not merely the
same affine point. The identity is (1, 1, 0); affine rows cannot hold it.
Without a usable GPU every function raises RuntimeError (no CPU fallback).
"""

from __future__ import annotations

from collections.abc import Iterable
from typing import Any

import numpy as np

from zkgpu import gpu_backend as _gb
from zkgpu.curve import Jacobian, Point, from_jacobian, to_jacobian
from zkgpu.field256_gpu import _CSRC, BN254_FQ, LIMBS
from zkgpu.gpu_backend import GPU_AVAILABLE as _gpu_available, DeviceArray

_compiled: dict[str, Any] = {}  # filled once the self-test passes


def cuda_source() -> str:
    """BN254_FQ's field source followed by the point formulas (zk_curve.hpp)."""
    return BN254_FQ.cuda_source() + (_CSRC / "zk_curve.hpp").read_text()


hip_source = cuda_source


def _kernels() -> dict[str, Any]:
    if not _compiled:
        _gb.require_gpu("curve_gpu", _gpu_available)
        _compiled.update({"add": "zk_g1_add", "add_affine": "zk_g1_add_affine", "double": "zk_g1_double"})
        try:
            _self_test()
        except Exception:
            _compiled.clear()
            raise
    return _compiled


def _self_test() -> None:
    """
    Every branch against exact Jacobian values: general Z, equal points,
    inverse points, the identity on either side, once per process. The
    expected triples are fixed data (small multiples of G), so the check needs
    no CPU point arithmetic.
    """
    from zkgpu import _selftest_vectors as v

    checks = {
        "add": (add(jacobian_to_device(v.ADD_A), jacobian_to_device(v.ADD_B)), v.ADD_OUT),
        "add_affine": (add_affine(jacobian_to_device(v.MADD_A), _affine_rows(v.MADD_B)), v.MADD_OUT),
        "double": (double(jacobian_to_device(v.DBL_A)), v.DBL_OUT),
    }
    for op, (rows, expected) in checks.items():
        if jacobian_ints(rows) != [tuple(t) for t in expected]:
            raise RuntimeError(f"curve_gpu: the {op} kernel failed its self-test: libzkgpu_hip.so computed wrong points")


def compile_kernels() -> None:
    """Load and self-test the point kernels now (RuntimeError if they fail)."""
    _kernels()


# ------------------------------------------------------------------
# Conversion
# ------------------------------------------------------------------


def jacobian_to_device(points: Iterable[Jacobian]) -> DeviceArray:
    """(n, 3, 4) rows for Jacobian (X, Y, Z) int triples."""
    flat = [c for point in points for c in point]
    if len(flat) % 3:
        raise ValueError("Jacobian points must be (X, Y, Z) triples")
    return BN254_FQ.from_ints(flat).reshape(-1, 3, LIMBS)


def to_device(points: Iterable[Point]) -> DeviceArray:
    """(n, 3, 4) Jacobian rows for affine Points (Z = 1; the identity becomes (1, 1, 0))."""
    return jacobian_to_device(to_jacobian(p) for p in points)


def _affine_rows(pairs: Iterable[tuple[int, int]]) -> DeviceArray:
    flat = [c for xy in pairs for c in xy]
    return BN254_FQ.from_ints(flat).reshape(-1, 2, LIMBS)


def affine_to_device(points: Iterable[Point]) -> DeviceArray:
    """(n, 2, 4) affine rows; the identity is rejected."""
    pairs = []
    for p in points:
        if p.is_identity():
            raise ValueError("affine rows cannot hold the point at infinity; leave it out (it adds nothing)")
        pairs.append(p.affine())
    return _affine_rows(pairs)


def jacobian_ints(x: Any) -> list[Jacobian]:
    """The exact (X, Y, Z) coordinates of Jacobian rows, as Python ints."""
    _point_count(x, "x", 3)
    flat = BN254_FQ.to_ints(x.reshape(-1, LIMBS))
    return [(flat[k], flat[k + 1], flat[k + 2]) for k in range(0, len(flat), 3)]


def from_device(x: Any) -> list[Point]:
    """Affine Points for Jacobian rows; each is re-validated on the curve by Point."""
    return [from_jacobian(j) for j in jacobian_ints(x)]


# ------------------------------------------------------------------
# Arithmetic
# ------------------------------------------------------------------


def add(a: Any, b: Any) -> DeviceArray:
    """Element-wise a + b of Jacobian rows (either may be one broadcast point)."""
    return _launch_binary("add", a, b, 3)


def add_affine(a: Any, b: Any) -> DeviceArray:
    """Element-wise a + b of Jacobian rows a and affine rows b (mixed addition)."""
    return _launch_binary("add_affine", a, b, 2)


def double(a: Any) -> DeviceArray:
    """Element-wise 2a of Jacobian rows."""
    n = _point_count(a, "a", 3)
    name = _kernels()["double"]
    src = _gb.asarray(a)
    out = _gb.empty((n, 3, LIMBS))
    if n:
        _gb.call(name, src.ptr, out.ptr, n, _gb.current_stream())
    return out


# ------------------------------------------------------------------
# Internals
# ------------------------------------------------------------------


def _point_count(x: Any, name: str, coords: int) -> int:
    """Points in x after checking it is an (n, coords, 4) uint64 array (curve_gpu.py:422-430)."""
    if getattr(x, "dtype", None) != np.uint64 or getattr(x, "ndim", None) != 3 or tuple(x.shape[1:]) != (coords, LIMBS):
        kind = "Jacobian" if coords == 3 else "affine"
        raise TypeError(
            f"{name} must be an (n, {coords}, {LIMBS}) uint64 array of {kind} points, "
            f"got {getattr(x, 'dtype', type(x).__name__)} with shape {getattr(x, 'shape', None)}"
        )
    return int(x.shape[0])


def _broadcast(a: Any, b: Any, b_coords: int) -> int:
    rows_a, rows_b = _point_count(a, "a", 3), _point_count(b, "b", b_coords)
    if rows_a == rows_b or rows_b == 1:
        return rows_a
    if rows_a == 1:
        return rows_b
    raise ValueError(f"cannot broadcast {rows_a} points against {rows_b} points")


def _launch_binary(op: str, a: Any, b: Any, b_coords: int) -> DeviceArray:
    n = _broadcast(a, b, b_coords)
    name = _kernels()[op]
    a, b = _gb.asarray(a), _gb.asarray(b)
    out = _gb.empty((n, 3, LIMBS))
    if n:
        _gb.call(name, a.ptr, 1 if a.shape[0] == 1 else 0, b.ptr, 1 if b.shape[0] == 1 else 0, out.ptr, n, _gb.current_stream())
    return out
