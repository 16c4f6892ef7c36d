"""
zkgpu.msm_gpu: Pippenger multi-scalar multiplication over BN254 G1 on the MI355X.

msm_pippenger_gpu keeps the reference's signature, validation order, errors
and return type (src/zkgpu/msm_gpu.py:398-449): Python ints (or object
arrays) in, an affine zkgpu.curve.Point out, Point.identity() for the point at
infinity or an empty input. Everything between runs on the device
(csrc/zk_msm.hip): signed-digit recoding, a counting sort of the bucket keys,
load-balanced bucket accumulation, the bucket reduction and the window
combination; the host converts the single Jacobian result to affine.

Points supplied by the caller are validated on the curve on the device, and
an off-curve point raises the reference's ValueError for the first bad index.

msm_g1 is the device-resident form: points as (n, 2, 4) Montgomery rows and
scalars as (n, 4) plain limbs already in HBM (e.g. a shared SRS prefix and a
witness vector), with no host conversion. Without a usable GPU both raise
RuntimeError: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import operator
from collections.abc import Iterable
from typing import Any, SupportsIndex

import numpy as np

from zkgpu import gpu_backend as _gb
from zkgpu.curve import P as _P, Point, from_jacobian
from zkgpu.field import FIELD_MODULUS as _GROUP_ORDER
from zkgpu.field256_gpu import BN254_FQ, LIMBS as _FIELD_LIMBS, _ints_to_limbs, _limbs_to_ints
from zkgpu.gpu_backend import GPU_AVAILABLE as _gpu_available, DeviceArray
from zkgpu.msm import _check_window_size

# Sorted bucket entries per thread in the first accumulation level (the
# reference's name for its chunk length, msm_gpu.py:75); 0 lets the library pick.
_CHUNK_LEN = 0
_SCALAR_BITS = _GROUP_ORDER.bit_length()

_compiled: dict[str, Any] = {}


def _kernels() -> dict[str, Any]:
    if not _compiled:
        _gb.require_gpu("msm_gpu", _gpu_available)
        _compiled["msm"] = "zk_msm_g1"
        try:
            _self_test()
        except Exception:
            _compiled.clear()
            raise
    return _compiled


def _self_test() -> None:
    """A small MSM with every bucket case (several points, P and -P, P twice, digit 0) against fixed data."""
    from zkgpu import _selftest_vectors as v

    xs = [p[0] for p in v.MSM_POINTS]
    ys = [p[1] for p in v.MSM_POINTS]
    got = _msm_ints(xs, ys, [s % _GROUP_ORDER for s in v.MSM_SCALARS], v.MSM_WINDOW, 2)
    if got.is_identity() or list(got.affine()) != list(v.MSM_OUT):
        raise RuntimeError("msm_gpu: the MSM kernels failed their self-test: libzkgpu_hip.so computed a wrong sum")


def compile_kernels() -> None:
    """Load and self-test the MSM kernels now (RuntimeError if they fail)."""
    _kernels()


def _as_ints(values: Iterable[SupportsIndex], name: str) -> list[int]:
    """Plain Python ints; operator.index rejects floats instead of truncating them."""
    try:
        return [operator.index(v) for v in values]
    except TypeError as exc:
        raise TypeError(f"{name} must contain integers: {exc}") from None


def _result(jac_rows: DeviceArray) -> Point:
    x, y, z = _limbs_to_ints(BN254_FQ.to_limbs(jac_rows.reshape(3, _FIELD_LIMBS)).get())
    return from_jacobian((x, y, z))


def _tuned(n: int, window: int, chunk: int) -> tuple[int, int]:
    """0 means automatic: this GPU's autotuned window and accumulation chunk, else the cost model."""
    from zkgpu.autotune import lookup

    return (window or lookup("msm_window", n)), (chunk or lookup("msm_chunk", n))


def _run(points: DeviceArray, scalars: DeviceArray, n: int, window: int, chunk: int) -> DeviceArray:
    out = _gb.empty((3, _FIELD_LIMBS))
    window, chunk = _tuned(n, window, chunk)
    _gb.call("zk_msm_g1", points.ptr, scalars.ptr, n, window, chunk, out.ptr, _gb.current_stream())
    return out


def _check_points(points: DeviceArray, xs: list[int], ys: list[int]) -> None:
    bad = ctypes.c_int64(-1)
    _gb.call("zk_g1_on_curve", points.ptr, points.shape[0], ctypes.byref(bad), _gb.current_stream())
    if bad.value >= 0:
        i = bad.value
        raise ValueError(f"({xs[i] % _P}, {ys[i] % _P}) is not on the BN254 curve y^2 = x^3 + 3")


def _msm_ints(xs: list[int], ys: list[int], reduced: list[int], window: int, chunk: int) -> Point:
    n = len(xs)
    coords = _ints_to_limbs(c % _P for xy in zip(xs, ys) for c in xy)
    points = BN254_FQ.from_limbs(coords).reshape(n, 2, _FIELD_LIMBS)
    _check_points(points, xs, ys)
    scalars = _gb.asarray(_ints_to_limbs(reduced))
    return _result(_run(points, scalars, n, window, chunk))


def msm_pippenger_gpu(
    points_x: Iterable[int],
    points_y: Iterable[int],
    scalars: Iterable[int],
    *,
    window_size: int | None = None,
) -> Point:
    """
    Pippenger MSM over BN254 G1 on the GPU: sum of s_i * (x_i, y_i).

    Parameters
    ----------
    points_x, points_y : affine coordinates (Python ints or an object array)
    scalars            : any integers; reduced mod r, which leaves every s_i * P_i unchanged
    window_size        : bits per signed digit, 1..16 (default: this GPU's autotuned
                         window, else the device cost model)

    Raises
    ------
    ValueError : if the inputs differ in length, window_size is out of range,
                 or a point is not on the curve
    TypeError  : if a coordinate or scalar is not an integer
    RuntimeError : if no GPU is usable
    """
    if window_size is not None:
        window_size = _check_window_size(window_size)
    xs, ys = _as_ints(points_x, "points_x"), _as_ints(points_y, "points_y")
    reduced = [s % _GROUP_ORDER for s in _as_ints(scalars, "scalars")]
    if not len(xs) == len(ys) == len(reduced):
        raise ValueError(
            f"Length mismatch: {len(xs)} x-coordinates, {len(ys)} y-coordinates, {len(reduced)} scalars."
        )
    if not reduced:
        return Point.identity()
    _gb.require_gpu("msm_pippenger_gpu", _gpu_available)
    _kernels()
    return _msm_ints(xs, ys, reduced, window_size or 0, _CHUNK_LEN)


def msm_g1(points: Any, scalars: Any, *, window_size: int | None = None) -> Point:
    """
    MSM of device-resident inputs: points (n, 2, 4) Montgomery BN254_FQ rows
    (finite, on the curve), scalars (n, 4) plain uint64 limbs (< 2^256,
    reduced mod r on the device). Returns the affine Point.
    """
    if getattr(points, "ndim", None) != 3 or tuple(points.shape[1:]) != (2, _FIELD_LIMBS) or points.dtype != np.uint64:
        raise TypeError(f"points must be an (n, 2, 4) uint64 array, got {getattr(points, 'shape', None)}")
    if getattr(scalars, "ndim", None) != 2 or scalars.shape[1] != 4 or scalars.dtype != np.uint64:
        raise TypeError(f"scalars must be an (n, 4) uint64 array, got {getattr(scalars, 'shape', None)}")
    if points.shape[0] != scalars.shape[0]:
        raise ValueError(f"Length mismatch: {points.shape[0]} points, {scalars.shape[0]} scalars.")
    if window_size is not None:
        window_size = _check_window_size(window_size)
    n = points.shape[0]
    if n == 0:
        return Point.identity()
    _gb.require_gpu("msm_g1", _gpu_available)
    _kernels()
    return _result(_run(_gb.asarray(points), _gb.asarray(scalars), n, window_size or 0, _CHUNK_LEN))


def msm_g1_jacobian(points: DeviceArray, scalars: DeviceArray, out: DeviceArray, *, window_size: int = 0, chunk_len: int = 0) -> None:
    """Enqueue an MSM whose Jacobian result (3, 4) Montgomery rows stays on the device in out."""
    n = points.shape[0]
    window_size, chunk_len = _tuned(n, window_size, chunk_len)
    _gb.call("zk_msm_g1", points.ptr, scalars.ptr, n, window_size, chunk_len, out.ptr, _gb.current_stream())
