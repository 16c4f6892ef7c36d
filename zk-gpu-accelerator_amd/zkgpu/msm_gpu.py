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
from collections.abc import Iterable, Sequence
from typing import Any, SupportsIndex

import numpy as np

from zkgpu import gpu_backend as _gb
from zkgpu.curve import P as _P, Point, from_jacobian
from zkgpu.field import FIELD_MODULUS as _GROUP_ORDER
from zkgpu.field256_gpu import BN254_FQ, LIMBS as _FIELD_LIMBS, _ingest, _limbs_to_ints, _rows_from_ints
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


def _check_points(points: DeviceArray, xs: Sequence[Any], ys: Sequence[Any]) -> None:
    """ValueError naming the first off-curve point (msm_gpu.py:441-442 checks each Point on the host)."""
    bad = ctypes.c_int64(-1)
    _gb.call("zk_g1_on_curve", points.ptr, points.shape[0], ctypes.byref(bad), _gb.current_stream())
    if bad.value >= 0:
        i = bad.value
        x, y = (operator.index(xs[i]) % _P, operator.index(ys[i]) % _P) if xs is not None else \
            tuple(BN254_FQ.to_ints(points[i]))
        raise ValueError(f"({x}, {y}) is not on the BN254 curve y^2 = x^3 + 3")


def _msm_rows(xrows: np.ndarray, yrows: np.ndarray, srows: np.ndarray, xs: Sequence[Any], ys: Sequence[Any],
              window: int, chunk: int) -> Point:
    """The MSM of host rows: coordinates below 2^256 (reduced mod p on the device), scalars below
    2^256 (reduced mod r on the device)."""
    n = xrows.shape[0]
    coords = np.empty((n, 2, _FIELD_LIMBS), dtype=np.uint64)
    coords[:, 0] = xrows
    coords[:, 1] = yrows
    points = BN254_FQ.from_limbs(coords.reshape(-1, _FIELD_LIMBS)).reshape(n, 2, _FIELD_LIMBS)
    _check_points(points, xs, ys)
    scalars = _gb.asarray(np.ascontiguousarray(srows, dtype=np.uint64))
    return _result(_run(points, scalars, n, window, chunk))


def _msm_ints(xs: list[int], ys: list[int], reduced: list[int], window: int, chunk: int) -> Point:
    xrows, xs = _rows_from_ints(xs, _P)
    yrows, ys = _rows_from_ints(ys, _P)
    srows, _ = _rows_from_ints(reduced, _GROUP_ORDER)
    return _msm_rows(xrows, yrows, srows, xs, ys, window, chunk)


def _scalar_rows(scalars: Any) -> DeviceArray:
    """Scalars as (n, 4) uint64 device rows: a DeviceArray or NumPy array of plain limbs as it is,
    anything else read as integers (reduced mod r)."""
    if isinstance(scalars, DeviceArray) or (isinstance(scalars, np.ndarray) and scalars.dtype == np.uint64):
        if scalars.ndim != 2 or scalars.shape[1] != _FIELD_LIMBS:
            raise TypeError(f"scalars must be an (n, 4) uint64 array of plain limbs, got shape {scalars.shape}")
        return _gb.asarray(np.ascontiguousarray(scalars) if isinstance(scalars, np.ndarray) else scalars)
    rows, _ = _rows_from_ints(scalars, _GROUP_ORDER, "scalars")
    return _gb.asarray(np.ascontiguousarray(rows, dtype=np.uint64))


def _msm_resident(points: Any, points_y: Any, scalars: Any, window: int) -> Point:
    """msm_pippenger_gpu over device-resident points: an SRS (an MSM over its first n points, with
    its window tables) or (n, 2, 4) Montgomery rows already in HBM."""
    from zkgpu.batch import SRS

    if points_y is not None:
        raise TypeError("points_y must be None when points_x holds device-resident points (an SRS or (n, 2, 4) rows)")
    _gb.require_gpu("msm_pippenger_gpu", _gpu_available)
    _kernels()
    sc = _scalar_rows(scalars)
    n = int(sc.shape[0])
    if isinstance(points, SRS):
        if n > points.n:
            raise ValueError(f"Length mismatch: an SRS of {points.n} points, {n} scalars.")
        if n == 0:
            return Point.identity()
        out = _gb.empty((3, _FIELD_LIMBS))
        ptrs = (ctypes.c_void_p * 1)(sc.ptr)
        sizes = (ctypes.c_int64 * 1)(n)
        wins = (ctypes.c_int32 * 1)(window)
        _gb.call("zk_msm_g1_batch", points.handle, 1, ptrs, sizes, wins, out.ptr, _gb.current_stream())
        return _result(out)
    if points.ndim != 3 or tuple(points.shape[1:]) != (2, _FIELD_LIMBS) or points.dtype != np.uint64:
        raise TypeError(f"points must be an (n, 2, 4) uint64 array, got {points.shape}")
    if points.shape[0] != n:
        raise ValueError(f"Length mismatch: {points.shape[0]} points, {n} scalars.")
    if n == 0:
        return Point.identity()
    _check_points(points, None, None)
    return _result(_run(points, sc, n, window, _CHUNK_LEN))


def _as_sequence(values: Any) -> Sequence[Any]:
    return values if isinstance(values, Sequence) else list(values)


def _msm_staged(points_x: Any, points_y: Any, scalars: Any, window: int) -> Point:
    """msm_pippenger_gpu of host integers. C threads convert each input straight into one
    page-locked block (x and y interleaved as (n, 2, 4) rows, then the scalar rows); the points go
    up in parts as they are converted, and are reduced, brought into Montgomery form and checked
    on the curve while the scalars are converted, which go up in parts too. The errors and their
    order are the reference's (msm_gpu.py:432-442): integer types of x, y and the scalars, the
    lengths, then the first point off the curve."""
    xs, ys = _as_sequence(points_x), _as_sequence(points_y)
    limbs = isinstance(scalars, (DeviceArray, np.ndarray)) and getattr(scalars, "dtype", None) == np.uint64
    ss = scalars if limbs else _as_sequence(scalars)
    m = max(len(xs), len(ys))
    with _gb.PinnedBuffer(64 * m + (0 if limbs else 32 * len(ss))) as buf:
        shapes_ok = not limbs or (ss.ndim == 2 and ss.shape[1] == _FIELD_LIMBS)
        n = (int(ss.shape[0]) if shapes_ok else 0) if limbs else len(ss)
        points = bad = sc = None
        raw = None
        if shapes_ok and len(xs) == len(ys) == n > 0:
            # the points go up in parts, each while the next one is converted
            _kernels()
            raw = _gb.empty((2 * n, _FIELD_LIMBS))
            part = max(1 << 16, -(-n // 4))
            try:
                for lo in range(0, n, part):
                    hi = min(n, lo + part)
                    _ingest.rows_into(xs, _P, "points_x", buf.view, 2, 0, lo, hi)
                    _ingest.rows_into(ys, _P, "points_y", buf.view, 2, 1, lo, hi)
                    buf.upload(raw[2 * lo:2 * hi], 64 * lo, 64 * (hi - lo))
            except TypeError:
                raw = None  # a non-integer: reported below in the reference's order, all of x first
        if raw is None:
            _ingest.rows_into(xs, _P, "points_x", buf.view, 2, 0)
            _ingest.rows_into(ys, _P, "points_y", buf.view, 2, 1)
        if not shapes_ok:
            raise TypeError(f"scalars must be an (n, 4) uint64 array of plain limbs, got shape {ss.shape}")
        if raw is not None:
            # the points' device work runs while the scalars are converted
            points = BN254_FQ.from_limbs(raw).reshape(n, 2, _FIELD_LIMBS)
            bad = _gb.empty((1,))
            _gb.call("zk_g1_on_curve_async", points.ptr, n, bad.ptr, _gb.current_stream())
            sc = _gb.empty((n, _FIELD_LIMBS))
        if limbs:
            pass
        elif sc is None:
            _ingest.rows_into(ss, _GROUP_ORDER, "scalars", buf.view, 1, 2 * m)
        else:
            part = max(1 << 16, -(-n // 4))
            for lo in range(0, n, part):
                hi = min(n, lo + part)
                _ingest.rows_into(ss, _GROUP_ORDER, "scalars", buf.view, 1, 2 * m, lo, hi)
                buf.upload(sc[lo:hi], 64 * m + 32 * lo, 32 * (hi - lo))
        if not len(xs) == len(ys) == n:
            raise ValueError(f"Length mismatch: {len(xs)} x-coordinates, {len(ys)} y-coordinates, {n} scalars.")
        if n == 0:
            return Point.identity()
        if limbs:
            sc = _gb.asarray(np.ascontiguousarray(ss) if isinstance(ss, np.ndarray) else ss)
        out = _gb.empty((3, _FIELD_LIMBS))
        w, chunk = _tuned(n, window, _CHUNK_LEN)
        _gb.call("zk_msm_g1", points.ptr, sc.ptr, n, w, chunk, out.ptr, _gb.current_stream())
        first_bad = int(bad.get()[0])
        if first_bad < n:
            x, y = operator.index(xs[first_bad]) % _P, operator.index(ys[first_bad]) % _P
            raise ValueError(f"({x}, {y}) is not on the BN254 curve y^2 = x^3 + 3")
        return _result(out)


def msm_pippenger_gpu(
    points_x: Iterable[int] | Any,
    points_y: Iterable[int] | None,
    scalars: Iterable[int] | Any,
    *,
    window_size: int | None = None,
) -> Point:
    """
    Pippenger MSM over BN254 G1 on the GPU: sum of s_i * (x_i, y_i).

    Parameters
    ----------
    points_x, points_y : affine coordinates (Python ints or an object array); or points_x a
                         device-resident basis (a zkgpu.batch.SRS, summed over its first n points,
                         or (n, 2, 4) Montgomery rows in HBM) with points_y None
    scalars            : any integers, reduced mod r, which leaves every s_i * P_i unchanged;
                         or (n, 4) uint64 plain limbs (a DeviceArray or NumPy array)
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
    from zkgpu.batch import SRS

    if isinstance(points_x, (SRS, DeviceArray)):
        return _msm_resident(points_x, points_y, scalars, window_size or 0)
    if _gpu_available and _ingest is not None:
        return _msm_staged(points_x, points_y, scalars, window_size or 0)
    xrows, xs = _rows_from_ints(points_x, _P, "points_x")
    yrows, ys = _rows_from_ints(points_y, _P, "points_y")
    if isinstance(scalars, (DeviceArray, np.ndarray)) and getattr(scalars, "dtype", None) == np.uint64:
        srows = scalars
    else:
        srows, _ = _rows_from_ints(scalars, _GROUP_ORDER, "scalars")
    if not xrows.shape[0] == yrows.shape[0] == srows.shape[0]:
        raise ValueError(
            f"Length mismatch: {xrows.shape[0]} x-coordinates, {yrows.shape[0]} y-coordinates, {srows.shape[0]} scalars."
        )
    if xrows.shape[0] == 0:
        return Point.identity()
    _gb.require_gpu("msm_pippenger_gpu", _gpu_available)
    _kernels()
    if isinstance(srows, DeviceArray):
        n = xrows.shape[0]
        coords = np.empty((n, 2, _FIELD_LIMBS), dtype=np.uint64)
        coords[:, 0], coords[:, 1] = xrows, yrows
        points = BN254_FQ.from_limbs(coords.reshape(-1, _FIELD_LIMBS)).reshape(n, 2, _FIELD_LIMBS)
        _check_points(points, xs, ys)
        return _result(_run(points, srows, n, window_size or 0, _CHUNK_LEN))
    return _msm_rows(xrows, yrows, srows, xs, ys, window_size or 0, _CHUNK_LEN)


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
