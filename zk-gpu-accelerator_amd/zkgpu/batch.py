"""
zkgpu.batch: many MSMs and NTTs of uneven sizes as one launch sequence.

SRS holds a device-resident basis, converted once into the kernels' internal
layout (the reference converts and uploads its points on every call,
src/zkgpu/msm_gpu.py:348). QuotientBatch runs the Groth16 quotient of several
jobs of mixed sizes as one launch sequence. MsmBatch plans a set of MSMs over prefixes of an
SRS: one digit sort, one load-balanced bucket accumulation, one bucket
reduction and one window combination for all of them together, each MSM with
its own window width. NttBatch runs several Fr NTTs of different sizes as one
launch per pass index. Both allocate everything at plan time, so launch()
enqueues a fixed kernel sequence that can be recorded once into a HIP graph
(capture()) and replayed.

Window widths default to the autotuned table (zkgpu.autotune) when one exists
for this GPU, else to the library's cost model.
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence
from typing import Any

import numpy as np

from zkgpu import gpu_backend as _gb
from zkgpu.curve import Point, from_jacobian
from zkgpu.field256_gpu import BN254_FQ, _limbs_to_ints
from zkgpu.gpu_backend import DeviceArray


def _ptr_array(ptrs: Sequence[int]) -> Any:
    return (ctypes.c_void_p * max(len(ptrs), 1))(*ptrs)


class SRS:
    """A device-resident G1 basis: n affine Montgomery points (finite, on the curve)."""

    def __init__(self, points: Any) -> None:
        if getattr(points, "ndim", None) != 3 or tuple(points.shape[1:]) != (2, 4) or points.dtype != np.uint64:
            raise TypeError(f"points must be an (n, 2, 4) uint64 array, got {getattr(points, 'shape', None)}")
        _gb.require_gpu("SRS")
        self.points = _gb.asarray(points)
        self.n = int(points.shape[0])
        h = ctypes.c_void_p()
        _gb.call("zk_srs_create", self.points.ptr, self.n, ctypes.byref(h), _gb.current_stream())
        self.handle = h.value

    def precompute(self, windows: Sequence[int] = (16,)) -> SRS:
        """
        Add window tables (2^(c w) P_i for every window of width c): MSMs planned afterwards put
        all windows' digits into one set of buckets, with no window combination. W x 64 bytes per
        point each (W = floor(254 / c) + 1); add them before creating plans.
        """
        for c in windows:
            _gb.call("zk_srs_precompute", self.handle, int(c), _gb.current_stream())
        self.tables = sorted(set(getattr(self, "tables", [])) | {int(c) for c in windows})
        return self

    @classmethod
    def from_keys(cls, keys: np.ndarray) -> SRS:
        """The basis k_i * G for plain (n, 4) scalar rows k_i, generated on the GPU."""
        from zkgpu.workloads import srs_points

        return cls(srs_points(keys))

    def __del__(self) -> None:
        if _gb.LIB is not None and getattr(self, "handle", None):
            _gb.LIB.zk_srs_destroy(self.handle)
            self.handle = None


def _tuned_window(n: int) -> int:
    from zkgpu.autotune import lookup

    return lookup("msm_window", n)


def _tuned_chunk(sizes: Sequence[int]) -> int:
    from zkgpu.autotune import lookup

    return lookup("msm_chunk", max(sizes)) if sizes else 0


class MsmBatch:
    """
    count MSMs over prefixes of one SRS. scalars[j] is an (n_j, 4) uint64 device
    array of plain scalars (read at every launch). Results land in self.out,
    a (count, 3, 4) Jacobian device array; results() converts them to Points.
    """

    def __init__(self, srs: SRS, scalars: Sequence[DeviceArray], *, windows: Sequence[int] | None = None,
                 chunk_len: int | None = None) -> None:
        _gb.require_gpu("MsmBatch")
        self.srs = srs
        self.scalars = [_gb.asarray(s) for s in scalars]
        self.sizes = [int(s.shape[0]) for s in self.scalars]
        if any(n > srs.n for n in self.sizes):
            raise ValueError(f"an MSM is larger than the SRS ({srs.n} points)")
        count = len(self.scalars)
        if windows is not None:
            self.windows = list(windows)
        elif getattr(srs, "tables", None):
            self.windows = [0] * count  # the library picks each job's window table
        else:
            self.windows = [_tuned_window(n) for n in self.sizes]
        if chunk_len is None:
            chunk_len = _tuned_chunk(self.sizes)
        self.out = _gb.empty((max(count, 1), 3, 4))
        h = ctypes.c_void_p()
        sizes = (ctypes.c_int64 * max(count, 1))(*self.sizes)
        wins = (ctypes.c_int32 * max(count, 1))(*self.windows)
        _gb.call(
            "zk_msm_plan_create", srs.handle, count, _ptr_array([s.ptr for s in self.scalars]), sizes, wins,
            chunk_len, self.out.ptr, ctypes.byref(h), _gb.current_stream(),
        )
        self.handle = h.value
        self.count = count

    def launch(self) -> None:
        _gb.call("zk_msm_plan_launch", self.handle, _gb.current_stream())

    @property
    def nbytes(self) -> int:
        """Device bytes the plan holds."""
        out = ctypes.c_uint64()
        _gb.call("zk_msm_plan_bytes", self.handle, ctypes.byref(out))
        return int(out.value)

    def results(self) -> list[Point]:
        rows = BN254_FQ.to_limbs(self.out.reshape(-1, 4)).get()
        flat = _limbs_to_ints(rows)
        return [from_jacobian((flat[3 * j], flat[3 * j + 1], flat[3 * j + 2])) for j in range(self.count)]

    def __del__(self) -> None:
        if _gb.LIB is not None and getattr(self, "handle", None):
            _gb.LIB.zk_msm_plan_destroy(self.handle)
            self.handle = None


class NttBatch:
    """Several Fr NTTs of different sizes (inputs and outputs (n_j, 4) Montgomery device rows)."""

    def __init__(self, inputs: Sequence[DeviceArray], outputs: Sequence[DeviceArray], inverse: Sequence[bool]) -> None:
        _gb.require_gpu("NttBatch")
        if not len(inputs) == len(outputs) == len(inverse):
            raise ValueError("inputs, outputs and inverse must have the same length")
        self.inputs = [_gb.asarray(x) for x in inputs]
        self.outputs = list(outputs)
        log_n = []
        for x, y in zip(self.inputs, self.outputs):
            n = int(x.shape[0])
            if n < 2 or n & (n - 1) or tuple(y.shape) != (n, 4):
                raise ValueError("batched NTT sizes must be powers of two >= 2 with matching (n, 4) outputs")
            log_n.append(n.bit_length() - 1)
        count = len(self.inputs)
        h = ctypes.c_void_p()
        _gb.call(
            "zk_ntt_plan_create", count, _ptr_array([x.ptr for x in self.inputs]), _ptr_array([y.ptr for y in self.outputs]),
            (ctypes.c_int32 * max(count, 1))(*log_n), (ctypes.c_int32 * max(count, 1))(*[1 if i else 0 for i in inverse]),
            ctypes.byref(h), _gb.current_stream(),
        )
        self.handle = h.value

    def launch(self) -> None:
        _gb.call("zk_ntt_plan_launch", self.handle, _gb.current_stream())

    def __del__(self) -> None:
        if _gb.LIB is not None and getattr(self, "handle", None):
            _gb.LIB.zk_ntt_plan_destroy(self.handle)
            self.handle = None


class QuotientBatch:
    """
    The Groth16 quotient h_j = (A_j B_j - C_j) / Z_H on the coset g_j H for several jobs of mixed
    sizes (Radix2Domain.quotient, one launch per stage and pass index for all of them). a, b, c
    are (n_j, 4) Montgomery rows on the device; results land in h (allocated here unless given).
    """

    def __init__(self, a: Sequence[DeviceArray], b: Sequence[DeviceArray], c: Sequence[DeviceArray], *,
                 shifts: Sequence[int] | int = 7, h: Sequence[DeviceArray] | None = None) -> None:
        _gb.require_gpu("QuotientBatch")
        if not len(a) == len(b) == len(c):
            raise ValueError("a, b and c must have the same length")
        self.a, self.b, self.c = ([_gb.asarray(x) for x in v] for v in (a, b, c))
        count = len(self.a)
        log_n = []
        for x, y, z in zip(self.a, self.b, self.c):
            n = int(x.shape[0])
            if n < 2 or n & (n - 1) or tuple(x.shape) != (n, 4) or tuple(y.shape) != (n, 4) or tuple(z.shape) != (n, 4):
                raise ValueError("quotient sizes must be powers of two >= 2 with (n, 4) rows for a, b and c")
            log_n.append(n.bit_length() - 1)
        self.log_n = log_n
        if isinstance(shifts, int):
            shifts = [shifts] * count
        if len(shifts) != count:
            raise ValueError("one coset shift per job")
        self.h = list(h) if h is not None else [_gb.empty((1 << k, 4)) for k in log_n]
        shift_rows = np.zeros((max(count, 1), 4), dtype=np.uint64)
        for j, g in enumerate(shifts):
            shift_rows[j] = [(int(g) >> (64 * k)) & (2**64 - 1) for k in range(4)]
        hnd = ctypes.c_void_p()
        _gb.call(
            "zk_groth16_plan_create", count, _ptr_array([x.ptr for x in self.a]), _ptr_array([x.ptr for x in self.b]),
            _ptr_array([x.ptr for x in self.c]), _ptr_array([x.ptr for x in self.h]),
            (ctypes.c_int32 * max(count, 1))(*log_n), shift_rows.ctypes.data, ctypes.byref(hnd), _gb.current_stream(),
        )
        self.handle = hnd.value

    def launch(self) -> None:
        _gb.call("zk_groth16_plan_launch", self.handle, _gb.current_stream())

    def __del__(self) -> None:
        if _gb.LIB is not None and getattr(self, "handle", None):
            _gb.LIB.zk_groth16_plan_destroy(self.handle)
            self.handle = None


def quotient_workspace_bytes(log_n: Sequence[int]) -> int:
    """Device bytes a QuotientBatch over these sizes holds (besides its inputs and outputs)."""
    out = ctypes.c_uint64()
    _gb.call("zk_groth16_workspace_bytes", len(log_n), (ctypes.c_int32 * max(len(log_n), 1))(*log_n), ctypes.byref(out))
    return int(out.value)


def msm_workspace_bytes(sizes: Sequence[int], windows: Sequence[int] | None = None, *, srs: SRS | None = None,
                        chunk_len: int | None = None) -> int:
    """Device bytes an MsmBatch over these sizes holds (besides its scalars and results): over
    `srs` the plan's own window-table choice applies; chunk_len defaults to the tuned chunk of the
    largest size, as MsmBatch picks it."""
    out = ctypes.c_uint64()
    count = len(sizes)
    if windows is None:
        windows = [0] * count if srs is not None and getattr(srs, "tables", None) else [_tuned_window(n) for n in sizes]
    if chunk_len is None:
        chunk_len = _tuned_chunk(sizes)
    wins = (ctypes.c_int32 * max(count, 1))(*windows)
    _gb.call("zk_msm_workspace_bytes", srs.handle if srs is not None else None, count,
             (ctypes.c_int64 * max(count, 1))(*sizes), wins, int(chunk_len), ctypes.byref(out))
    return int(out.value)


def tuned_chunks() -> list[int]:
    """Every accumulation chunk length the tuning table of this GPU may pick (0 = persistent grid)."""
    from zkgpu.autotune import load

    table = load() or {}
    return sorted({0} | {int(v) for v in table.get("msm_chunk", {}).values()})


def capture(work: Sequence[Any], on: _gb.Stream) -> _gb.Graph:
    """Record the launch sequence of several plans (MsmBatch / NttBatch) into one HIP graph."""
    with _gb.Graph.capture(on) as box:
        for w in work:
            w.launch()
    return box[0]
