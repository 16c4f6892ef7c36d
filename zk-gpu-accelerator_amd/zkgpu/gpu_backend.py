"""
zkgpu.gpu_backend: the HIP runtime under every GPU module, decided once at import.

Replaces the reference's CuPy/NVRTC backend (src/zkgpu/gpu_backend.py:79-169)
with a loader for libzkgpu_hip.so, the ahead-of-time gfx950 build of this
package's hand-written HIP kernels, reached through its C ABI (include/zkgpu.h).

Detection stages (the first failure wins), recorded in GPU_UNAVAILABLE_REASON:
  1. the shared library exists and loads (it is built by __graft_entry__.build)
  2. the HIP runtime sees at least one device
  3. a probe kernel runs and returns the expected values

There is no CPU fallback anywhere in this package: every GPU entry point calls
require_gpu(), which raises RuntimeError with the reason instead of computing
on the CPU. GpuFallbackWarning and warn_cpu_fallback keep the reference's names
for callers that import them, and are never emitted by this package.

Device memory is exposed as DeviceArray, the subset of a CuPy ndarray that the
reference's callers use: dtype, shape, ndim, size, reshape (a view), row
slicing (a view), get() / tolist() to the host, and copy().
"""

from __future__ import annotations

import contextlib
import ctypes
import threading
import warnings
from collections.abc import Iterator, Sequence
from pathlib import Path
from typing import Any

import numpy as np

_LIB_NAME = "libzkgpu_hip.so"
LIB_PATH = Path(__file__).resolve().parent / "_lib" / _LIB_NAME


class GpuFallbackWarning(UserWarning):
    """Kept for API compatibility with the reference; this package never falls back to the CPU."""


class ZkError(RuntimeError):
    """A call into libzkgpu_hip.so failed; the message is the library's own."""


# ----------------------------------------------------------------------
# Library loading and the C ABI signatures
# ----------------------------------------------------------------------

_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)

# name -> argument types; every entry returns an int status (0 = success)
SIGNATURES: dict[str, list[Any]] = {
    "zk_init": [_i32],
    "zk_device_count": [ctypes.POINTER(_i32)],
    "zk_device_info": [ctypes.c_char_p, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32)],
    "zk_mem_info": [_u64p, _u64p],
    "zk_alloc": [ctypes.POINTER(_vp), _u64, _vp],
    "zk_free": [_vp],
    "zk_memcpy_h2d": [_vp, _vp, _u64, _vp],
    "zk_memcpy_d2h": [_vp, _vp, _u64, _vp],
    "zk_memcpy_d2d": [_vp, _vp, _u64, _vp],
    "zk_memset": [_vp, _i32, _u64, _vp],
    "zk_host_alloc": [ctypes.POINTER(_vp), _u64],
    "zk_host_free": [_vp],
    "zk_memcpy_h2d_async": [_vp, _vp, _u64, _vp],
    "zk_copy": [_vp, _vp, _u64, _vp],
    "zk_stream_create": [ctypes.POINTER(_vp)],
    "zk_stream_destroy": [_vp],
    "zk_stream_sync": [_vp],
    "zk_device_sync": [],
    "zk_event_create": [ctypes.POINTER(_vp)],
    "zk_event_destroy": [_vp],
    "zk_event_record": [_vp, _vp],
    "zk_event_elapsed_ms": [_vp, _vp, ctypes.POINTER(ctypes.c_float)],
    "zk_event_query": [_vp, ctypes.POINTER(_i32)],
    "zk_event_sync": [_vp],
    "zk_stream_wait_event": [_vp, _vp],
    "zk_selftest": [],
    "zk_graph_begin": [_vp],
    "zk_graph_end": [_vp, ctypes.POINTER(_vp)],
    "zk_graph_launch": [_vp, _vp],
    "zk_graph_destroy": [_vp],
    "zk_valu_peak": [ctypes.POINTER(ctypes.c_double)],
    "zk_product_peak": [ctypes.POINTER(ctypes.c_double)],
    "zk_msm_default_window": [_i64, ctypes.POINTER(_i32)],
    "zk_msm_set_glv": [_i32],
    "zk_msm_get_glv": [ctypes.POINTER(_i32)],
    "zk_msm_set_accumulate": [_i32],
    "zk_msm_get_accumulate": [ctypes.POINTER(_i32)],
    "zk_msm_set_segment": [_i32],
    "zk_msm_get_segment": [ctypes.POINTER(_i32)],
    "zk_prof_enable": [_i32],
    "zk_prof_read": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), _i64p, _i32],
    "zk_field_register": [_u64p, ctypes.POINTER(_i32)],
    "zk_field_to_mont": [_i32, _vp, _vp, _i64, _vp],
    "zk_field_from_mont": [_i32, _vp, _vp, _i64, _vp],
    "zk_field_mul": [_i32, _vp, _i32, _vp, _i32, _vp, _i64, _vp],
    "zk_field_add": [_i32, _vp, _i32, _vp, _i32, _vp, _i64, _vp],
    "zk_field_sub": [_i32, _vp, _i32, _vp, _i32, _vp, _i64, _vp],
    "zk_ntt": [_i32, _vp, _vp, _i32, _i32, _u64p, _u64p, _vp],
    "zk_ntt_cache_clear": [],
    "zk_ntt_cache_stats": [_u64p, _u64p, _u64p],
    "zk_f64_binary": [_i32, _vp, _i32, _i64p, _vp, _i32, _i64p, _i32, _i64p, _vp, _u64, _vp],
    "zk_f64_pow": [_vp, _i32, _i64p, _i32, _i64p, _u64p, _i32, _vp, _u64, _vp],
    "zk_f64_any_zero": [_vp, _i32, _i64p, _i32, _i64p, _u64, ctypes.POINTER(_i32), _vp],
    "zk_ntt64": [_vp, _i32, _vp, _i32, _u64, _u64, _i32, _vp],
    "zk_ntt_set_fusion": [_i32, _i32],
    "zk_ntt_get_fusion": [_i32, ctypes.POINTER(_i32)],
    "zk_groth16_quotient": [_vp, _vp, _vp, _vp, _i32, _u64p, _vp],
    "zk_groth16_plan_create": [_i32, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp), _vp],
    "zk_groth16_plan_launch": [_vp, _vp],
    "zk_groth16_plan_destroy": [_vp],
    "zk_groth16_workspace_bytes": [_i32, _vp, ctypes.POINTER(_u64)],
    "zk_g1_add": [_vp, _i32, _vp, _i32, _vp, _i64, _vp],
    "zk_g1_add_affine": [_vp, _i32, _vp, _i32, _vp, _i64, _vp],
    "zk_g1_double": [_vp, _vp, _i64, _vp],
    "zk_g1_on_curve": [_vp, _i64, _i64p, _vp],
    "zk_g1_on_curve_async": [_vp, _i64, _vp, _vp],
    "zk_g1_fixed_base": [_vp, _i64, _vp, _vp],
    "zk_msm_g1": [_vp, _vp, _i64, _i32, _i32, _vp, _vp],
    "zk_msm_g1_entries": [_vp, _vp, _i64, _i32, _i32, _vp, _u64p, _vp],
    "zk_msm_g1_noalloc": [_vp, _vp, _i64, _i32, _i32, _vp, _vp],
    "zk_msm_noalloc_max_bytes": [ctypes.POINTER(_u64)],
    "zk_srs_create": [_vp, _i64, ctypes.POINTER(_vp), _vp],
    "zk_srs_destroy": [_vp],
    "zk_srs_precompute": [_vp, _i32, _vp],
    "zk_msm_plan_create": [_vp, _i32, _vp, _vp, _vp, _i32, _vp, ctypes.POINTER(_vp), _vp],
    "zk_msm_plan_launch": [_vp, _vp],
    "zk_msm_plan_destroy": [_vp],
    "zk_msm_workspace_bytes": [_vp, _i32, _vp, _vp, _i32, ctypes.POINTER(_u64)],
    "zk_msm_plan_bytes": [_vp, ctypes.POINTER(_u64)],
    "zk_msm_g1_batch": [_vp, _i32, _vp, _vp, _vp, _vp, _vp],
    "zk_ntt_batch": [_i32, _vp, _vp, _vp, _vp, _vp],
    "zk_ntt_plan_create": [_i32, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp), _vp],
    "zk_ntt_plan_launch": [_vp, _vp],
    "zk_ntt_plan_destroy": [_vp],
}
STRING_RETURNS = {"zk_last_error": []}


def _load() -> tuple[Any, str | None]:
    if not LIB_PATH.exists():
        return None, f"{_LIB_NAME} is not built (run __graft_entry__.build())"
    try:
        lib = ctypes.CDLL(str(LIB_PATH))
    except OSError as exc:
        return None, f"{_LIB_NAME} failed to load: {exc}"
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.zk_last_error.argtypes = []
    lib.zk_last_error.restype = ctypes.c_char_p
    return lib, None


def _detect(lib: Any) -> str | None:
    count = ctypes.c_int(0)
    if lib.zk_device_count(ctypes.byref(count)) != 0 or count.value == 0:
        return "the HIP runtime found no GPU device"
    if lib.zk_init(0) != 0:
        return f"HIP initialisation failed: {lib.zk_last_error().decode()}"
    if lib.zk_selftest() != 0:
        return f"the probe kernel failed: {lib.zk_last_error().decode()}"
    return None


LIB, _load_reason = _load()
GPU_UNAVAILABLE_REASON: str | None = _load_reason if LIB is None else _detect(LIB)
GPU_AVAILABLE: bool = GPU_UNAVAILABLE_REASON is None


def warn_cpu_fallback(func_name: str, fallback: str) -> None:
    """The reference's warning helper; unused here because nothing falls back."""
    warnings.warn(f"{func_name}: GPU unavailable ({GPU_UNAVAILABLE_REASON}); {fallback}", GpuFallbackWarning, stacklevel=3)


def require_gpu(func_name: str, available: bool | None = None) -> None:
    """Raise unless the GPU path can run: this package never computes on the CPU instead."""
    ok = GPU_AVAILABLE if available is None else available
    if not ok:
        raise RuntimeError(
            f"{func_name}: GPU unavailable ({GPU_UNAVAILABLE_REASON}); zkgpu has no CPU fallback"
        )


def check(status: int, what: str = "") -> None:
    """Raise ZkError with the library's message for a non-zero status."""
    if status != 0:
        msg = LIB.zk_last_error().decode() if LIB is not None else "library not loaded"
        raise ZkError(f"{what}: {msg}" if what else msg)


def call(name: str, *args: Any) -> None:
    check(getattr(LIB, name)(*args), name)


# ----------------------------------------------------------------------
# Streams
# ----------------------------------------------------------------------

_tls = threading.local()


def current_stream() -> int:
    """The stream every op launches on: 0 (the null stream) unless inside stream()."""
    return getattr(_tls, "stream", 0)


@contextlib.contextmanager
def stream(handle: int) -> Iterator[int]:
    """Launch the ops inside this block on *handle* (from Stream)."""
    prev = current_stream()
    _tls.stream = handle
    try:
        yield handle
    finally:
        _tls.stream = prev


class Stream:
    """An owned non-blocking HIP stream (all streams here use the default priority)."""

    def __init__(self) -> None:
        h = ctypes.c_void_p()
        call("zk_stream_create", ctypes.byref(h))
        self.handle = h.value or 0

    def synchronize(self) -> None:
        call("zk_stream_sync", self.handle)

    def __del__(self) -> None:
        if LIB is not None and getattr(self, "handle", 0):
            LIB.zk_stream_destroy(self.handle)
            self.handle = 0


class Event:
    """A HIP event for ordering and timing work across streams."""

    def __init__(self) -> None:
        h = ctypes.c_void_p()
        call("zk_event_create", ctypes.byref(h))
        self.handle = h.value

    def record(self, stream_handle: int | None = None) -> Event:
        call("zk_event_record", self.handle, current_stream() if stream_handle is None else stream_handle)
        return self

    def elapsed_ms(self, end: Event) -> float:
        out = ctypes.c_float()
        call("zk_event_elapsed_ms", self.handle, end.handle, ctypes.byref(out))
        return float(out.value)

    def done(self) -> bool:
        """True once all work recorded before this event has finished (never blocks)."""
        out = ctypes.c_int()
        call("zk_event_query", self.handle, ctypes.byref(out))
        return bool(out.value)

    def synchronize(self) -> None:
        call("zk_event_sync", self.handle)

    def __del__(self) -> None:
        if LIB is not None and getattr(self, "handle", None):
            LIB.zk_event_destroy(self.handle)
            self.handle = None


def prof_enable(on: bool = True) -> None:
    """Record HIP events around the library's main kernels (see prof_read)."""
    call("zk_prof_enable", 1 if on else 0)


def prof_read(name: str, reset: bool = False) -> tuple[float, int]:
    """(total ms, launches) recorded for kernel `name` since the last reset."""
    total, count = ctypes.c_double(), ctypes.c_int64()
    call("zk_prof_read", name.encode(), ctypes.byref(total), ctypes.byref(count), 1 if reset else 0)
    return float(total.value), int(count.value)


class Graph:
    """A HIP graph recorded from a created stream; replay() re-runs the recorded launches."""

    def __init__(self, handle: int, stream_handle: int) -> None:
        self.handle = handle
        self.stream = stream_handle

    @classmethod
    @contextlib.contextmanager
    def capture(cls, on: Stream) -> Iterator[list]:
        """with Graph.capture(s) as box: ... launches on s ...; box[0] is the Graph."""
        box: list = []
        call("zk_graph_begin", on.handle)
        ok = False
        try:
            with stream(on.handle):
                yield box
            ok = True
        finally:
            h = ctypes.c_void_p()
            status = LIB.zk_graph_end(on.handle, ctypes.byref(h))
            if ok:
                check(status, "zk_graph_end")
                box.append(cls(h.value, on.handle))

    def replay(self, stream_handle: int | None = None) -> None:
        call("zk_graph_launch", self.handle, self.stream if stream_handle is None else stream_handle)

    def __del__(self) -> None:
        if LIB is not None and getattr(self, "handle", None):
            LIB.zk_graph_destroy(self.handle)
            self.handle = None


def synchronize() -> None:
    call("zk_device_sync")


def mem_info() -> tuple[int, int]:
    """(free, total) bytes of device memory."""
    free, total = ctypes.c_uint64(), ctypes.c_uint64()
    call("zk_mem_info", ctypes.byref(free), ctypes.byref(total))
    return int(free.value), int(total.value)


# ----------------------------------------------------------------------
# Device arrays
# ----------------------------------------------------------------------


class _Allocation:
    """One block of device memory, returned to the library's cache when collected."""

    __slots__ = ("ptr", "nbytes", "__weakref__")

    def __init__(self, nbytes: int) -> None:
        self.nbytes = nbytes
        self.ptr = 0
        if nbytes:
            p = ctypes.c_void_p()
            call("zk_alloc", ctypes.byref(p), max(nbytes, 1), current_stream())
            self.ptr = p.value or 0

    def __del__(self) -> None:
        if self.ptr and LIB is not None:
            LIB.zk_free(self.ptr)
            self.ptr = 0


class DeviceArray:
    """
    A C-contiguous array in device memory (the CuPy-ndarray subset the API returns).

    Views (reshape, row slices) share the allocation. Work is ordered on the
    stream that is current when it is launched; get() waits for that stream.
    """

    __slots__ = ("_alloc", "_offset", "shape", "dtype")

    def __init__(self, shape: Sequence[int] | int, dtype: Any = np.uint64, *, _alloc: _Allocation | None = None, _offset: int = 0):
        self.shape = (int(shape),) if isinstance(shape, (int, np.integer)) else tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self._alloc = _alloc if _alloc is not None else _Allocation(self.nbytes)
        self._offset = _offset

    # -- shape -----------------------------------------------------------
    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def size(self) -> int:
        out = 1
        for s in self.shape:
            out *= s
        return out

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    @property
    def ptr(self) -> int:
        """Device address of the first element (0 for an empty array)."""
        return (self._alloc.ptr + self._offset) if self._alloc.ptr else 0

    def __len__(self) -> int:
        if not self.shape:
            raise TypeError("len() of a 0-d array")
        return self.shape[0]

    def reshape(self, *shape: Any) -> DeviceArray:
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        shape = list(shape)
        if shape.count(-1) > 1:
            raise ValueError("can only specify one unknown dimension")
        if -1 in shape:
            known = 1
            for s in shape:
                if s != -1:
                    known *= s
            if known == 0 or self.size % known:
                raise ValueError(f"cannot reshape array of size {self.size} into shape {tuple(shape)}")
            shape[shape.index(-1)] = self.size // known
        new = 1
        for s in shape:
            new *= s
        if new != self.size:
            raise ValueError(f"cannot reshape array of size {self.size} into shape {tuple(shape)}")
        return DeviceArray(tuple(shape), self.dtype, _alloc=self._alloc, _offset=self._offset)

    def __getitem__(self, key: Any) -> DeviceArray:
        """Contiguous views along the first axis: a[i], a[i:j] (step 1)."""
        if not self.shape:
            raise IndexError("0-d arrays cannot be indexed")
        rows = self.shape[0]
        row_bytes = self.nbytes // rows if rows else 0
        if isinstance(key, (int, np.integer)):
            i = int(key)
            if i < 0:
                i += rows
            if not 0 <= i < rows:
                raise IndexError(f"index {key} is out of bounds for axis 0 with size {rows}")
            return DeviceArray(self.shape[1:], self.dtype, _alloc=self._alloc, _offset=self._offset + i * row_bytes)
        if isinstance(key, slice):
            start, stop, step = key.indices(rows)
            if step != 1:
                raise IndexError("only step-1 slices give contiguous device views")
            count = max(0, stop - start)
            return DeviceArray((count, *self.shape[1:]), self.dtype, _alloc=self._alloc, _offset=self._offset + start * row_bytes)
        raise IndexError(f"unsupported index {key!r}")

    # -- transfers ---------------------------------------------------------
    def get(self) -> np.ndarray:
        """A host copy (waits for the current stream)."""
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            call("zk_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes, current_stream())
        return out

    def __array__(self, dtype: Any = None, copy: Any = None) -> np.ndarray:
        host = self.get()
        return host if dtype is None else host.astype(dtype)

    def tolist(self) -> list[Any]:
        return self.get().tolist()

    def copy(self) -> DeviceArray:
        out = DeviceArray(self.shape, self.dtype)
        if self.nbytes:
            call("zk_memcpy_d2d", out.ptr, self.ptr, self.nbytes, current_stream())
        return out

    def __repr__(self) -> str:
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype.name})"


def empty(shape: Sequence[int] | int, dtype: Any = np.uint64) -> DeviceArray:
    return DeviceArray(shape, dtype)


def zeros(shape: Sequence[int] | int, dtype: Any = np.uint64) -> DeviceArray:
    out = DeviceArray(shape, dtype)
    if out.nbytes:
        call("zk_memset", out.ptr, 0, out.nbytes, current_stream())
    return out


def asarray(host: Any, dtype: Any = None) -> DeviceArray:
    """Upload a host array (a DeviceArray is returned unchanged)."""
    if isinstance(host, DeviceArray):
        return host
    arr = np.asarray(host, dtype=dtype)
    if not arr.flags.c_contiguous:
        arr = arr.copy(order="C")  # (np.ascontiguousarray would also turn a 0-d array into 1-d)
    out = DeviceArray(arr.shape, arr.dtype)
    if out.nbytes:
        call("zk_memcpy_h2d", out.ptr, arr.ctypes.data, out.nbytes, current_stream())
    return out


class PinnedBuffer:
    """
    Page-locked host staging memory (zk_host_alloc) for uploads at PCIe speed, used as a context
    manager: `view` is a writable buffer over the block, upload() enqueues an asynchronous copy to
    device memory on the current stream, and leaving the block waits for that stream before the
    block goes back to the library's cache (zk_host_free).
    """

    def __init__(self, nbytes: int) -> None:
        require_gpu("PinnedBuffer")
        p = ctypes.c_void_p()
        call("zk_host_alloc", ctypes.byref(p), max(int(nbytes), 1))
        self.ptr = p.value or 0
        self.nbytes = int(nbytes)
        self.view = (ctypes.c_char * max(self.nbytes, 1)).from_address(self.ptr)

    def upload(self, dst: DeviceArray, offset: int, nbytes: int) -> None:
        """Copy bytes [offset, offset + nbytes) of the block to the start of dst, asynchronously."""
        if offset < 0 or nbytes < 0 or offset + nbytes > self.nbytes or nbytes > dst.nbytes:
            raise ValueError("PinnedBuffer.upload: range outside the block or the destination")
        if nbytes:
            call("zk_memcpy_h2d_async", dst.ptr, self.ptr + offset, nbytes, current_stream())

    def release(self) -> None:
        if self.ptr and LIB is not None:
            call("zk_stream_sync", current_stream())
            LIB.zk_host_free(self.ptr)
        self.ptr = 0

    def __enter__(self) -> "PinnedBuffer":
        return self

    def __exit__(self, *exc: Any) -> None:
        self.release()

    def __del__(self) -> None:
        if self.ptr and LIB is not None:
            LIB.zk_stream_sync(current_stream())
            LIB.zk_host_free(self.ptr)
            self.ptr = 0


def ascontiguousarray(x: Any) -> DeviceArray:
    """Device arrays are always contiguous; host arrays are uploaded."""
    return asarray(x)


def is_device_array(x: Any) -> bool:
    return isinstance(x, DeviceArray)


def _to_host(x: Any) -> np.ndarray:
    return x.get() if isinstance(x, DeviceArray) else np.asarray(x)


def u64_array(values: Sequence[int]) -> Any:
    """A ctypes uint64[len] initialised from Python ints (for small constant arguments)."""
    arr = (ctypes.c_uint64 * len(values))(*[int(v) & (2**64 - 1) for v in values])
    return arr


def limbs_of(value: int) -> Any:
    """ctypes uint64[4] of a value below 2^256, little-endian."""
    return u64_array([(value >> (64 * j)) & (2**64 - 1) for j in range(4)])


def _apply_tuning() -> None:
    """Hand this GPU's autotuned NTT fusion depths (zkgpu/tuning/<device>.json) to the library."""
    if GPU_AVAILABLE:
        from zkgpu import autotune

        autotune.apply()


_apply_tuning()
