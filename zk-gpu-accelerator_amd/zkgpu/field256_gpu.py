"""
zkgpu.field256_gpu: 254-bit prime-field arithmetic and the NTT on the MI355X.

The reference's Field256 API (src/zkgpu/field256_gpu.py:269-512), backed by
hand-written gfx950 kernels in libzkgpu_hip.so instead of NVRTC-compiled
CUDA strings.

Representation (unchanged from the reference, :12-20): a batch of n
elements is an (n, 4) uint64 DeviceArray of a * 2^256 mod p, least
significant limb first. Inside the kernels each element is nine 29-bit limbs
(so a Montgomery product needs no carry handling inside a column); rows in
HBM keep the 4 x 64-bit layout, so arrays move between this package and any
code that reads the reference's format.

Contract (as the reference's): the modulus is odd and below 2^255; rows are
canonical; a single-row operand broadcasts. Without a usable GPU every
method raises RuntimeError with the reason: there is no CPU fallback.

Field256.ntt is the reference's transform value for value (ntt.ntt, natural
order in and out, omega from the smallest quadratic non-residue, the inverse
scaled by n^-1), computed by multi-stage LDS passes (see csrc/zk_ntt.hip).
"""

from __future__ import annotations

import ctypes
import operator
from collections.abc import Iterable, Sequence
from pathlib import Path
from typing import Any

import numpy as np

from zkgpu import gpu_backend as _gb
from zkgpu.curve import P as BN254_BASE_MODULUS
from zkgpu.field import FIELD_MODULUS as BN254_SCALAR_MODULUS
from zkgpu.gpu_backend import GPU_AVAILABLE as _gpu_available, DeviceArray
from zkgpu.ntt import get_ntt_root

LIMBS = 4
_LIMB_BITS = 64
_R_BITS = LIMBS * _LIMB_BITS  # R = 2^256
_CSRC = Path(__file__).resolve().parent.parent / "csrc"


try:  # bulk int <-> row conversion in C (csrc/zk_ingest.c); host bookkeeping only
    from zkgpu import _ingest
except ImportError:  # not built: the same conversions one element at a time
    _ingest = None


def _ints_to_limbs(values: Iterable[int]) -> np.ndarray:
    """(n, 4) uint64 little-endian limbs of values in [0, 2^256)."""
    if _ingest is not None:
        buf, _ = _ingest.rows_from_ints(values, 2**_R_BITS, "")
    else:
        buf = b"".join(v.to_bytes(32, "little") for v in values)
    return np.frombuffer(buf, dtype="<u8").astype(np.uint64).reshape(-1, LIMBS)


def _rows_from_ints(values: Iterable[Any], modulus: int, name: str = "") -> tuple[np.ndarray, Sequence[Any]]:
    """(n, 4) uint64 rows of integers (anything operator.index accepts) and the sequence they were
    read from. Values in [0, 2^256) are kept as they are (the device reduces them); others are
    reduced mod `modulus` here. A non-integer raises TypeError: operator.index's, or with a name,
    "<name> must contain integers: ..." (msm_gpu.py:385-390)."""
    if _ingest is not None:
        buf, items = _ingest.rows_from_ints(values, modulus, name)
        return np.frombuffer(buf, dtype="<u8").reshape(-1, LIMBS), items
    items = values if isinstance(values, Sequence) else list(values)
    try:
        ints = [operator.index(v) for v in items]
    except TypeError as exc:
        if not name:
            raise
        raise TypeError(f"{name} must contain integers: {exc}") from None
    kept = [v if 0 <= v < 2**_R_BITS else v % modulus for v in ints]
    return _ints_to_limbs(kept), items


def _limbs_to_ints(limbs: np.ndarray) -> list[int]:
    if _ingest is not None:
        return _ingest.ints_from_rows(np.ascontiguousarray(limbs, dtype="<u8"))
    buf = np.ascontiguousarray(limbs, dtype="<u8").tobytes()
    return [int.from_bytes(buf[k : k + 32], "little") for k in range(0, len(buf), 32)]


def _row_count(x: Any, name: str) -> int:
    """Rows of x after checking it is an (n, 4) uint64 array (field256_gpu.py:250-257)."""
    if getattr(x, "dtype", None) != np.uint64 or getattr(x, "ndim", None) != 2 or x.shape[1] != LIMBS:
        raise TypeError(
            f"{name} must be an (n, {LIMBS}) uint64 array from Field256.from_ints, "
            f"got {getattr(x, 'dtype', type(x).__name__)} with shape {getattr(x, 'shape', None)}"
        )
    return int(x.shape[0])


def _broadcast_rows(a: Any, b: Any) -> int:
    rows_a, rows_b = _row_count(a, "a"), _row_count(b, "b")
    if rows_a == rows_b or rows_b == 1:
        return rows_a
    if rows_a == 1:
        return rows_b
    raise ValueError(f"cannot broadcast {rows_a} rows against {rows_b} rows")


def _stream() -> int:
    return _gb.current_stream()


class Field256:
    """
    Arithmetic modulo an odd modulus below 2^255 on (n, 4) uint64 rows.

    BN254's r and p use kernels with the modulus compiled in; any other
    modulus is registered with the library on first use. The kernels are
    self-tested against Python integers before the first result is returned.
    """

    def __init__(self, modulus: int) -> None:
        modulus = operator.index(modulus)
        if not (3 <= modulus < 2 ** (_R_BITS - 1) and modulus % 2 == 1):
            raise ValueError(f"modulus must be odd and in [3, 2^255), got {modulus}")
        self.modulus = modulus
        self._r2 = pow(2, 2 * _R_BITS, modulus)
        self._r_inv = pow(2**_R_BITS, -1, modulus)
        self._n_inv = -pow(modulus, -1, 2**_LIMB_BITS) % 2**_LIMB_BITS
        self._kernels: dict[str, Any] | None = None

    def __repr__(self) -> str:
        return f"Field256({self.modulus:#x})"

    # ------------------------------------------------------------------
    # Conversion
    # ------------------------------------------------------------------

    def from_ints(self, values: Iterable[int]) -> DeviceArray:
        """Montgomery-form rows for any ints (reduced mod the modulus: negative values and those of
        2^256 and more on the host, the rest by the conversion kernel)."""
        rows, _ = _rows_from_ints(values, self.modulus)
        _gb.require_gpu("Field256.from_ints", _gpu_available)
        fid = self._compile()["id"]
        return self._to_mont(fid, _gb.asarray(np.ascontiguousarray(rows, dtype=np.uint64)))

    def from_limbs(self, plain: Any) -> DeviceArray:
        """Montgomery-form rows for plain (n, 4) uint64 limbs (host or device), each < 2^256."""
        _row_count(plain, "plain")
        _gb.require_gpu("Field256.from_limbs", _gpu_available)
        return self._to_mont(self._compile()["id"], _gb.asarray(plain))

    def _to_mont(self, fid: int, plain: DeviceArray) -> DeviceArray:
        n = plain.shape[0]
        out = _gb.empty((n, LIMBS))
        if n:
            _gb.call("zk_field_to_mont", fid, plain.ptr, out.ptr, n, _stream())
        return out

    def to_limbs(self, x: Any) -> DeviceArray:
        """Plain canonical values of Montgomery-form rows, as (n, 4) uint64 device rows."""
        n = _row_count(x, "x")
        _gb.require_gpu("Field256.to_limbs", _gpu_available)
        fid = self._compile()["id"]
        src = _gb.asarray(x)
        out = _gb.empty((n, LIMBS))
        if n:
            _gb.call("zk_field_from_mont", fid, src.ptr, out.ptr, n, _stream())
        return out

    def to_ints(self, x: Any) -> list[int]:
        """The values of Montgomery-form rows, as Python ints in [0, modulus)."""
        return _limbs_to_ints(self.to_limbs(x).get())

    # ------------------------------------------------------------------
    # Arithmetic
    # ------------------------------------------------------------------

    def mul(self, a: Any, b: Any) -> DeviceArray:
        """Element-wise product of Montgomery-form rows (either may be one broadcast row)."""
        return self._launch("zk_field_mul", a, b)

    def add(self, a: Any, b: Any) -> DeviceArray:
        return self._launch("zk_field_add", a, b)

    def sub(self, a: Any, b: Any) -> DeviceArray:
        """Element-wise a - b."""
        return self._launch("zk_field_sub", a, b)

    # ------------------------------------------------------------------
    # NTT
    # ------------------------------------------------------------------

    def ntt(self, x: Any, *, invert: bool = False) -> DeviceArray:
        """
        Radix-2 NTT of Montgomery-form rows x; returns new rows, x is unchanged.

        Exactly zkgpu.ntt's transform (and Radix2Domain.fft / ifft for
        BN254_FR). Raises ValueError if the row count is not a power of two,
        the modulus is not an odd prime, or the size does not divide modulus - 1.
        """
        n = _row_count(x, "x")
        if n == 0:
            _gb.require_gpu("Field256.ntt", _gpu_available)
            return _gb.empty((0, LIMBS))
        if n & (n - 1):
            raise ValueError(f"NTT length must be a power of 2, got {n}")
        omega = get_ntt_root(n, self.modulus)  # validates the modulus and the size first
        _gb.require_gpu("Field256.ntt", _gpu_available)
        if invert:
            omega = pow(omega, -1, self.modulus)
        return self._ntt(_gb.asarray(x), n, omega, invert, None)

    def coset_ntt(self, x: Any, coset: int, *, invert: bool = False) -> DeviceArray:
        """
        Forward: NTT of x_j * g^j (evaluations on the coset g*H). Inverse: the
        inverse NTT followed by out_j * g^-j. Radix2Domain.coset_fft / coset_ifft
        (src/zkgpu/domain.py:122-135) with the scaling fused into the transform.
        """
        n = _row_count(x, "x")
        if n == 0:
            _gb.require_gpu("Field256.coset_ntt", _gpu_available)
            return _gb.empty((0, LIMBS))
        if n & (n - 1):
            raise ValueError(f"NTT length must be a power of 2, got {n}")
        omega = get_ntt_root(n, self.modulus)
        _gb.require_gpu("Field256.coset_ntt", _gpu_available)
        if invert:
            omega = pow(omega, -1, self.modulus)
        return self._ntt(_gb.asarray(x), n, omega, invert, coset % self.modulus)

    def _ntt(self, src: DeviceArray, n: int, omega: int, invert: bool, coset: int | None) -> DeviceArray:
        fid = self._compile()["id"]
        out = _gb.empty((n, LIMBS))
        coset_arg = _gb.limbs_of(coset) if coset is not None else None
        _gb.call("zk_ntt", fid, src.ptr, out.ptr, n.bit_length() - 1, 1 if invert else 0, _gb.limbs_of(omega), coset_arg, _stream())
        return out

    # ------------------------------------------------------------------
    # Kernel set
    # ------------------------------------------------------------------

    def _constant(self, value: int) -> DeviceArray:
        """One device row holding value as plain limbs."""
        return _gb.asarray(_ints_to_limbs([value]))

    def cuda_source(self) -> str:
        """
        The kernels' field source for this modulus: its constants as #defines
        (ZK_P0..3, ZK_ONE0..3 = 2^256 mod p, ZK_NINV = -p^-1 mod 2^64, the
        reference's names, field256_gpu.py:397-413), then the HIP device code.
        """

        def limbs(value: int) -> list[int]:
            return [(value >> (_LIMB_BITS * j)) & (2**_LIMB_BITS - 1) for j in range(LIMBS)]

        defines = "".join(f"#define ZK_P{j} {limb:#x}ULL\n" for j, limb in enumerate(limbs(self.modulus)))
        one = 2**_R_BITS % self.modulus
        defines += "".join(f"#define ZK_ONE{j} {limb:#x}ULL\n" for j, limb in enumerate(limbs(one)))
        defines += f"#define ZK_NINV {self._n_inv:#x}ULL\n"
        body = "".join((_CSRC / name).read_text() for name in ("zk_constants.hpp", "zk_field.hpp"))
        return defines + body

    hip_source = cuda_source

    def compile(self) -> None:
        """Load and self-test the kernels now instead of on first use (RuntimeError if they fail)."""
        _gb.require_gpu("Field256.compile", _gpu_available)
        self._compile()

    def _field_id(self) -> int:
        if self.modulus == BN254_SCALAR_MODULUS:
            return 0
        if self.modulus == BN254_BASE_MODULUS:
            return 1
        fid = ctypes.c_int(-1)
        _gb.call("zk_field_register", _gb.limbs_of(self.modulus), ctypes.byref(fid))
        return int(fid.value)

    def _compile(self) -> dict[str, Any]:
        if self._kernels is None:
            self._kernels = {"id": self._field_id()}
            try:
                self._self_test()
            except Exception:
                self._kernels = None  # never keep kernels that failed their self-test
                raise
        return self._kernels

    def _self_test(self) -> None:
        """Carry-heavy operands (0, 1, p - 1, all-ones limbs) against Python ints, once per field."""
        p = self.modulus
        a = [0, 1, 2, p - 1, p - 2, p // 2, (2**128 - 1) % p, (2**192 - 1) % p]
        b = [p - 1, p - 1, 3, p - 1, 5, p // 3, (2**64 - 1) % p, 2**254 % p]
        x, y = self.from_ints(a), self.from_ints(b)
        checks = {
            "mul": (self.mul(x, y), [u * v % p for u, v in zip(a, b)]),
            "add": (self.add(x, y), [(u + v) % p for u, v in zip(a, b)]),
            "sub": (self.sub(x, y), [(u - v) % p for u, v in zip(a, b)]),
        }
        try:
            w = get_ntt_root(len(a), p)
        except ValueError:
            pass  # no NTT of this size modulo p
        else:
            dft = [sum(v * pow(w, i * k, p) for i, v in enumerate(a)) % p for k in range(len(a))]
            checks["ntt"] = (self.ntt(x), dft)
        for op, (rows, expected) in checks.items():
            if self.to_ints(rows) != expected:
                raise RuntimeError(
                    f"{self!r}: the {op} kernel failed its self-test: libzkgpu_hip.so computed wrong "
                    "field elements on this GPU"
                )

    def _launch(self, name: str, a: Any, b: Any) -> DeviceArray:
        n = _broadcast_rows(a, b)
        _gb.require_gpu(f"Field256.{name.rsplit('_', 1)[-1]}", _gpu_available)
        fid = self._compile()["id"]
        a, b = _gb.asarray(a), _gb.asarray(b)
        out = _gb.empty((n, LIMBS))
        if n:
            _gb.call(name, fid, a.ptr, 1 if a.shape[0] == 1 else 0, b.ptr, 1 if b.shape[0] == 1 else 0, out.ptr, n, _stream())
        return out


class _NttTables:
    """
    The reference's _ntt_tables (field256_gpu.py:492-508) as a cache handle.

    Twiddles live in libzkgpu_hip.so's own cache (per field, size, root and
    coset); calling this returns the reference's table layout for inspection,
    and cache_clear() empties the library cache so the next transform
    rebuilds its tables.
    """

    def __call__(self, field: Field256, n: int, invert: bool) -> tuple[DeviceArray, DeviceArray]:
        x = np.zeros((n, LIMBS), dtype=np.uint64)
        if n > 1:
            x[1] = _ints_to_limbs([2**_R_BITS % field.modulus])[0]  # the polynomial X in Montgomery form
        powers = field.ntt(x, invert=False)
        if invert:
            inv_powers = field.ntt(x, invert=True)
            powers = field.mul(inv_powers, field.from_ints([n]))  # undo the n^-1 scale: omega^-k
        half = max(n // 2, 1)
        return powers[:half], field.from_ints([pow(n, -1, field.modulus)])

    def cache_clear(self) -> None:
        if _gb.LIB is not None:
            _gb.call("zk_ntt_cache_clear")


_ntt_tables = _NttTables()

BN254_FR = Field256(BN254_SCALAR_MODULUS)  # scalar field r: NTT and polynomial arithmetic
BN254_FQ = Field256(BN254_BASE_MODULUS)  # base field p: curve point coordinates
