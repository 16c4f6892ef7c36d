"""
ctypes binding of the C restatement (oracle/c/zkref.c) (test infrastructure only).

Plain (non-Montgomery) canonical values cross this boundary as (n, 4)
little-endian uint64 rows. Build with `make -C oracle`.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

from zkref.field import R_SCALAR, ints_to_limbs, limbs_to_ints
from zkref.ntt import get_ntt_root

_LIB_PATH = Path(__file__).resolve().parent.parent / "lib" / "libzkref.so"
_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library, its signatures declared before it is shared (callers may be threads)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"{_LIB_PATH} is not built; run `make -C oracle`")
        dll = ctypes.CDLL(str(_LIB_PATH))
        u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
        L = ctypes.c_long
        dll.zr_ntt.argtypes = [u64p, u64p, ctypes.c_int, u64p, ctypes.c_void_p]
        dll.zr_scale_by_powers.argtypes = [u64p, u64p, L, u64p]
        dll.zr_pointwise_quotient.argtypes = [u64p, u64p, u64p, u64p, u64p, L]
        dll.zr_g1_op.argtypes = [ctypes.c_int, u64p, L, u64p, L, u64p, L]
        dll.zr_msm_pippenger.argtypes = [u64p, u64p, L, ctypes.c_int, u64p, ctypes.POINTER(ctypes.c_int)]
        dll.zr_fixed_base_g1.argtypes = [u64p, u64p, L, ctypes.c_int]
        dll.zr_dot_mod_r.argtypes = [u64p, u64p, L, u64p]
        _lib = dll
    return _lib


def _limbs(v) -> np.ndarray:
    return np.ascontiguousarray(v, dtype=np.uint64)


def ntt_rows(values: np.ndarray, invert: bool = False) -> np.ndarray:
    """ntt(values, r, invert) on plain (n, 4) rows; values must be canonical."""
    values = _limbs(values).reshape(-1, 4)
    n = values.shape[0]
    log_n = n.bit_length() - 1
    w = get_ntt_root(n, R_SCALAR)
    scale = None
    if invert:
        w = pow(w, -1, R_SCALAR)
        scale = ints_to_limbs([pow(n, -1, R_SCALAR)])
    out = np.empty_like(values)
    lib().zr_ntt(values, out, log_n, ints_to_limbs([w]), scale.ctypes.data if scale is not None else None)
    return out


def scale_by_powers(values: np.ndarray, g: int) -> np.ndarray:
    values = _limbs(values).reshape(-1, 4)
    out = np.empty_like(values)
    lib().zr_scale_by_powers(values, out, values.shape[0], ints_to_limbs([g % R_SCALAR]))
    return out


def groth16_quotient(a: np.ndarray, b: np.ndarray, c: np.ndarray, shift: int) -> np.ndarray:
    """zkref.domain.groth16_quotient on plain rows."""
    n = _limbs(a).reshape(-1, 4).shape[0]
    z_inv = pow((pow(shift, n, R_SCALAR) - 1) % R_SCALAR, -1, R_SCALAR)
    g_inv = pow(shift, -1, R_SCALAR)
    cos = [ntt_rows(scale_by_powers(ntt_rows(v, invert=True), shift)) for v in (a, b, c)]
    num = np.empty_like(cos[0])
    lib().zr_pointwise_quotient(cos[0], cos[1], cos[2], ints_to_limbs([z_inv]), num, n)
    return scale_by_powers(ntt_rows(num, invert=True), g_inv)


def g1_op(op: str, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """Element-wise add / add_affine / double on plain Jacobian (n,3,4) rows (b broadcasts when one row)."""
    a = _limbs(a).reshape(-1, 12)
    code = {"add": 0, "add_affine": 1, "double": 2}[op]
    width = 12 if op == "add" else 8
    if b is None:
        b = np.zeros((1, width), dtype=np.uint64)
    b = _limbs(b).reshape(-1, width)
    n = max(a.shape[0], b.shape[0])
    out = np.empty((n, 12), dtype=np.uint64)
    lib().zr_g1_op(code, a, 12 if a.shape[0] > 1 else 0, b, width if b.shape[0] > 1 else 0, out, n)
    return out.reshape(n, 3, 4)


def msm_pippenger(points_xy: np.ndarray, scalars: np.ndarray, w: int) -> tuple[int, int] | None:
    """sum s_i P_i for plain affine points (n,2,4) and reduced scalars (n,4); None for the identity."""
    pts = _limbs(points_xy).reshape(-1, 8)
    sc = _limbs(scalars).reshape(-1, 4)
    out = np.zeros(8, dtype=np.uint64)
    fin = ctypes.c_int(0)
    lib().zr_msm_pippenger(pts, sc, pts.shape[0], w, out, ctypes.byref(fin))
    if not fin.value:
        return None
    x, y = limbs_to_ints(out.reshape(2, 4))
    return x, y


def fixed_base_g1(ks: np.ndarray, threads: int | None = None) -> np.ndarray:
    """Affine k_i * G as plain (n, 2, 4) rows (all-zero rows for the identity)."""
    ks = _limbs(ks).reshape(-1, 4)
    out = np.empty((ks.shape[0], 8), dtype=np.uint64)
    t = threads or min(16, len(os.sched_getaffinity(0)))
    lib().zr_fixed_base_g1(ks, out, ks.shape[0], t)
    return out.reshape(-1, 2, 4)


def dot_mod_r(keys: np.ndarray, scalars: np.ndarray) -> int:
    """sum k_i s_i mod r over plain (n, 4) rows (any values below 2^256): for a basis P_i = k_i G
    the MSM of the scalars s_i is this multiple of G."""
    k = _limbs(keys).reshape(-1, 4)
    s = _limbs(scalars).reshape(-1, 4)
    if k.shape != s.shape:
        raise ValueError("keys and scalars must have the same number of rows")
    out = np.zeros(4, dtype=np.uint64)
    lib().zr_dot_mod_r(k, s, k.shape[0], out)
    return limbs_to_ints(out.reshape(1, 4))[0]
