"""
zkgpu.field_gpu: elementwise arithmetic over 64-bit prime fields on the device.

The reference's module of the same name (src/zkgpu/field_gpu.py:1-230), with its contract:

- the modulus must satisfy 2 <= mod < 2^63 (ValueError otherwise, TypeError if it is not an
  integer: operator.index, so 7.0 is refused rather than truncated);
- inputs are arrays of any integer dtype (TypeError "<name> must be an integer array, got dtype
  ..." for others), widened to int64 as astype(int64) does and reduced into [0, mod) first, so
  negative and unreduced values give field-correct results;
- operands broadcast against each other (NumPy rules);
- results are int64 device arrays with every element in [0, mod);
- a negative exponent in gpu_batch_pow raises the inverse to |exp| through Fermat's little
  theorem (so the modulus must be prime), and a base that is 0 mod mod then raises
  ValueError("base is not invertible for the given modulus"), as Python's pow() does.

Operands may be zkgpu DeviceArrays or host arrays (uploaded). The work runs in libzkgpu_hip.so:
zk_f64_binary (one Barrett-reduced kernel for every modulus below 2^63, broadcasting by
strides), zk_f64_pow (the whole square-and-multiply chain in one launch per 4096 exponent bits,
where the reference launches two kernels per bit) and zk_f64_any_zero. There is no CPU
fallback: without a usable GPU every call raises RuntimeError.
"""

from __future__ import annotations

import ctypes
import math
import operator
from typing import Any

import numpy as np

from zkgpu import gpu_backend as _gb

_INT64_MAX = 2**63 - 1

# The largest modulus whose residues multiply without overflowing int64, (mod - 1)^2 <= 2^63 - 1:
# where the reference switches to its wide product. Every modulus takes the same path here.
_INT64_PRODUCT_SAFE_MOD = math.isqrt(_INT64_MAX) + 1  # 3_037_000_500

# dtype codes of include/zkgpu.h (ZK_I8 ... ZK_U64)
_DTYPE_CODES = {
    np.dtype(np.int8): 0,
    np.dtype(np.int16): 1,
    np.dtype(np.int32): 2,
    np.dtype(np.int64): 3,
    np.dtype(np.uint8): 4,
    np.dtype(np.uint16): 5,
    np.dtype(np.uint32): 6,
    np.dtype(np.uint64): 7,
}
_MAX_DIMS = 8


def _check_modulus(mod: Any) -> int:
    """mod as an int, or an error if it cannot be an int64 field modulus (field_gpu.py:107-112)."""
    mod = operator.index(mod)
    if not 2 <= mod <= _INT64_MAX:
        raise ValueError(f"mod must satisfy 2 <= mod < 2^63 to fit in int64, got {mod}")
    return mod


def _check_integer_array(x: Any, name: str) -> None:
    """Refuse non-integer arrays, which an int64 kernel would truncate (field_gpu.py:115-119)."""
    dtype = x.dtype if hasattr(x, "dtype") else np.asarray(x).dtype
    if dtype.kind not in "iu":
        raise TypeError(f"{name} must be an integer array, got dtype {dtype}")


def _on_device(x: Any) -> _gb.DeviceArray:
    dev = x if isinstance(x, _gb.DeviceArray) else _gb.asarray(np.asarray(x))
    if dev.dtype not in _DTYPE_CODES:
        raise TypeError(f"unsupported integer dtype {dev.dtype}")
    return dev


def _strides_for(shape: tuple[int, ...], out: tuple[int, ...]) -> list[int]:
    """Element strides of a contiguous array of `shape` read at each index of `out` (0 where broadcast)."""
    full = (1,) * (len(out) - len(shape)) + tuple(shape)
    strides = [0] * len(out)
    step = 1
    for d in range(len(out) - 1, -1, -1):
        strides[d] = 0 if full[d] == 1 else step
        step *= full[d]
    return strides


def _launch_shape(out: tuple[int, ...], operands: list[_gb.DeviceArray]) -> tuple[list[int], list[list[int]]]:
    """The output shape and per-operand strides with unit dimensions dropped and adjacent
    dimensions merged where every operand steps through them as one, at most 8 left."""
    strides = [_strides_for(x.shape, out) for x in operands]
    dims = [d for d in range(len(out)) if out[d] != 1]
    shape = [out[d] for d in dims]
    per = [[s[d] for d in dims] for s in strides]
    i = len(shape) - 2
    while i >= 0:
        if all(p[i] == p[i + 1] * shape[i + 1] for p in per):
            shape[i] *= shape[i + 1]
            del shape[i + 1]
            for p in per:
                p[i] = p.pop(i + 1)  # the merged dimension steps like the inner one
        i -= 1
    if not shape:
        return [1], [[0] for _ in operands]
    if len(shape) > _MAX_DIMS:
        raise ValueError(f"broadcasting these operands needs {len(shape)} dimensions; at most {_MAX_DIMS} are supported")
    return shape, per


def _i64(values: list[int]) -> Any:
    return (ctypes.c_int64 * len(values))(*values)


def _binary(op: int, a: Any, b: Any, mod: int, func: str) -> _gb.DeviceArray:
    _gb.require_gpu(func)
    da, db = _on_device(a), _on_device(b)
    out_shape = np.broadcast_shapes(da.shape, db.shape)
    out = _gb.empty(out_shape, np.int64)
    if out.size == 0:
        return out
    shape, (sa, sb) = _launch_shape(out_shape, [da, db])
    _gb.call(
        "zk_f64_binary", op, da.ptr, _DTYPE_CODES[da.dtype], _i64(sa), db.ptr, _DTYPE_CODES[db.dtype], _i64(sb),
        len(shape), _i64(shape), out.ptr, mod, _gb.current_stream(),
    )
    return out


def gpu_batch_mul(a: Any, b: Any, mod: int) -> _gb.DeviceArray:
    """
    Element-wise modular multiplication on the GPU: result[i] = (a[i] * b[i]) % mod.

    Parameters
    ----------
    a, b : integer arrays (any values; reduced mod *mod* first), broadcast against each other
    mod  : modulus with 2 <= mod < 2^63

    Returns
    -------
    int64 DeviceArray with every element in [0, mod)

    Raises
    ------
    ValueError   : if *mod* is outside [2, 2^63)
    TypeError    : if *mod* is not an integer or an input is not an integer array
    RuntimeError : if no GPU is usable
    """
    mod = _check_modulus(mod)
    _check_integer_array(a, "a")
    _check_integer_array(b, "b")
    return _binary(0, a, b, mod, "gpu_batch_mul")


def gpu_batch_add(a: Any, b: Any, mod: int) -> _gb.DeviceArray:
    """
    Element-wise modular addition on the GPU: result[i] = (a[i] + b[i]) % mod.

    Same contract as gpu_batch_mul.
    """
    mod = _check_modulus(mod)
    _check_integer_array(a, "a")
    _check_integer_array(b, "b")
    return _binary(1, a, b, mod, "gpu_batch_add")


def gpu_batch_pow(bases: Any, exp: int, mod: int) -> _gb.DeviceArray:
    """
    Element-wise modular exponentiation: result[i] = bases[i]^exp % mod.

    A negative *exp* raises the inverse to |exp|, the inverse taken by Fermat's little theorem
    (a^-1 = a^(mod-2)), so *mod* must be prime; exp = 0 gives 1 everywhere.

    Raises
    ------
    ValueError   : if exp < 0 and some base is 0 mod *mod* (zero has no inverse), as Python's
                   pow() does; also for a modulus outside [2, 2^63)
    TypeError    : if *exp* or *mod* is not an integer or *bases* is not an integer array
    RuntimeError : if no GPU is usable
    """
    mod = _check_modulus(mod)
    exp = operator.index(exp)
    _check_integer_array(bases, "bases")
    _gb.require_gpu("gpu_batch_pow")
    base = _on_device(bases)
    out = _gb.empty(base.shape, np.int64)
    if out.size == 0:
        return out
    shape, (sb,) = _launch_shape(base.shape, [base])
    code, strides, dims = _DTYPE_CODES[base.dtype], _i64(sb), _i64(shape)
    if exp < 0:
        zero = ctypes.c_int(0)
        _gb.call("zk_f64_any_zero", base.ptr, code, strides, len(shape), dims, mod, ctypes.byref(zero), _gb.current_stream())
        if zero.value:
            raise ValueError("base is not invertible for the given modulus")
        exp = (mod - 2) * -exp  # (b^(mod-2))^|exp|
    words = [(exp >> (64 * i)) & (2**64 - 1) for i in range(max(1, (exp.bit_length() + 63) // 64))]
    _gb.call(
        "zk_f64_pow", base.ptr, code, strides, len(shape), dims, _gb.u64_array(words), len(words), out.ptr, mod,
        _gb.current_stream(),
    )
    return out
