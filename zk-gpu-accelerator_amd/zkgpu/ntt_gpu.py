"""
zkgpu.ntt_gpu: the number theoretic transform of int64 vectors on the device.

The reference's module of the same name (src/zkgpu/ntt_gpu.py:1-155): ntt_gpu(a, mod, invert)
for an odd prime mod < 2^63 with len(a) a power of two dividing mod - 1, the root from
zkgpu.ntt.get_ntt_root (the CPU NTT's, so both transforms agree element for element), inputs of
any values reduced mod *mod* first, the input left unchanged, and the CPU NTT's validation
errors.

The transform is zk_ntt64 in libzkgpu_hip.so: a multi-pass decomposition with at most 2^10
points per pass, each pass one HBM round trip through LDS (two passes at 2^20, where the
reference launches five CuPy kernels per stage). Its twiddle tables are built once per
(size, modulus, direction) and cached on the device with the 256-bit NTT's tables;
_twiddle_table.cache_clear() frees the ones no live plan holds. There is no CPU fallback:
without a usable GPU, ntt_gpu raises RuntimeError.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from zkgpu import gpu_backend as _gb
from zkgpu.field_gpu import _DTYPE_CODES
from zkgpu.ntt import get_ntt_root

_INT64_MAX = 2**63 - 1


class _TwiddleTables:
    """Handle on the cached device tables, standing in for the reference's lru_cache
    (ntt_gpu.py:65-79): cache_clear() frees every table no live plan holds."""

    def cache_clear(self) -> None:
        if _gb.GPU_AVAILABLE:
            _gb.call("zk_ntt_cache_clear")


_twiddle_table = _TwiddleTables()


def ntt_gpu(a: Any, mod: int, invert: bool = False) -> _gb.DeviceArray:
    """
    NTT on the GPU: X_k = sum_j a_j omega^(jk) mod *mod* (inverse: omega^-1 and a factor n^-1).

    Parameters
    ----------
    a      : 1-D array (a DeviceArray of an integer dtype, or host values converted as
             numpy.asarray(a, dtype=int64)), length a power of 2; any values, reduced mod *mod*
    mod    : odd prime below 2^63 with len(a) dividing mod - 1
    invert : if True, compute the inverse NTT

    Returns
    -------
    int64 DeviceArray of the same length with values in [0, mod)

    Raises
    ------
    ValueError   : mod >= 2^63, or the same length/modulus errors as zkgpu.ntt.ntt
    TypeError    : a device array of a non-integer dtype
    RuntimeError : if no GPU is usable
    """
    if mod > _INT64_MAX:
        raise ValueError(
            f"ntt_gpu works on int64 elements, so mod must be below 2^63, got {mod}; "
            "use zkgpu.ntt.ntt for larger moduli such as BN254's r"
        )
    _gb.require_gpu("ntt_gpu")
    if isinstance(a, _gb.DeviceArray):
        if a.dtype.kind not in "iu":
            raise TypeError(f"a must be an integer array, got dtype {a.dtype}")
        x = a
    else:
        x = _gb.asarray(np.asarray(a, dtype=np.int64))
    if x.ndim != 1:
        raise ValueError(f"ntt_gpu transforms a 1-D array, got shape {x.shape}")
    n = x.shape[0]
    if n & (n - 1):  # lets n == 0 through
        raise ValueError(f"NTT length must be a power of 2, got {n}")
    out = _gb.empty(n, np.int64)
    if n == 0:
        return out
    omega = get_ntt_root(n, mod)  # validates mod and n as the CPU NTT does
    _gb.call(
        "zk_ntt64", x.ptr, _DTYPE_CODES[x.dtype], out.ptr, n.bit_length() - 1, mod, omega, 1 if invert else 0,
        _gb.current_stream(),
    )
    return out
