"""
Pippenger parameters and argument checks (host logic of msm_pippenger_gpu).

The reference's window cost model and bounds (src/zkgpu/msm.py:35-129):
w = argmin over 1..16 of ceil(254 / w) * (n + 2^(w+1)). The GPU path honours
an explicit window_size; with none it picks its own (signed digits halve the
bucket count, so the device cost model differs, see zkgpu.msm_gpu).
"""

from __future__ import annotations

import operator

from zkgpu.field import FIELD_MODULUS

SCALAR_BITS = FIELD_MODULUS.bit_length()  # 254
MAX_WINDOW_SIZE = 16


def pippenger_window_size(n: int) -> int:
    n = operator.index(n)
    costs = {w: -(-SCALAR_BITS // w) * (n + (2 << w)) for w in range(1, MAX_WINDOW_SIZE + 1)}
    return min(costs, key=lambda w: (costs[w], w))


def _check_window_size(window_size: int) -> int:
    w = operator.index(window_size)
    if w < 1 or w > MAX_WINDOW_SIZE:
        raise ValueError(f"window_size must be in 1..{MAX_WINDOW_SIZE}, got {w}")
    return w
