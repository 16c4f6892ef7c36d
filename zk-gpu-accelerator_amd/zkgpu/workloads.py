"""
Seeded synthetic inputs for the benchmark configurations (no datasets exist for these).

- ntt_coefficients: uniform Fr values from random.Random(seed) (the 2^12 NTT
  case uses seed 0x5EED01).
- srs_keys / srs_points: a basis P_i = k_i * G with k_i uniform in [1, r); the
  points are computed on the GPU by zk_g1_fixed_base.
- witness_scalars: the witness-like mix per index, u = rng.random():
  u < 0.5 uniform in [0, r), u < 0.8 uniform in {0, 1}, else uniform below 2^32.
- witness_rows_fast: the same mix drawn with numpy (device pools of the
  served workload).
- log_uniform_sizes: sizes 2^U with U uniform in [lo, hi] for the batched and
  served workloads.
- batched_workload / batched_msm_scalars / batched_ntt_rows: the batched form's
  jobs (64 MSMs with sizes log-uniform in [2^10, 2^20] over prefixes of the
  2^20-point basis with seed 0x5EED04, witness-like scalars with a seed per
  MSM; 32 forward NTTs of 2^k rows, k uniform in {10..20}).
All generators return plain little-endian (n, 4) uint64 limbs unless noted.
"""

from __future__ import annotations

import math
import random

import numpy as np

from zkgpu.field import FIELD_MODULUS as R


def _limbs(values: list[int]) -> np.ndarray:
    buf = b"".join(v.to_bytes(32, "little") for v in values)
    return np.frombuffer(buf, dtype="<u8").astype(np.uint64).reshape(-1, 4)


def ntt_coefficients(n: int, seed: int = 0x5EED01) -> list[int]:
    rng = random.Random(seed)
    return [rng.randrange(R) for _ in range(n)]


def uniform_rows(n: int, seed: int) -> np.ndarray:
    """n uniform Fr values as plain rows (fast numpy generator, rejection-sampled below r)."""
    gen = np.random.default_rng(seed)
    out = np.empty((n, 4), dtype=np.uint64)
    filled = 0
    r_top = R >> 192
    while filled < n:
        need = n - filled
        cand = gen.integers(0, 2**64, size=(need + need // 3 + 16, 4), dtype=np.uint64, endpoint=False)
        cand[:, 3] &= np.uint64((1 << 62) - 1)
        ok = cand[:, 3] < np.uint64(r_top)  # top limb strictly below r's keeps the value below r
        take = cand[ok][:need]
        out[filled : filled + len(take)] = take
        filled += len(take)
    return out


def srs_keys(n: int, seed: int = 0x5EED02) -> list[int]:
    rng = random.Random(seed)
    return [rng.randrange(1, R) for _ in range(n)]


def witness_scalars(n: int, seed: int = 0x5EED02) -> list[int]:
    rng = random.Random(seed + 1)
    out = []
    for _ in range(n):
        u = rng.random()
        if u < 0.5:
            out.append(rng.randrange(R))
        elif u < 0.8:
            out.append(rng.randrange(2))
        else:
            out.append(rng.randrange(1 << 32))
    return out


def witness_rows(n: int, seed: int = 0x5EED02) -> np.ndarray:
    return _limbs(witness_scalars(n, seed))


def witness_rows_fast(n: int, seed: int) -> np.ndarray:
    """The witness-like mix as plain rows from numpy's generator (for large device pools)."""
    gen = np.random.default_rng(seed)
    u = gen.random(n)
    out = uniform_rows(n, seed ^ 0x9E3779B9)
    small = u >= 0.5
    out[small] = 0
    boolean = small & (u < 0.8)
    out[boolean, 0] = gen.integers(0, 2, size=int(boolean.sum()), dtype=np.uint64)
    word = u >= 0.8
    out[word, 0] = gen.integers(0, 1 << 32, size=int(word.sum()), dtype=np.uint64)
    return out


def key_rows(n: int, seed: int = 0x5EED02) -> np.ndarray:
    return _limbs(srs_keys(n, seed))


def srs_points(keys: np.ndarray):
    """Affine Montgomery rows (n, 2, 4) of k_i * G on the device (zk_g1_fixed_base)."""
    from zkgpu import gpu_backend as gb

    gb.require_gpu("workloads.srs_points")
    k = gb.asarray(np.ascontiguousarray(keys, dtype=np.uint64))
    n = k.shape[0]
    out = gb.empty((n, 2, 4))
    gb.call("zk_g1_fixed_base", k.ptr, n, out.ptr, gb.current_stream())
    return out


def log_uniform_sizes(count: int, lo: int, hi: int, seed: int) -> list[int]:
    """floor(2^U) with U uniform in [lo, hi] (the batched MSM sizes)."""
    rng = random.Random(seed)
    return [int(math.floor(2 ** rng.uniform(lo, hi))) for _ in range(count)]


def log_uniform_pow2(count: int, lo: int, hi: int, seed: int) -> list[int]:
    """2^k with k uniform in {lo..hi} (NTT and job sizes)."""
    rng = random.Random(seed)
    return [1 << rng.randint(lo, hi) for _ in range(count)]


def closed_form_scalar(keys: list[int], scalars: list[int]) -> int:
    """sum s_i k_i mod r: for P_i = k_i G the MSM equals this multiple of G."""
    return sum(k * s for k, s in zip(keys, scalars)) % R


# ---- the batched form ---------------------------------------------------------------------

BATCHED_SRS_LOG = 20
BATCHED_SRS_SEED = 0x5EED04


def batched_workload() -> tuple[list[int], list[int]]:
    """(MSM sizes, NTT sizes): 64 MSMs floor(2^U), U uniform in [10, 20], and 32 NTTs of 2^k
    rows, k uniform in {10..20}."""
    return log_uniform_sizes(64, 10, 20, 0x5EED10), log_uniform_pow2(32, 10, 20, 0x5EED11)


def batched_msm_scalars(j: int, n: int) -> np.ndarray:
    """Witness-like plain scalar rows of MSM j of the batched form."""
    return witness_rows_fast(n, 0x5EED1000 + j)


def batched_ntt_rows(i: int, n: int) -> np.ndarray:
    """Uniform plain Fr rows, the input of NTT i of the batched form."""
    return uniform_rows(n, 0x5EED2000 + i)
