"""
zkgpu.autotune: launch parameters measured on the GPU, kept per device.

Three tables, each keyed by log2 of the problem size:

- msm_window:  the signed-digit window c of an MSM of n points (windows of c
               bits, 2^(c-1) buckets each); the library's cost model
               (zk_msm_default_window) is the fallback;
- msm_chunk:   sorted bucket entries per accumulation thread (the tile of the
               bucket accumulation; 0 = one persistent wave of workgroups over
               the whole GPU);
- ntt_fusion:  radix-2 stages fused per LDS-staged NTT pass (which fixes the
               number of HBM sweeps and the columns per tile);

and one device-wide choice:

- msm_accumulate: the bucket accumulation kernel (Jacobian or XYZZ bucket
               coordinates, serial or low-latency products; zk_msm_set_accumulate),
               measured on the largest MSM.

tune() measures every candidate with HIP events (median of several runs) and
writes zkgpu/tuning/<device>.json; lookup() reads the table of the GPU in use,
and apply() hands the NTT fusion depths and the accumulation kernel to the
library (gpu_backend calls it once the device is up). MSMs over an SRS with
window tables take their window from the table chosen per job instead. The reference picks its MSM window from a closed-form
rule (src/zkgpu/msm.py:33-50) and has no NTT fusion at all; that rule stays the
fallback when no table exists for this GPU.

Usage on the GPU:  python -m zkgpu.autotune [--quick] [--no-write]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import re
import statistics
import sys
import time
from pathlib import Path
from typing import Any

TUNING_DIR = Path(__file__).resolve().parent / "tuning"
_CACHE: dict[str, Any] = {}


def device_key() -> str:
    """A file-name-safe name of the current GPU, e.g. 'AMD_Instinct_MI355X_gfx950'."""
    from zkgpu import gpu_backend as gb

    gb.require_gpu("autotune.device_key")
    buf = ctypes.create_string_buffer(256)
    cus, clk = ctypes.c_int(), ctypes.c_int()
    gb.call("zk_device_info", buf, 256, ctypes.byref(cus), ctypes.byref(clk))
    name = buf.value.decode(errors="replace")
    name = name.replace(":sramecc+:xnack-", "")
    return re.sub(r"[^A-Za-z0-9]+", "_", name).strip("_")


def load(device: str | None = None) -> dict[str, Any] | None:
    key = device or device_key()
    if key not in _CACHE:
        path = TUNING_DIR / f"{key}.json"
        _CACHE[key] = json.loads(path.read_text()) if path.exists() else None
    return _CACHE[key]


def lookup(kind: str, n: int, device: str | None = None) -> int:
    """The tuned value of `kind` for size n (nearest tuned power of two), or 0 if none."""
    table = load(device)
    if not table or kind not in table or n <= 0:
        return 0
    entries = {int(k): int(v) for k, v in table[kind].items()}
    if not entries:
        return 0
    lg = math.log2(n)
    best = min(entries, key=lambda k: (abs(k - lg), k))
    return entries[best]


def apply(device: str | None = None) -> int:
    """Hand the tuned NTT fusion depths and accumulation kernel to the library; returns how many were set."""
    from zkgpu import gpu_backend as gb

    table = load(device)
    if not table:
        return 0
    count = 0
    for k, v in table.get("ntt_fusion", {}).items():
        gb.call("zk_ntt_set_fusion", int(k), int(v))
        count += 1
    if "msm_accumulate" in table:
        gb.call("zk_msm_set_accumulate", int(table["msm_accumulate"]))
        count += 1
    return count


# ---- measurement -------------------------------------------------------------------------


def _median_ms(fn, reps: int) -> float:
    from zkgpu import gpu_backend as gb

    fn()
    gb.synchronize()
    times = []
    for _ in range(reps):
        a, b = gb.Event(), gb.Event()
        a.record()
        fn()
        b.record()
        times.append(a.elapsed_ms(b))
    return statistics.median(times)


def _distinct_fusions(log_n: int) -> dict[tuple, int]:
    """Fusion depths 4..10 grouped by the pass plan they give (pass count only matters)."""
    plans: dict[tuple, int] = {}
    for k in range(4, 11):
        passes = 1 if (log_n <= 10 and k >= log_n) else -(-log_n // k)
        plans.setdefault((passes,), k)
    return plans


def tune_ntt(log_sizes: range, reps: int, log) -> tuple[dict[str, int], dict[str, Any]]:
    from zkgpu import gpu_backend as gb
    from zkgpu.field import FIELD_MODULUS as R
    from zkgpu.field256_gpu import BN254_FR
    from zkgpu.ntt import get_ntt_root
    from zkgpu.workloads import uniform_rows

    best: dict[str, int] = {}
    seen: dict[str, Any] = {}
    for lg in log_sizes:
        n = 1 << lg
        x = BN254_FR.from_limbs(uniform_rows(n, 0x7A11 + lg))
        y = gb.empty((n, 4))
        omega = gb.limbs_of(get_ntt_root(n, R))
        results = {}
        for _, k in sorted(_distinct_fusions(lg).items()):
            gb.call("zk_ntt_set_fusion", lg, k)
            results[k] = _median_ms(
                lambda: gb.call("zk_ntt", 0, x.ptr, y.ptr, lg, 0, omega, None, gb.current_stream()), reps)
        k_best = min(results, key=results.get)
        gb.call("zk_ntt_set_fusion", lg, k_best)
        best[str(lg)] = k_best
        seen[str(lg)] = {str(k): round(v, 4) for k, v in results.items()}
        log(f"ntt 2^{lg}: " + ", ".join(f"{k}:{v:.3f}ms" for k, v in results.items()) + f" -> {k_best}")
    return best, seen


def tune_msm(log_sizes: range, reps: int, log) -> tuple[dict[str, int], dict[str, int], int, dict[str, Any]]:
    from zkgpu import gpu_backend as gb
    from zkgpu.batch import SRS, MsmBatch
    from zkgpu.workloads import key_rows, srs_points, witness_rows_fast

    top = 1 << max(log_sizes)
    srs = SRS(srs_points(key_rows(top, 0x7A12)))
    pool = gb.asarray(witness_rows_fast(top, 0x7A13))
    windows: dict[str, int] = {}
    chunks: dict[str, int] = {}
    seen: dict[str, Any] = {}
    accumulate = 0
    for lg in log_sizes:
        n = 1 << lg
        c0 = ctypes.c_int()
        gb.call("zk_msm_default_window", n, ctypes.byref(c0))
        results = {}
        for c in range(max(4, c0.value - 3), min(20, c0.value + 3) + 1):
            plan = MsmBatch(srs, [pool[:n]], windows=[c])
            results[c] = _median_ms(plan.launch, reps)
            del plan
        c_best = min(results, key=results.get)
        chunk_res = {}
        for chunk in (0, 4, 8, 16, 32):
            plan = MsmBatch(srs, [pool[:n]], windows=[c_best], chunk_len=chunk)
            chunk_res[chunk] = _median_ms(plan.launch, reps)
            del plan
        k_best = min(chunk_res, key=chunk_res.get)
        windows[str(lg)] = c_best
        chunks[str(lg)] = k_best
        if lg == max(log_sizes):
            var_res = {}
            for v in range(4):
                gb.call("zk_msm_set_accumulate", v)
                plan = MsmBatch(srs, [pool[:n]], windows=[c_best], chunk_len=k_best)
                var_res[v] = _median_ms(plan.launch, reps)
                del plan
            accumulate = min(var_res, key=var_res.get)
            gb.call("zk_msm_set_accumulate", accumulate)
            seen["accumulate_ms"] = {str(v): round(t, 4) for v, t in var_res.items()}
            log(f"msm 2^{lg}: accumulation variant {accumulate} " + ", ".join(f"{v}:{t:.3f}ms" for v, t in var_res.items()))
        seen[str(lg)] = {"window_ms": {str(c): round(v, 4) for c, v in results.items()},
                         "chunk_ms": {str(k): round(v, 4) for k, v in chunk_res.items()}}
        log(f"msm 2^{lg}: window {c_best} ({results[c_best]:.3f} ms; model {c0.value}: {results.get(c0.value, float('nan')):.3f} ms), "
            f"chunk {k_best} ({chunk_res[k_best]:.3f} ms)")
    return windows, chunks, accumulate, seen


def tune(*, quick: bool = False, write: bool = True, log=print) -> dict[str, Any]:
    from zkgpu import gpu_backend as gb

    gb.require_gpu("autotune.tune")
    reps = 3 if quick else 5
    ntt_sizes = range(10, 23, 2 if quick else 1)
    msm_sizes = range(10, 21, 2 if quick else 1)
    t0 = time.time()
    fusion, ntt_ms = tune_ntt(ntt_sizes, reps, log)
    win, chunk, accumulate, msm_ms = tune_msm(msm_sizes, reps, log)
    table = {
        "device": device_key(),
        "created": time.strftime("%Y-%m-%d %H:%M:%S"),
        "tuning_seconds": round(time.time() - t0, 1),
        "msm_window": win,
        "msm_chunk": chunk,
        "ntt_fusion": fusion,
        "msm_accumulate": accumulate,
        "measured_ms": {"ntt": ntt_ms, "msm": msm_ms},
    }
    if write:
        TUNING_DIR.mkdir(exist_ok=True)
        path = TUNING_DIR / f"{table['device']}.json"
        path.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
        _CACHE.pop(table["device"], None)
        log(f"wrote {path}")
    return table


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[1])
    ap.add_argument("--quick", action="store_true", help="every other size, fewer repetitions")
    ap.add_argument("--no-write", action="store_true", help="print the table without writing it")
    args = ap.parse_args(argv)
    table = tune(quick=args.quick, write=not args.no_write, log=lambda *a: print(*a, file=sys.stderr, flush=True))
    print(json.dumps({k: table[k] for k in ("device", "msm_window", "msm_chunk", "ntt_fusion", "msm_accumulate")}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
