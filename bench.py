#!/usr/bin/env python3
"""
bench.py: the BN254 prover core on one MI355X.

A step is one plain G1 MSM over 2^20 device-resident points P_i = k_i G
(seeded k_i in [1, r)) with witness-like scalars (50% uniform in [0, r), 30% in
{0, 1}, 20% below 2^32), through zk_msm_g1: the headline G1 MSM Mpts/s at 2^20.
The same run measures the Fr NTT at 2^22 (Melem/s, forward), the Groth16
quotient pipeline at 2^22, the batched form against the plain per-call path
on the same mixed-size workload, and the served form's p50/p99 latency, and
prints ONE JSON line.

roofline: the bucket accumulation (k_accumulate) is the step's main
throughput kernel; its duration comes from HIP events recorded on its own
stream around each launch (zk_prof_*). Algorithmic work per launch = (nonzero
signed digits) x 11 Fq multiplications x 136 32-bit multiply-adds, the cost
basis of a mixed addition on 8x32-bit limbs; the peak is gfx950's measured
v_mad_u64_u32 rate (zk_valu_peak). Its HBM traffic per launch comes from the
committed byte-counter summary profiles/k_accumulate_traffic.json (rocprofv3
FETCH_SIZE and WRITE_SIZE runs scaled by the zk_copy probe, --traffic-probe).
cpu_baseline: the oracle's Python-int msm_pippenger (the reference's execution
model: one core) on a 2^18-point sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "zk-gpu-accelerator_amd"))

HBM_PEAK_GBPS = 8000.0
SRS_TABLES = (12, 16)  # window tables of the batched/served SRS


def timed(fn, steps: int, warmup: int, gb) -> float:
    """Milliseconds per step over exactly `steps` calls, synchronised on both sides."""
    for _ in range(warmup):
        fn()
    gb.synchronize()
    start, end = gb.Event(), gb.Event()
    start.record()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    end.record()
    gb.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    return max(start.elapsed_ms(end), 1e-9) / steps if steps else wall


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=10,
                    help="untimed steps first (the GPU's clocks take some milliseconds of load to settle)")
    ap.add_argument("--log-msm", type=int, default=20)
    ap.add_argument("--log-ntt", type=int, default=22)
    ap.add_argument("--no-extras", action="store_true", help="only the headline MSM")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--ntt-only", action="store_true",
                    help="only the forward NTT at --log-ntt (for the pass kernel's trace and byte counters)")
    ap.add_argument("--field-only", action="store_true",
                    help="only the 2^24 Fr products (for k_binary's trace and counters)")
    ap.add_argument("--ntt64-only", action="store_true",
                    help="only the 64-bit NTT at 2^23 (for k_ntt64_pass's trace and byte counters)")
    ap.add_argument("--serve-jobs", type=int, default=2000, help="jobs of the served workload")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="an extra Poisson arrival rate (jobs/s) beside 0.5, 0.8 and 0.95 of the measured capacity")
    ap.add_argument("--serve-budget-gib", type=float, default=128.0, help="HBM budget of the server")
    ap.add_argument("--plain-sample", type=int, default=200, help="jobs timed through the per-call path")
    ap.add_argument("--msm-sizes", metavar="LOGS",
                    help="only plain MSMs at these log2 sizes (comma-separated), --warmup untimed and --steps timed "
                         "each (for a per-size kernel trace); one JSON line")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="copy 256 MiB with zk_copy first (the known-size kernel for scaling byte counters)")
    ap.add_argument("--suite", metavar="JSON",
                    help="instead of the bench line: the reference's benchmark suite at its published sizes "
                         "(Fr products, Fr NTT, G1 MSM, the Groth16 demo), written as JSON in its results schema")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
            "run with --gpus 1",
            file=sys.stderr,
        )
        return 2

    from zkgpu import gpu_backend as gb
    from zkgpu import workloads
    from zkgpu.field256_gpu import BN254_FR
    from zkgpu.msm_gpu import msm_g1_jacobian

    gb.require_gpu("bench")
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    if args.suite:
        _reference_suite(Path(args.suite), gb, log)
        return 0

    if args.msm_sizes:
        print(json.dumps(_msm_sizes(args, gb, workloads, msm_g1_jacobian, log)), flush=True)
        return 0

    if args.ntt64_only:
        if args.traffic_probe:
            _copy_probe(gb)
        print(json.dumps(_ntt64(args, gb, log)), flush=True)
        return 0

    if args.field_only:
        print(json.dumps(_field_mul(args, gb, BN254_FR, workloads, log)), flush=True)
        return 0

    if args.ntt_only:
        if args.traffic_probe:
            _copy_probe(gb)
        print(json.dumps(_ntt_and_pipeline(args, gb, BN254_FR, workloads, log, quotient=False)), flush=True)
        return 0

    # ---- headline: plain MSM at 2^20 --------------------------------------
    n = 1 << args.log_msm
    log(f"setup: basis of {n} points k_i G on the GPU, witness-like scalars")
    keys = workloads.key_rows(n, 0x5EED04)
    points = workloads.srs_points(keys)
    scalars_host = workloads.witness_rows(n, 0x5EED04)
    scalars = gb.asarray(scalars_host)
    out = gb.empty((3, 4))
    window = int(os.environ.get("ZK_MSM_WINDOW", "0"))

    def msm_step() -> None:
        msm_g1_jacobian(points, scalars, out, window_size=window)

    if args.traffic_probe:
        _copy_probe(gb)
    log("msm: warmup + timed")
    gb.prof_enable(True)
    for _ in range(args.warmup):
        msm_step()
    gb.synchronize()
    gb.prof_read("msm_accumulate", reset=True)
    ms = timed(msm_step, args.steps, 0, gb)
    acc_ms, acc_launches = gb.prof_read("msm_accumulate")
    gb.prof_read("", reset=True)
    # the timers' events leave about 10 us idle around each timed launch (12 us per MSM here), so
    # they stay off for everything below except the NTT's own pass-timing loop
    gb.prof_enable(False)
    from zkgpu.msm_gpu import _result

    result = _result(out)
    if not result.is_on_curve():
        raise RuntimeError("bench: MSM result is not on the curve")
    from zkgpu.autotune import lookup

    c_used = window or lookup("msm_window", n) or _default_window(gb, n)
    entries = _msm_entries(gb, points, scalars, n, window)
    mads_per_entry = 11 * 136
    acc_avg = acc_ms / max(acc_launches, 1)
    achieved_tmads = entries * mads_per_entry / (acc_avg * 1e-3) / 1e12
    peaks = _valu_peaks(gb)
    msm_mpts = n / (ms * 1e-3) / 1e6
    log(f"msm 2^{args.log_msm}: {ms:.3f} ms/step, {msm_mpts:.2f} Mpts/s; accumulate {acc_avg:.3f} ms, {entries} entries "
        f"({achieved_tmads:.2f} Tmad/s; peak {peaks['derived']:.2f} derived, {peaks['measured_max']:.2f} measured)")

    traffic = _traffic("k_accumulate", "msm_points", n)
    extras: dict = {}
    if not args.no_extras:
        extras.update(_public_call(gb, points, scalars_host, ms, log))
        extras.update(_msm_2p16(args, gb, workloads, msm_g1_jacobian, log))
        extras.update(_ntt_and_pipeline(args, gb, BN254_FR, workloads, log))
        extras.update(_field_mul(args, gb, BN254_FR, workloads, log))
        extras.update(_ntt64(args, gb, log))
        from zkgpu.batch import SRS

        log("srs: device-resident basis with window tables of widths 12 and 16 (setup, not timed)")
        t0 = time.perf_counter()
        srs = SRS(points).precompute(SRS_TABLES)
        gb.synchronize()
        srs_s = time.perf_counter() - t0
        extras.update(_batched(args, gb, points, srs, workloads, log))
        extras.update(_served(args, gb, srs, log))
        extras["srs"] = {"points": n, "window_tables": list(SRS_TABLES), "setup_s": round(srs_s, 3)}
    cpu = None if args.no_cpu else _cpu_baseline(points, scalars_host, gb, log)

    line = {
        "metric": "G1 MSM Mpts/s @2^20, Fr NTT Melem/s @2^22, batched/plain speedup, p99 ms",
        "value": round(msm_mpts, 4),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (29-bit-limb Montgomery Fq, exact)",
        "data": "synthetic: P_i = k_i G with seeded k_i, witness-like seeded scalars",
        "config": {
            "workload": f"plain BN254 G1 Pippenger MSM, 2^{args.log_msm} device-resident points, witness-like scalars",
            "msm_points": n,
            "window_bits_signed": c_used,
            "autotuned": bool(lookup("msm_window", n)),
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": "k_accumulate",
            "bound": "valu",
            "achieved": round(achieved_tmads, 4),
            **_peak_fields(achieved_tmads, peaks),
            "unit": "Tmad/s",
            "traffic": traffic.get("bytes_per_launch") if traffic else None,
            "traffic_source": traffic.get("source") if traffic else None,
            "algorithmic_bytes_per_launch": entries * 72,
            "avg_launch_ms": round(acc_avg, 4),
            "units_per_launch": entries,
            "units_source": "the plan's sorted-entry count (zk_msm_g1_entries on the timed inputs)",
            "counters_source": "profiles/k_accumulate_msm_2p20_counters.txt",
            "work_per_unit": "one mixed addition: 11 Fq mul x 136 32-bit mad; 72 bytes (key, point ref, affine point)",
            "product_ceiling": {**_product_fields(entries * MADD_MADS / 162 / (acc_avg * 1e-3), gb),
                                "products_per_unit": round(MADD_MADS / 162, 3),
                                "note": "madd-2008-s as executed (1,467 multiply-adds = 9.06 products of 162) "
                                        "against zk_product_peak"},
        },
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(line), flush=True)
    return 0


def _public_call(gb, points, scalars_host, resident_ms: float, log) -> dict:
    """The reference's public call at the headline size: msm_pippenger_gpu with Python-int
    coordinates and scalars (host conversion, PCIe upload, on-device validation, the MSM, the
    result back as a Point), median of 3 after one warm-up; the lists are built untimed."""
    import statistics

    from zkgpu.field256_gpu import BN254_FQ, _limbs_to_ints
    from zkgpu.msm_gpu import msm_pippenger_gpu

    n = points.shape[0]
    coords = BN254_FQ.to_ints(points.reshape(-1, 4))
    xs, ys = coords[0::2], coords[1::2]
    sc = _limbs_to_ints(scalars_host)
    msm_pippenger_gpu(xs, ys, sc)
    samples = []
    for _ in range(3):
        t0 = time.perf_counter()
        msm_pippenger_gpu(xs, ys, sc)
        samples.append((time.perf_counter() - t0) * 1e3)
    ms = statistics.median(samples)
    log(f"public call, 2^{n.bit_length() - 1} Python ints: {ms:.1f} ms ({ms / resident_ms:.0f}x the device-resident step)")
    return {"msm_public_python_ints": {"n": n, "ms": round(ms, 2), "vs_resident": round(ms / resident_ms, 1),
                                       "note": "msm_pippenger_gpu(xs, ys, scalars) with Python ints: host conversion, "
                                               "PCIe upload, on-device validation, MSM, result as a Point"}}


def _msm_sizes(args, gb, workloads, msm_g1_jacobian, log) -> dict:
    """Plain MSMs at several sizes (P_i = k_i G, k_i seeded 0x5EED02, witness-like scalars), each
    timed alone: the latency of one MSM by size."""
    res = {}
    for lg in (int(v) for v in args.msm_sizes.split(",")):
        n = 1 << lg
        pts = workloads.srs_points(workloads.key_rows(n, 0x5EED02))
        sc = gb.asarray(workloads.witness_rows_fast(n, 0x5EED02))
        out = gb.empty((3, 4))
        ms = timed(lambda: msm_g1_jacobian(pts, sc, out), args.steps, args.warmup, gb)
        log(f"msm 2^{lg}: {ms:.4f} ms")
        res[str(lg)] = round(ms, 4)
    return {"msm_ms_by_log_size": res}


def _msm_2p16(args, gb, workloads, msm_g1_jacobian, log) -> dict:
    """The plain MSM at 2^16 points (P_i = k_i G, k_i seeded 0x5EED02, witness-like scalars)."""
    n = 1 << 16
    pts = workloads.srs_points(workloads.key_rows(n, 0x5EED02))
    sc = gb.asarray(workloads.witness_rows(n, 0x5EED02))
    out = gb.empty((3, 4))
    ms = timed(lambda: msm_g1_jacobian(pts, sc, out), args.steps, 10, gb)
    log(f"msm 2^16: {ms:.3f} ms, {n / (ms * 1e-3) / 1e6:.2f} Mpts/s")
    return {"msm_2p16": {"ms": round(ms, 4), "mpts_s": round(n / (ms * 1e-3) / 1e6, 3)}}


def _copy_probe(gb) -> None:
    """256 MiB through zk_copy: the known-size kernel that scales the byte counters."""
    probe = 256 << 20
    a, b = gb.empty(probe // 8), gb.empty(probe // 8)
    gb.call("zk_copy", b.ptr, a.ptr, probe, gb.current_stream())
    gb.synchronize()
    del a, b


def _traffic(kernel: str, key: str, size: int) -> dict | None:
    """HBM bytes per launch of `kernel` from the committed byte-counter summary, if one exists
    for this workload size (`key` names the size field)."""
    path = ROOT / "profiles" / f"{kernel}_traffic.json"
    if not path.exists():
        return None
    data = json.loads(path.read_text())
    if data.get(key) != size:
        return None
    return {"bytes_per_launch": data["bytes_per_launch"], "source": f"profiles/{path.name}"}


def _default_window(gb, n: int) -> int:
    import ctypes

    c = ctypes.c_int(0)
    gb.call("zk_msm_default_window", n, ctypes.byref(c))
    return int(c.value)


def _valu_peak(gb) -> float:
    """gfx950 v_mad_u64_u32 throughput in T/s, measured on this GPU (zk_valu_peak)."""
    import ctypes

    out = ctypes.c_double(0)
    gb.call("zk_valu_peak", ctypes.byref(out))
    return float(out.value)


_PEAKS: dict = {}


def _valu_peaks(gb, runs: int = 5) -> dict:
    """The VALU multiply-add peak two ways: derived from the device's compute units and peak engine
    clock (64 lanes per CU, one v_mad_u64_u32 per lane per clock), and the maximum of `runs`
    zk_valu_peak measurements (a dependent-chain microbenchmark; it swings with clocks, so the
    maximum, not one sample). Measured once per process."""
    import ctypes

    if not _PEAKS:
        name = ctypes.create_string_buffer(256)
        cus, clk = ctypes.c_int(), ctypes.c_int()
        gb.call("zk_device_info", name, 256, ctypes.byref(cus), ctypes.byref(clk))
        derived = cus.value * 64 * clk.value * 1e3 / 1e12
        samples = [_valu_peak(gb) for _ in range(runs)]
        _PEAKS.update(derived=derived, measured_max=max(samples), samples=[round(x, 3) for x in samples],
                      source=f"{cus.value} CUs x 64 lanes x {clk.value // 1000} MHz; max of {runs} zk_valu_peak runs")
    return _PEAKS


def _product_peak(gb, runs: int = 3) -> float:
    """Measured 9-limb Montgomery products per second (10^9) on every lane with no memory traffic
    (zk_product_peak): the ceiling the field arithmetic sets. Maximum of `runs`, once per process."""
    import ctypes

    if "products" not in _PEAKS:
        out = ctypes.c_double(0)
        samples = []
        for _ in range(runs):
            gb.call("zk_product_peak", ctypes.byref(out))
            samples.append(float(out.value))
        _PEAKS["products"] = max(samples)
    return _PEAKS["products"]


# multiply-adds of one madd-2008-s in XYZZ coordinates: 6 products of 162 (two 81-term halves), two
# squarings of 126, and one pair of products that share a reduction (3 halves, 243)
MADD_MADS = 6 * 162 + 2 * 126 + 243


def _product_fields(products_per_s: float, gb) -> dict:
    peak = _product_peak(gb)
    return {"measured_gproducts_s": round(peak, 2), "achieved_gproducts_s": round(products_per_s / 1e9, 2),
            "frac": round(products_per_s / 1e9 / peak, 4)}


def _peak_fields(achieved: float, peaks: dict) -> dict:
    return {
        "peak": round(peaks["derived"], 4),
        "frac": round(achieved / peaks["derived"], 4),
        "peak_measured_max": round(peaks["measured_max"], 4),
        "frac_vs_measured_max": round(achieved / peaks["measured_max"], 4),
        "peak_samples": peaks["samples"],
        "peak_source": peaks["source"],
    }


def _msm_entries(gb, points, scalars, n: int, window: int) -> int:
    """The (bucket, point) entries the accumulation processes for these inputs: the plan's own
    sorted-entry count, read after one zk_msm_g1_entries call."""
    import ctypes

    out = gb.empty((3, 4))
    m = ctypes.c_uint64(0)
    gb.call("zk_msm_g1_entries", points.ptr, scalars.ptr, n, window, 0, out.ptr, ctypes.byref(m), gb.current_stream())
    return int(m.value)


def _field_mul(args, gb, Fr, workloads, log) -> dict:
    """Element-wise Fr Montgomery products over 2^24 device-resident rows (the reference's
    Field256.mul benchmark, field256_gpu.py:325-341, at a size that fills the GPU): 64 bytes in
    and 32 out per product, so the kernel is bound by HBM, with its VALU rate beside it (two
    29-bit-limb Montgomery products per element, since the rows keep the reference's 2^256 form:
    a b 2^-256 = mont(a, mont(b, 2^266)))."""
    n = 1 << 24
    log("field: 2^24 Fr products, uniform rows")
    a = Fr.from_limbs(workloads.uniform_rows(n, 0x5EED41))
    b = Fr.from_limbs(workloads.uniform_rows(n, 0x5EED42))
    out = gb.empty((n, 4))
    fid = Fr._compile()["id"]

    def mul_step() -> None:
        gb.call("zk_field_mul", fid, a.ptr, 0, b.ptr, 0, out.ptr, n, gb.current_stream())

    ms = timed(mul_step, args.steps, 20, gb)
    gbps = 96.0 * n / (ms * 1e-3) / 1e9
    gmul = n / (ms * 1e-3) / 1e9
    tmads = n * 2 * 136 / (ms * 1e-3) / 1e12
    log(f"field mul 2^24: {ms:.3f} ms, {gmul:.2f} G mul/s, {gbps:.0f} GB/s, {tmads:.2f} Tmad/s")
    del a, b, out
    return {
        "field_mul": {
            "n": n,
            "ms": round(ms, 4),
            "gmul_s": round(gmul, 3),
            "roofline": {"kernel": "k_binary<mul>", "bound": "hbm", "achieved": round(gbps, 1), "peak": HBM_PEAK_GBPS,
                         "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4),
                         "algorithmic_bytes_per_launch": 96 * n, "valu_tmad_s": round(tmads, 3),
                         "avg_launch_ms": round(ms, 4),
                         "trace_and_counters_source": "profiles/k_binary_fr_mul_2p24_counters.txt"},
        }
    }


def _ntt64(args, gb, log, log_n: int = 23) -> dict:
    """The 64-bit NTT (the reference's ntt_gpu, int64 modulo 998244353) at 2^23 elements, its largest
    size for that modulus, device
    resident. Each pass reads and writes every element once (16 bytes per element per pass, the
    algorithmic bytes of a launch); pass times from the library's kernel timers."""
    from zkgpu.ntt_gpu import ntt_gpu

    mod = 998_244_353
    n = 1 << log_n
    x = gb.asarray(np.random.default_rng(0x5EED64).integers(0, mod, n, dtype=np.int64))
    ntt_gpu(x, mod)
    ms = timed(lambda: ntt_gpu(x, mod), args.steps, 10, gb)
    gb.prof_enable(True)
    ntt_gpu(x, mod)
    gb.synchronize()
    gb.prof_read("k_ntt64_pass", reset=True)
    timed(lambda: ntt_gpu(x, mod), args.steps, 0, gb)
    pass_ms, passes = gb.prof_read("k_ntt64_pass", reset=True)
    gb.prof_enable(False)
    per_pass = pass_ms / max(passes, 1)
    gbps = 16.0 * n / (per_pass * 1e-3) / 1e9
    log(f"ntt64 2^{log_n}: {ms:.3f} ms ({n / (ms * 1e-3) / 1e6:.0f} Melem/s), pass {per_pass:.4f} ms, {gbps:.0f} GB/s")
    return {"ntt64": {"log_n": log_n, "mod": mod, "ms": round(ms, 4), "melem_s": round(n / (ms * 1e-3) / 1e6, 1),
                      "roofline": {"kernel": "k_ntt64_pass", "bound": "hbm", "achieved": round(gbps, 1),
                                   "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4),
                                   "algorithmic_bytes_per_launch": 16 * n, "avg_launch_ms": round(per_pass, 4),
                                   "passes": max(1, round(passes / args.steps))}}}


def _ntt_and_pipeline(args, gb, Fr, workloads, log, quotient: bool = True) -> dict:
    n = 1 << args.log_ntt
    log(f"ntt: 2^{args.log_ntt} forward, uniform Fr rows")
    x = Fr.from_limbs(workloads.uniform_rows(n, 0x5EED03))
    out = gb.empty((n, 4))
    from zkgpu.ntt import get_ntt_root
    from zkgpu.field import FIELD_MODULUS as R

    omega = gb.limbs_of(get_ntt_root(n, R))

    def ntt_step() -> None:
        gb.call("zk_ntt", 0, x.ptr, out.ptr, args.log_ntt, 0, omega, None, gb.current_stream())

    ntt_step()
    gb.synchronize()
    # the transform as a caller sees it (no kernel timers), after 20 untimed transforms: the
    # GPU's clocks take some milliseconds of sustained load to settle
    ms = timed(ntt_step, args.steps, 20, gb)
    # the passes' kernel time, from the library's HIP-event timers over a second timed loop
    gb.prof_enable(True)
    ntt_step()
    gb.synchronize()
    gb.prof_read("ntt_pass", reset=True)
    timed(ntt_step, args.steps, 0, gb)
    pass_ms, passes = gb.prof_read("ntt_pass", reset=True)
    gb.prof_enable(False)
    per_pass = pass_ms / max(passes, 1)
    passes_per_ntt = max(1, round(passes / args.steps))
    ntt_pass_ms = per_pass * passes_per_ntt  # kernel time of one transform
    melem = n / (ms * 1e-3) / 1e6
    hbm_gbps = 64.0 * n / (ms * 1e-3) / 1e9
    pass_gbps = 64.0 * n / (per_pass * 1e-3) / 1e9  # each pass reads and writes every row once
    # one Montgomery product per butterfly, n/2 butterflies per stage, log n stages
    ntt_tmads = args.log_ntt * (n // 2) * 136 / (ntt_pass_ms * 1e-3) / 1e12
    peaks = _valu_peaks(gb)
    log(f"ntt 2^{args.log_ntt}: {ms:.3f} ms, {melem:.1f} Melem/s; pass {per_pass:.3f} ms ({pass_gbps:.0f} GB/s, "
        f"{ntt_tmads:.2f} Tmad/s; peak {peaks['derived']:.2f} derived)")

    traffic = _traffic("k_ntt_pass", "ntt_log_n", args.log_ntt)
    extras = {}
    if quotient:
        log("groth16 quotient pipeline at the same size (7 NTTs)")
        a, b, c = (Fr.from_limbs(workloads.uniform_rows(n, s)) for s in (0x5EED31, 0x5EED32, 0x5EED33))
        h = gb.empty((n, 4))
        shift = gb.limbs_of(7)

        def pipe_step() -> None:
            gb.call("zk_groth16_quotient", a.ptr, b.ptr, c.ptr, h.ptr, args.log_ntt, shift, gb.current_stream())

        pipe_step()
        pms = timed(pipe_step, max(1, args.steps // 2), 3, gb)
        log(f"groth16 2^{args.log_ntt}: {pms:.3f} ms")
        extras["groth16_quotient"] = {"log_n": args.log_ntt, "ms": round(pms, 4), "coset_shift": 7}
    return {
        "ntt": {
            "log_n": args.log_ntt,
            "ms": round(ms, 4),
            "melem_s": round(melem, 2),
            "algorithmic_gbps": round(hbm_gbps, 1),
            "pass_ms": round(per_pass, 4),
            "passes": passes_per_ntt,
            "roofline": {
                "kernel": "k_ntt_pass",
                "bound": "valu",
                "achieved": round(ntt_tmads, 4),
                **_peak_fields(ntt_tmads, peaks),
                "unit": "Tmad/s",
                "product_ceiling": {**_product_fields(args.log_ntt * (n // 2) / (ntt_pass_ms * 1e-3), gb),
                                    "products_per_unit": 1.0,
                                    "note": "one 9-limb Montgomery product per butterfly against zk_product_peak"},
                "work_per_unit": "one butterfly: 1 Fr mul x 136 32-bit mad; n/2 butterflies x log n stages",
                "hbm": {"achieved": round(pass_gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                        "frac": round(pass_gbps / HBM_PEAK_GBPS, 4), "bytes_per_pass": 64 * n},
                "traffic": traffic.get("bytes_per_launch") if traffic else None,
                "traffic_source": traffic.get("source") if traffic else None,
                "algorithmic_bytes_per_launch": 64 * n,
                "avg_launch_ms": round(per_pass, 4),
                "counters_source": "profiles/k_ntt_pass_fr_2p22_counters.txt",
            },
        },
        **extras,
    }


def _batched(args, gb, points, srs, workloads, log) -> dict:
    """
    The batched form: 64 MSMs with sizes log-uniform in [2^10, 2^20] over prefixes of the
    2^20-point basis (resident, with window tables 12 and 16) plus 32 forward NTTs of sizes 2^k,
    k uniform in {10..20}, as one graph replay (the workload the full-size parity test checks:
    workloads.batched_workload and its per-job seeds). Two per-call baselines run the same work
    one call at a time, each MSM read back before the next starts (as msm_pippenger_gpu returns a
    point) and one zk_ntt per transform:
      raw    - zk_msm_g1 over the raw points (no precomputation),
      tables - a one-job zk_msm_g1_batch over the SRS with its window tables, the batched form's
               own precomputation without the batching.
    """
    from zkgpu.batch import MsmBatch, NttBatch, capture
    from zkgpu.field256_gpu import BN254_FR
    from zkgpu.field import FIELD_MODULUS as R
    from zkgpu.ntt import get_ntt_root

    import ctypes

    msm_sizes, ntt_sizes = workloads.batched_workload()
    log(f"batched: 64 MSMs ({sum(msm_sizes)} points) + 32 NTTs ({sum(ntt_sizes)} elements)")
    scal = [gb.asarray(workloads.batched_msm_scalars(j, n)) for j, n in enumerate(msm_sizes)]
    ntt_in = [BN254_FR.from_limbs(workloads.batched_ntt_rows(i, n)) for i, n in enumerate(ntt_sizes)]
    ntt_out = [gb.empty((n, 4)) for n in ntt_sizes]
    outs = [gb.empty((3, 4)) for _ in msm_sizes]
    omegas = {n: gb.limbs_of(get_ntt_root(n, R)) for n in set(ntt_sizes)}
    one = [((ctypes.c_void_p * 1)(s.ptr), (ctypes.c_int64 * 1)(n)) for s, n in zip(scal, msm_sizes)]

    def ntts() -> None:
        for x, y, n in zip(ntt_in, ntt_out, ntt_sizes):
            gb.call("zk_ntt", 0, x.ptr, y.ptr, n.bit_length() - 1, 0, omegas[n], None, gb.current_stream())

    def plain_raw() -> None:
        for s, n, o in zip(scal, msm_sizes, outs):
            gb.call("zk_msm_g1", points.ptr, s.ptr, n, 0, 0, o.ptr, gb.current_stream())
            o.get()
        ntts()
        gb.synchronize()

    def plain_tables() -> None:
        for (ptrs, sizes), o in zip(one, outs):
            gb.call("zk_msm_g1_batch", srs.handle, 1, ptrs, sizes, None, o.ptr, gb.current_stream())
            o.get()
        ntts()
        gb.synchronize()

    st = gb.Stream()
    with gb.stream(st.handle):
        msm = MsmBatch(srs, scal)
        ntt = NttBatch(ntt_in, ntt_out, [False] * len(ntt_sizes))
    st.synchronize()
    graph = capture([msm, ntt], st)

    def batched() -> None:
        with gb.stream(st.handle):
            graph.replay()
            msm.out.get()  # every result back on the host, as in the per-call paths
        st.synchronize()

    def clock(fn, reps: int) -> float:
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) * 1e3 / reps

    reps = max(2, args.steps // 2)
    raw_ms = clock(plain_raw, reps)
    tables_ms = clock(plain_tables, reps)
    batch_ms = clock(batched, reps)
    # the batched results equal the per-call ones (as group elements); the full-size parity test
    # pins every one of them to the oracle's closed form
    from zkgpu.msm_gpu import _result

    got = msm.results()
    for k in range(len(msm_sizes)):
        want = _result(outs[k])
        if (got[k].is_identity() != want.is_identity()) or (not want.is_identity() and got[k].affine() != want.affine()):
            raise RuntimeError(f"bench: batched MSM {k} differs from the per-call result")
    log(f"batched: per-call raw {raw_ms:.2f} ms, per-call with tables {tables_ms:.2f} ms, batched graph "
        f"{batch_ms:.2f} ms: {raw_ms / batch_ms:.2f}x and {tables_ms / batch_ms:.2f}x")
    # the reference's public API, Python ints in (msm_pippenger_gpu), on a sample of the MSMs,
    # scaled to the workload by point count
    from zkgpu.field256_gpu import BN254_FQ, _limbs_to_ints
    from zkgpu.msm_gpu import msm_pippenger_gpu

    sample = [k for k in range(len(msm_sizes)) if msm_sizes[k] <= (1 << 16)][:6]
    coords = _limbs_to_ints(BN254_FQ.to_limbs(points[: max(msm_sizes[k] for k in sample)].reshape(-1, 4)).get())
    ints = {k: _limbs_to_ints(scal[k].get()) for k in sample}
    t0 = time.perf_counter()
    for k in sample:
        m = msm_sizes[k]
        msm_pippenger_gpu(coords[0 : 2 * m : 2], coords[1 : 2 * m : 2], ints[k])
    ints_s = time.perf_counter() - t0
    ints_ms_total = ints_s * 1e3 * sum(msm_sizes) / sum(msm_sizes[k] for k in sample)
    log(f"batched: public API with Python ints, {len(sample)} sampled MSMs: {ints_s:.2f} s; scaled {ints_ms_total:.0f} ms")
    return {
        "batched": {
            "workload": "64 G1 MSMs log-uniform in [2^10, 2^20] over one 2^20 SRS + 32 Fr NTTs 2^10..2^20",
            "per_call_raw_ms": round(raw_ms, 3),
            "per_call_tables_ms": round(tables_ms, 3),
            "batched_graph_ms": round(batch_ms, 3),
            "speedup": round(tables_ms / batch_ms, 3),
            "speedup_note": "against the per-call path over the same SRS and window tables (batching alone)",
            "speedup_vs_raw_per_call": round(raw_ms / batch_ms, 3),
            "per_call_python_ints_ms_est": round(ints_ms_total, 1),
            "per_call_python_ints_note": f"msm_pippenger_gpu with Python ints on {len(sample)} of the MSMs "
                                         "(sizes up to 2^16), scaled by point count; the NTTs excluded",
            "msm_windows": "per job: the SRS window table (widths %s) with the least work for its size"
                           % ", ".join(str(c) for c in SRS_TABLES),
        }
    }


def _served(args, gb, srs, log) -> dict:
    """
    The served form: Groth16 proof jobs (quotient at size n with coset shift 7 + three G1 MSMs
    of size n over the shared 2^20 SRS), n = 2^k with k uniform in {12..20}. Capacity comes from
    all jobs queued at once; then Poisson arrivals at 0.5, 0.8 and 0.95 of that capacity give the
    latency and batch-size curve (the headline p50/p99 at 0.8). The per-call paths (one quotient
    call and three MSM calls per job, in order, each result read back) are timed on the first
    jobs, over the raw points and over the SRS with its window tables.
    """
    from zkgpu.serve import ProofServer, WitnessPools, run_plain, served_workload

    pools = WitnessPools((1 << 20) + (1 << 16), seed=0x5EED05)
    budget = int(args.serve_budget_gib * 2**30)
    server = ProofServer(srs, hbm_budget=budget, slots=2, max_batch_jobs=64)
    log("served: warm-up (tables for every size)")
    server.serve([pools.job(100000 + k, k, 0.0) for k in range(12, 21)])
    gb.synchronize()
    # one untimed pass of the queued workload first: its batches allocate the workspace that
    # the server's stream caches then reuse, as in a server that has been running a while
    server.serve(served_workload(pools, args.serve_jobs, None, seed=0x5EED07))
    gb.synchronize()
    log(f"served: capacity, {args.serve_jobs} jobs queued at once")
    cap = server.serve(served_workload(pools, args.serve_jobs, None, seed=0x5EED06)).summary()
    log(f"served: capacity {cap['jobs_per_s']} jobs/s in {cap['batches']} batches")
    gb.synchronize()
    curve = []
    fracs = [0.5, 0.8, 0.95]
    rates = [round(f * cap["jobs_per_s"], 1) for f in fracs]
    if args.serve_rate:
        fracs.append(round(args.serve_rate / cap["jobs_per_s"], 3))
        rates.append(args.serve_rate)
    for f, rate in zip(fracs, rates):
        rep = server.serve(served_workload(pools, args.serve_jobs, rate, seed=0x5EED06)).summary()
        gb.synchronize()
        curve.append({"load": f, "rate_jobs_s": rate, "p50_ms": rep["p50_ms"], "p99_ms": rep["p99_ms"],
                      "mean_ms": rep.get("mean_ms"), "jobs_per_s": rep["jobs_per_s"],
                      "mean_batch_jobs": rep["mean_batch_jobs"], "batches": rep["batches"]})
        log(f"served: {rate} jobs/s ({f:.2f} of capacity): p50 {rep['p50_ms']} ms, p99 {rep['p99_ms']} ms, "
            f"mean batch {rep['mean_batch_jobs']} jobs")
    head = curve[1]
    sample = served_workload(pools, args.plain_sample, None, seed=0x5EED06)
    per_call = {}
    for name, kw in (("raw", {}), ("tables", {"srs": srs})):
        run_plain(sample[0], srs.points, **kw)
        gb.synchronize()
        t0 = time.perf_counter()
        for j in sample:
            run_plain(j, srs.points, **kw)
        gb.synchronize()
        per_call[name] = len(sample) / (time.perf_counter() - t0)
        log(f"served: per-call path ({name}) {per_call[name]:.1f} jobs/s on the first {len(sample)} jobs")
    return {
        "served": {
            "workload": f"{args.serve_jobs} Groth16 jobs, n = 2^k, k uniform in 12..20: quotient + 3 G1 MSMs over a 2^20 SRS",
            "rate_jobs_s": head["rate_jobs_s"],
            "p50_ms": head["p50_ms"],
            "p99_ms": head["p99_ms"],
            "latency_vs_load": curve,
            "capacity": cap,
            "per_call_tables_jobs_s": round(per_call["tables"], 2),
            "per_call_raw_jobs_s": round(per_call["raw"], 2),
            "per_call_sample_jobs": len(sample),
            "capacity_vs_per_call": round(cap["jobs_per_s"] / per_call["tables"], 3),
            "capacity_vs_per_call_note": "against the per-call path over the same SRS and window tables",
            "capacity_vs_raw_per_call": round(cap["jobs_per_s"] / per_call["raw"], 3),
        }
    }


def _host_cpu() -> tuple[str, int]:
    """The host CPU's model name and the cores this process may run on."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def _cpu_ntt_worker(reps: int) -> float:
    """Seconds for `reps` oracle NTTs of 2^12 (one worker process of the parallel-efficiency probe)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "zk-gpu-accelerator_amd"))
    from zkgpu.field import FIELD_MODULUS as R
    from zkgpu.workloads import ntt_coefficients
    from zkref.ntt import ntt

    a = ntt_coefficients(1 << 12)
    t0 = time.perf_counter()
    for _ in range(reps):
        ntt(a, R)
    return time.perf_counter() - t0


def _parallel_efficiency(procs: int, reps: int = 20) -> float:
    """Per-process speed with `procs` oracle processes running at once, relative to one alone
    (fresh interpreters that never touch the GPU)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor

    ctx = mp.get_context("spawn")
    with ProcessPoolExecutor(1, mp_context=ctx) as pool:
        pool.submit(_cpu_ntt_worker, 1).result()
        alone = pool.submit(_cpu_ntt_worker, reps).result()
    with ProcessPoolExecutor(procs, mp_context=ctx) as pool:
        list(pool.map(_cpu_ntt_worker, [1] * procs))
        together = list(pool.map(_cpu_ntt_worker, [reps] * procs))
    return alone / max(together)


def _cpu_baseline(points, scalars_host, gb, log) -> dict:
    """
    The reference's CPU path (the oracle's Python-int restatement of ntt, Radix2Domain's
    quotient and msm_pippenger; CPython, one core) on the host, per workload:
      headline - msm_pippenger on the first 2^18 points and scalars of the 2^20 workload;
      C1       - the NTT at 2^12, median of 5 after a warm-up;
      C2       - the MSM at 2^16, single-shot, checked against the closed form;
      C3       - the quotient pipeline at 2^22, estimated from a single-shot run at 2^16 scaled
                 by n log n (the full run takes minutes on one core);
      C4, C5   - estimates from cost models calibrated on the runs above (MSM seconds per
                 Pippenger bucket addition at the reference's window, NTT seconds per n log n),
                 on 1 core and on P worker processes with P the host cores available (at most
                 16), scaled by the parallel efficiency measured with P oracle processes at once.
    """
    import statistics

    sys.path.insert(0, str(ROOT / "oracle"))
    from zkgpu import workloads
    from zkgpu.field import FIELD_MODULUS as R
    from zkgpu.field256_gpu import BN254_FQ, _limbs_to_ints
    from zkgpu.msm import pippenger_window_size
    from zkref import fast
    from zkref.curve import G, Point
    from zkref.domain import groth16_quotient
    from zkref.field import limbs_to_ints
    from zkref.msm import msm_pippenger
    from zkref.ntt import ntt

    model, cores = _host_cpu()
    procs = min(cores, 16)

    def ops(n: int) -> int:  # bucket additions of the reference's Pippenger at its window
        w = pippenger_window_size(n)
        return -(-254 // w) * (n + (2 << w))

    # headline sample
    sample = 1 << 18
    coords = _limbs_to_ints(BN254_FQ.to_limbs(points[:sample].reshape(-1, 4)).get())
    pts = [Point(coords[2 * i], coords[2 * i + 1]) for i in range(sample)]
    sc = _limbs_to_ints(scalars_host[:sample])
    log(f"cpu baseline: oracle msm_pippenger on {sample} points (1 core, {model})")
    t0 = time.perf_counter()
    msm_pippenger(pts, sc)
    head_s = time.perf_counter() - t0
    del pts, coords

    # C1
    a = workloads.ntt_coefficients(1 << 12)
    ntt(a, R)
    c1 = []
    for _ in range(5):
        t0 = time.perf_counter()
        ntt(a, R)
        c1.append(time.perf_counter() - t0)
    c1_ms = statistics.median(c1) * 1e3

    # C2
    n2 = 1 << 16
    keys2 = workloads.key_rows(n2, 0x5EED02)
    flat = limbs_to_ints(fast.fixed_base_g1(keys2).reshape(-1, 4))
    pts2 = [Point(flat[2 * i], flat[2 * i + 1]) for i in range(n2)]
    sc2_rows = workloads.witness_rows(n2, 0x5EED02)
    sc2 = limbs_to_ints(sc2_rows)
    t0 = time.perf_counter()
    got2 = msm_pippenger(pts2, sc2)
    c2_s = time.perf_counter() - t0
    e2 = fast.dot_mod_r(keys2, sc2_rows)
    c2_ok = (got2.is_identity() and e2 == 0) or got2.affine() == (G * e2).affine()
    del pts2, flat

    # C3 at 2^16, scaled
    n3 = 1 << 16
    abc = [_limbs_to_ints(workloads.uniform_rows(n3, s_)) for s_ in (0x5EED31, 0x5EED32, 0x5EED33)]
    t0 = time.perf_counter()
    groth16_quotient(*abc, 7)
    c3_sample_s = time.perf_counter() - t0
    n3_full = 1 << 22
    scale3 = (n3_full * 22) / (n3 * 16)
    c3_est_s = c3_sample_s * scale3

    # cost models
    alpha = statistics.mean([head_s / ops(sample), c2_s / ops(n2)])  # s per bucket addition
    beta = statistics.mean([c1_ms * 1e-3 / (4096 * 12), c3_sample_s / 7 / (n3 * 16)])  # s per n log n of one NTT
    msm_sizes, ntt_sizes = workloads.batched_workload()
    c4_1core_s = sum(alpha * ops(n) for n in msm_sizes) + sum(beta * n * (n.bit_length() - 1) for n in ntt_sizes)
    job_s = [7 * beta * (1 << k) * k + 3 * alpha * ops(1 << k) for k in range(12, 21)]
    c5_1core_jobs_s = 1.0 / statistics.mean(job_s)
    log(f"cpu baseline: parallel efficiency with {procs} oracle processes")
    eff = _parallel_efficiency(procs)
    log(f"cpu baseline: headline sample {head_s:.1f} s, C1 {c1_ms:.1f} ms, C2 {c2_s:.1f} s, C3 est {c3_est_s:.0f} s, "
        f"C4 est {c4_1core_s:.0f} s (1 core), C5 est {c5_1core_jobs_s:.4f} jobs/s (1 core), efficiency {eff:.2f}")
    return {
        "value": round(sample / head_s / 1e6, 6),
        "unit": "Mpts/s",
        "cores": 1,
        "kind": "port",
        "sample": f"first {sample} points and scalars of the 2^20 workload, Python-int Pippenger (w={_ref_w(sample)}), "
                  f"{head_s:.1f} s",
        "host_cpu": model,
        "host_cores": cores,
        "configs": {
            "C1_ntt_2p12": {"median_ms": round(c1_ms, 3), "runs": 5, "cores": 1, "kind": "port"},
            "C2_msm_2p16": {"s": round(c2_s, 3), "mpts_s": round(n2 / c2_s / 1e6, 6), "cores": 1, "kind": "port",
                            "single_shot": True, "verified_closed_form": bool(c2_ok)},
            "C3_groth16_2p22": {"s_est": round(c3_est_s, 1), "cores": 1, "kind": "port", "estimate": True,
                                "basis": f"single-shot run at 2^16 ({c3_sample_s:.2f} s) x n log n ratio {scale3:.0f}"},
            "C4_batched": {"s_est_1core": round(c4_1core_s, 1), "s_est_P": round(c4_1core_s / (procs * eff), 1),
                           "P": procs, "parallel_efficiency": round(eff, 3), "estimate": True},
            "C5_served": {"jobs_s_est_1core": round(c5_1core_jobs_s, 5), "jobs_s_est_P": round(c5_1core_jobs_s * procs * eff, 4),
                          "P": procs, "parallel_efficiency": round(eff, 3), "estimate": True},
            "cost_model": {"msm_s_per_bucket_add": alpha, "ntt_s_per_nlogn": beta,
                           "note": "bucket additions = ceil(254/w) (n + 2^(w+1)) at the reference's window w; "
                                   "calibrated on the headline sample and C2 (MSM), C1 and the C3 sample (NTT)"},
        },
    }


def _ref_w(n: int) -> int:
    return min(range(1, 17), key=lambda w: (-(-254 // w)) * (n + 2 ** (w + 1)))



# ---- the reference's benchmark suite, in its results schema -------------------------------


def _median_ms(fn, gb, *, repeat: int = 5, warmup: int = 1) -> float:
    """Median wall-clock ms of fn after warm-up calls, the device synchronised before the clock stops
    (benchmarks/_harness.py: median of repeated samples)."""
    import statistics

    for _ in range(warmup):
        fn()
    gb.synchronize()
    samples = []
    for _ in range(repeat):
        t0 = time.perf_counter()
        fn()
        gb.synchronize()
        samples.append((time.perf_counter() - t0) * 1e3)
    return statistics.median(samples)


def _suite_meta(gb, command: str) -> dict:
    """The reference's metadata block (benchmarks/_harness.py:216-241) for this machine."""
    import ctypes
    import datetime
    import platform

    cpu = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    name = ctypes.create_string_buffer(256)
    cus, clk = ctypes.c_int(), ctypes.c_int()
    gb.call("zk_device_info", name, 256, ctypes.byref(cus), ctypes.byref(clk))
    _, total = gb.mem_info()
    return {
        "timestamp_utc": datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"),
        "git_commit": None,
        "git_dirty": None,
        "zkgpu": "MI355X build (libzkgpu_hip.so)",
        "python": platform.python_version(),
        "numpy": np.__version__,
        "os": platform.platform(),
        "cpu": cpu,
        "cpu_logical_cores": os.cpu_count(),
        "cpu_cores_used": 1,
        "gpu": {"name": name.value.decode(errors="replace"), "compute_units": cus.value,
                "clock_mhz": clk.value // 1000, "memory_gib": round(total / 2**30, 1)},
        "gpu_unavailable_reason": None,
        "quick": False,
        "runs": 1,
        "method": "median wall-clock time of 5 samples after one warm-up call; GPU timings synchronise the "
                  "device before the clock stops; 'gpu' starts from device-resident data, 'gpu_e2e_ints' "
                  "includes the Python-int conversion and host<->device copies; CPU columns are the oracle's "
                  "Python-int restatement of the reference's CPU path, one core",
        "command": command,
    }


def _reference_suite(out: Path, gb, log) -> None:
    """The reference's suite at its published sizes (benchmarks/bench_field256.py:35-63,
    bench_ntt256.py:53-99, bench_msm.py:36-95, demo/prover_simulation.py:199-245), with its inputs:
    uniform Fr values from random.Random(n); for the MSM P_i = P_0 + i S from random.Random(n)
    (S's multiple drawn first, then P_0's, then the scalars), kept as the multiples k_i of G so the
    basis is made on the device. Every GPU result is checked against the oracle; its Python-int
    run, on one core, is the CPU column."""
    import random

    sys.path.insert(0, str(ROOT / "oracle"))
    from zkgpu.batch import SRS
    from zkgpu.field import FIELD_MODULUS as R
    from zkgpu.field256_gpu import BN254_FQ as Fq, BN254_FR as Fr, _ints_to_limbs, _ntt_tables
    from zkgpu.msm import pippenger_window_size
    from zkgpu.msm_gpu import msm_pippenger_gpu
    from zkgpu.prover import make_inputs, prove_gpu
    from zkgpu.workloads import srs_points
    from zkref import fast
    from zkref.curve import G, Point
    from zkref.domain import groth16_quotient
    from zkref.field import limbs_to_ints
    from zkref.msm import msm_pippenger
    from zkref.ntt import ntt as ntt_cpu

    def cpu_ms(fn):
        t0 = time.perf_counter()
        out = fn()
        return (time.perf_counter() - t0) * 1e3, out

    def cpu_points(keys: list[int]) -> list[Point]:
        flat = limbs_to_ints(fast.fixed_base_g1(_ints_to_limbs(keys)).reshape(-1, 4))
        return [Point(flat[2 * i], flat[2 * i + 1]) for i in range(len(keys))]

    results: dict = {"meta": _suite_meta(gb, "python bench.py --suite " + str(out))}

    from zkgpu.field_gpu import gpu_batch_mul
    from zkgpu.ntt_gpu import _twiddle_table, ntt_gpu

    mod64 = 998_244_353
    rows = []
    for n in (10_000, 100_000, 1_000_000, 10_000_000):
        rng = np.random.default_rng(n)
        a = rng.integers(0, mod64, size=n, dtype=np.int64)
        b = rng.integers(0, mod64, size=n, dtype=np.int64)
        ms = {}
        ms["numpy"], want = cpu_ms(lambda: a * b % mod64)
        da, db = gb.asarray(a), gb.asarray(b)
        ms["gpu"] = _median_ms(lambda: gpu_batch_mul(da, db, mod64), gb)
        ms["gpu_e2e"] = _median_ms(lambda: gpu_batch_mul(gb.asarray(a), gb.asarray(b), mod64).get(), gb)
        ok = np.array_equal(gpu_batch_mul(da, db, mod64).get(), want)
        rows.append({"n": n, "mod": mod64, "verified": ok, "ms": ms})
        log(f"suite field n={n}: verified={ok} {ms}")
    results["field"] = rows

    rows = []
    for k in (14, 16, 18, 20):
        n = 1 << k
        host = np.random.default_rng(n).integers(0, mod64, size=n, dtype=np.int64)
        ms = {}
        ms["python"], want = cpu_ms(lambda: ntt_cpu(host.tolist(), mod64))
        dev = gb.asarray(host)
        ntt_gpu(gb.asarray(np.zeros(2, np.int64)), mod64)
        colds = []
        for _ in range(3):  # every sample pays the table build, as a first call does
            _twiddle_table.cache_clear()
            gb.synchronize()
            t0 = time.perf_counter()
            ntt_gpu(dev, mod64)
            gb.synchronize()
            colds.append((time.perf_counter() - t0) * 1e3)
        ms["gpu_cold"] = sorted(colds)[1]
        ms["gpu"] = _median_ms(lambda: ntt_gpu(dev, mod64), gb)
        ms["gpu_e2e"] = _median_ms(lambda: ntt_gpu(gb.asarray(host), mod64).get(), gb)
        ok = ntt_gpu(dev, mod64).tolist() == want
        rows.append({"n": n, "mod": mod64, "verified": ok, "ms": ms})
        log(f"suite ntt n=2^{k}: verified={ok} {ms}")
    results["ntt"] = rows

    rows = []
    for n in (10_000, 100_000, 1_000_000):
        rng = random.Random(n)
        a = [rng.randrange(R) for _ in range(n)]
        b = [rng.randrange(R) for _ in range(n)]
        ms = {}
        ms["python_ints"], want = cpu_ms(lambda: [x * y % R for x, y in zip(a, b)])
        da, db = Fr.from_ints(a), Fr.from_ints(b)
        ms["gpu"] = _median_ms(lambda: Fr.mul(da, db), gb)
        ms["gpu_e2e_ints"] = _median_ms(lambda: Fr.to_ints(Fr.mul(Fr.from_ints(a), Fr.from_ints(b))), gb, repeat=3)
        ok = Fr.to_ints(Fr.mul(da, db)) == want
        rows.append({"n": n, "field": "BN254 Fr", "verified": ok, "ms": ms})
        log(f"suite field256 n={n}: verified={ok} {ms}")
    results["field256"] = rows

    rows = []
    for k in (12, 14, 16, 18, 20):
        n = 1 << k
        rng = random.Random(n)
        vals = [rng.randrange(R) for _ in range(n)]
        ms = {}
        if k <= 18:
            ms["python"], want = cpu_ms(lambda: ntt_cpu(vals, R))
        else:  # the oracle's C transform checks the largest size
            want = limbs_to_ints(fast.ntt_rows(_ints_to_limbs(vals)))
        x = Fr.from_ints(vals)
        Fr.ntt(Fr.from_ints([0, 0]))
        colds = []
        for _ in range(3):  # every sample pays the table build, as a first call does
            _ntt_tables.cache_clear()
            gb.synchronize()
            t0 = time.perf_counter()
            Fr.ntt(x)
            gb.synchronize()
            colds.append((time.perf_counter() - t0) * 1e3)
        ms["gpu_cold"] = sorted(colds)[1]
        ms["gpu"] = _median_ms(lambda: Fr.ntt(x), gb)
        ms["gpu_e2e_ints"] = _median_ms(lambda: Fr.to_ints(Fr.ntt(Fr.from_ints(vals))), gb, repeat=3)
        ok = Fr.to_ints(Fr.ntt(x)) == want
        rows.append({"n": n, "field": "BN254 Fr", "verified": ok, "ms": ms})
        log(f"suite ntt256 n=2^{k}: verified={ok} {ms}")
    results["ntt256"] = rows

    rows = []
    for n in (100, 1_000, 10_000, 100_000):
        rng = random.Random(n)
        step, first = rng.randrange(1, R), rng.randrange(1, R)
        keys = [(first + i * step) % R for i in range(n)]
        scalars = [rng.randrange(R) for _ in range(n)]
        w = pippenger_window_size(n)
        pts_dev = srs_points(_ints_to_limbs(keys))
        coords = Fq.to_ints(pts_dev.reshape(2 * n, 4))
        xs, ys = coords[0::2], coords[1::2]
        ms = {}
        basis = cpu_points(keys)
        ms["pippenger_cpu"], want = cpu_ms(lambda: msm_pippenger(basis, scalars, window_size=w))
        got = msm_pippenger_gpu(xs, ys, scalars, window_size=w)
        ms["pippenger_gpu"] = _median_ms(lambda: msm_pippenger_gpu(xs, ys, scalars, window_size=w), gb, repeat=3)
        srs = SRS(pts_dev)
        dev_sc = gb.asarray(_ints_to_limbs(scalars))
        got_res = msm_pippenger_gpu(srs, None, dev_sc)
        ms["pippenger_gpu_resident"] = _median_ms(lambda: msm_pippenger_gpu(srs, None, dev_sc), gb)
        e = fast.dot_mod_r(_ints_to_limbs(keys), _ints_to_limbs(scalars))
        ok = got.affine() == want.affine() == got_res.affine() == (G * e).affine()
        rows.append({"n": n, "window_size": w, "verified": ok, "ms": ms})
        log(f"suite msm n={n}: verified={ok} {ms}")
    results["msm"] = rows

    m = 100_000
    inp = make_inputs(m, seed=0)
    run = prove_gpu(inp)
    ms = {"gpu_quotient": run.quotient_ms, **{f"gpu_msm_{k}": v for k, v in run.msm_ms.items()},
          "gpu_total": run.total_ms}
    ms["python_ints_quotient"], h_cpu = cpu_ms(lambda: groth16_quotient(inp.a, inp.b, inp.c, inp.domain.coset_offset))
    basis = cpu_points(inp.keys[:m])
    ms["python_ints_msm_A"], a_cpu = cpu_ms(lambda: msm_pippenger(basis, inp.witness))
    ok = run.h == h_cpu and run.msm_results["A"].affine() == a_cpu.affine()
    scal = inp.msm_scalars(run.h)
    for name, size in inp.msm_sizes().items():
        e = fast.dot_mod_r(_ints_to_limbs(inp.keys[:size]), _ints_to_limbs(scal[name]))
        ok = ok and run.msm_results[name].affine() == (G * e).affine()
    results["groth16_demo"] = [{"constraints": m, "n": inp.domain.size, "verified": ok, "ms": ms}]
    log(f"suite demo {m} constraints: verified={ok} {ms}")
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text(json.dumps(results, indent=2) + "\n")
    log(f"suite written to {out}")


if __name__ == "__main__":
    sys.exit(main())