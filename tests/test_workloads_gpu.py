"""
The batched and served forms at their stated workloads, against the oracle.

- Batched: 64 G1 MSMs with sizes log-uniform in [2^10, 2^20] over prefixes of one
  2^20-point basis P_i = k_i G (seed 0x5EED04) carrying window tables of widths 12 and
  16, witness-like scalars (a seed per MSM), plus 32 forward Fr NTTs of 2^10..2^20 rows,
  captured in one HIP graph and replayed twice with eager work in between. Every MSM
  of every replay equals the closed form (sum k_i s_i mod r) G, and every NTT equals
  the C oracle's ntt.
- Served: 2000 Groth16 jobs (quotient at size n with coset shift 7 + three G1 MSMs of
  size n over the same basis), n = 2^k with k uniform in {12..20}, Poisson arrivals at a
  rate that keeps the server loaded (so continuous batching engages). Every job's h and
  MSM results equal its solo run through the per-call path bit for bit (h rows) and as
  group elements (MSMs), and one job of each size class is pinned to the oracle's
  quotient and to the MSM closed form.

The oracle runs in worker threads (its C calls release the interpreter lock) while the
GPU works. The "same alone and in the mix" model is the reference's agreement check
(demo/prover_simulation.py:235-245).
"""

import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from zkgpu import gpu_backend as gb, workloads
from zkgpu.batch import SRS, MsmBatch, NttBatch, capture
from zkgpu.field256_gpu import BN254_FR
from zkgpu.serve import ProofServer, WitnessPools, run_plain, served_workload
from zkref import fast
from zkref.curve import G

pytestmark = pytest.mark.gpu

SERVE_RATE = 1000.0  # jobs/s: above the per-call path's capacity, below the server's


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def aff(p):
    return None if p is None or p.is_identity() else p.affine()


def closed_form(e):
    return None if e == 0 else (G * e).affine()


@pytest.fixture(scope="module")
def srs20():
    say("srs: 2^20 points k_i G (seed 0x5EED04) with window tables 12 and 16")
    keys = workloads.key_rows(1 << workloads.BATCHED_SRS_LOG, workloads.BATCHED_SRS_SEED)
    srs = SRS.from_keys(keys).precompute((12, 16))
    gb.synchronize()
    return keys, srs


def test_batched_workload_at_size(srs20):
    keys, srs = srs20
    msm_sizes, ntt_sizes = workloads.batched_workload()
    assert len(msm_sizes) == 64 and len(ntt_sizes) == 32
    assert min(msm_sizes) >= 1 << 10 and max(msm_sizes) <= 1 << 20
    say(f"batched: 64 MSMs ({sum(msm_sizes)} points) + 32 NTTs ({sum(ntt_sizes)} rows)")
    scal_host = [workloads.batched_msm_scalars(j, n) for j, n in enumerate(msm_sizes)]
    ntt_host = [workloads.batched_ntt_rows(i, n) for i, n in enumerate(ntt_sizes)]
    with ThreadPoolExecutor(8) as pool:
        dots = [pool.submit(fast.dot_mod_r, keys[:n], s) for n, s in zip(msm_sizes, scal_host)]
        ntts = [pool.submit(fast.ntt_rows, x) for x in ntt_host]
        st = gb.Stream()
        with gb.stream(st.handle):
            scal = [gb.asarray(s) for s in scal_host]
            ntt_in = [BN254_FR.from_limbs(x) for x in ntt_host]
            ntt_out = [gb.empty((n, 4)) for n in ntt_sizes]
            msm = MsmBatch(srs, scal)
            ntt = NttBatch(ntt_in, ntt_out, [False] * len(ntt_sizes))
        st.synchronize()
        graph = capture([msm, ntt], st)
        want_msm = [closed_form(d.result()) for d in dots]
        want_ntt = [f.result() for f in ntts]
    for rep in range(2):
        with gb.stream(st.handle):
            for o in ntt_out:
                gb.zeros(o.shape).copy()  # unrelated eager work between replays
            graph.replay()
        st.synchronize()
        got = [aff(p) for p in msm.results()]
        bad = [j for j in range(64) if got[j] != want_msm[j]]
        assert not bad, f"replay {rep}: MSMs {bad} (sizes {[msm_sizes[j] for j in bad]}) differ from the closed form"
        for i, (o, w) in enumerate(zip(ntt_out, want_ntt)):
            assert np.array_equal(BN254_FR.to_limbs(o).get(), w), f"replay {rep}: NTT {i} (n = {ntt_sizes[i]})"
        say(f"batched: replay {rep} matches the oracle")


def test_served_workload_at_size(srs20):
    keys, srs = srs20
    pools = WitnessPools((1 << 20) + (1 << 16), seed=0x5EED05)
    server = ProofServer(srs, hbm_budget=64 << 30, slots=2, max_batch_jobs=64)
    say("served: warm-up (tables for every size)")
    server.serve([pools.job(100000 + k, k, 0.0) for k in range(12, 21)])
    gb.synchronize()
    jobs = served_workload(pools, 2000, SERVE_RATE, seed=0x5EED06)
    assert sorted({j.log_n for j in jobs}) == list(range(12, 21))
    say(f"served: 2000 jobs, Poisson arrivals at {SERVE_RATE} jobs/s")
    rep = server.serve(jobs)
    s = rep.summary()
    say(f"served: {s}")
    assert s["jobs"] == 2000 and all(j.arrival <= j.admitted <= j.completed for j in jobs)
    assert s["mean_batch_jobs"] > 1.2, "continuous batching did not engage at this rate"

    # one job per size class against the oracle: quotient rows and the closed forms of its MSMs
    firsts = {}
    for j in jobs:
        firsts.setdefault(j.log_n, j)
    with ThreadPoolExecutor(9) as pool:
        checks = []
        for k, j in sorted(firsts.items()):
            plain = [BN254_FR.to_limbs(x).get() for x in (j.a, j.b, j.c)]
            quo = pool.submit(fast.groth16_quotient, *plain, j.shift)
            dots = [pool.submit(fast.dot_mod_r, keys[: j.n], sc.get()) for sc in j.scalars]
            checks.append((j, BN254_FR.to_limbs(j.h).get(), quo, dots))
        for j, got_h, quo, dots in checks:
            assert np.array_equal(got_h, quo.result()), f"job {j.job_id} (n = 2^{j.log_n}): quotient"
            assert [aff(p) for p in j.msm_results()] == [closed_form(d.result()) for d in dots], \
                f"job {j.job_id} (n = 2^{j.log_n}): MSMs"
    say("served: one job per size class matches the oracle")

    # every job alone through the per-call path (one quotient call, three MSM calls)
    for i, j in enumerate(jobs):
        alone = pools.job(j.job_id, j.log_n, 0.0)
        run_plain(alone, srs.points)
        assert np.array_equal(j.h.get(), alone.h.get()), f"job {j.job_id} (n = 2^{j.log_n}): h differs from its solo run"
        assert [aff(p) for p in j.msm_results()] == [aff(p) for p in alone.msm_results()], \
            f"job {j.job_id} (n = 2^{j.log_n}): MSMs differ from the solo run"
        j.h = None
        if i % 500 == 499:
            say(f"served: {i + 1} jobs equal their solo runs")
