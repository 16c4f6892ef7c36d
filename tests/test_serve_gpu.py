"""
The batched Groth16 quotient and the proof server on the GPU.

QuotientBatch must equal the C oracle's quotient for every job of a mixed-size batch
(bit-exact rows) and replay identically from a HIP graph. The server must give every job the
same results whatever batch it lands in: each job of a mixed run is recomputed alone through
the per-call path (one quotient call and three MSM calls) and compared bit for bit (h) and as
group elements (MSMs); one job is also pinned to the oracle's quotient and to the MSM's closed
form over the basis k_i G.
"""

import numpy as np
import pytest

from zkgpu import gpu_backend as gb, workloads
from zkgpu.batch import SRS, MsmBatch, QuotientBatch, capture, msm_workspace_bytes, quotient_workspace_bytes, tuned_chunks
from zkgpu.field256_gpu import BN254_FR
from zkgpu.serve import GpuExecutor, ProofServer, WitnessPools, run_plain, served_workload
from zkref import fast
from zkref.curve import G
from zkref.field import ints_to_limbs, limbs_to_ints

pytestmark = pytest.mark.gpu


def aff(p):
    return None if p is None or p.is_identity() else p.affine()


def quotient_inputs(sizes, seed):
    plain = [[workloads.uniform_rows(n, seed + 3 * j + k) for k in range(3)] for j, n in enumerate(sizes)]
    rows = [[BN254_FR.from_limbs(p) for p in trip] for trip in plain]
    return plain, rows


def test_quotient_batch_matches_oracle():
    sizes = [2, 16, 1024, 2048, 1 << 13, 1 << 16, 1 << 13]
    shifts = [7, 7, 5, 7, 11, 7, 7]
    plain, rows = quotient_inputs(sizes, 0x5EED40)
    qb = QuotientBatch([r[0] for r in rows], [r[1] for r in rows], [r[2] for r in rows], shifts=shifts)
    qb.launch()
    for j, (n, g) in enumerate(zip(sizes, shifts)):
        got = BN254_FR.to_limbs(qb.h[j]).get()
        assert np.array_equal(got, fast.groth16_quotient(*plain[j], g)), (j, n)


def test_quotient_batch_graph_replays():
    s = gb.Stream()
    sizes = [1 << 11, 1 << 14, 256]
    plain, rows = quotient_inputs(sizes, 0x5EED50)
    with gb.stream(s.handle):
        qb = QuotientBatch([r[0] for r in rows], [r[1] for r in rows], [r[2] for r in rows])
    s.synchronize()
    graph = capture([qb], s)
    want = [fast.groth16_quotient(*p, 7) for p in plain]
    for rep in range(3):
        with gb.stream(s.handle):
            for h in qb.h:
                gb.zeros(h.shape).copy()  # unrelated eager work between replays
            graph.replay()
        s.synchronize()
        for h, w in zip(qb.h, want):
            assert np.array_equal(BN254_FR.to_limbs(h).get(), w), rep


def test_workspace_sizes_grow_with_the_job():
    small = quotient_workspace_bytes([12]) + msm_workspace_bytes([1 << 12] * 3)
    large = quotient_workspace_bytes([18]) + msm_workspace_bytes([1 << 18] * 3)
    assert 0 < small < large
    assert quotient_workspace_bytes([18]) >= 3 * (1 << 18) * 32


@pytest.fixture(scope="module")
def served():
    keys = workloads.srs_keys(1 << 14, 0x5EED07)
    srs = SRS.from_keys(ints_to_limbs(keys))
    pools = WitnessPools((1 << 14) + 129, seed=0x5EED08)
    return keys, srs, pools


def check_against_alone(jobs, srs, pools):
    for j in jobs:
        alone = pools.job(j.job_id, j.log_n, 0.0)
        run_plain(alone, srs.points)
        assert np.array_equal(j.h.get(), alone.h.get()), j.job_id
        assert [aff(p) for p in j.msm_results()] == [aff(p) for p in alone.msm_results()], j.job_id


def test_served_results_equal_alone(served):
    keys, srs, pools = served
    jobs = served_workload(pools, 24, None, lo=8, hi=14, seed=11)
    server = ProofServer(srs, hbm_budget=1 << 31, slots=2, max_batch_jobs=5)
    rep = server.serve(jobs)
    assert rep.batches >= 5 and max(rep.batch_sizes) <= 5
    check_against_alone(jobs, srs, pools)

    # one job pinned to the oracle: its quotient and the closed form of its first MSM
    j = max(jobs, key=lambda x: x.n)
    plain = [BN254_FR.to_limbs(x).get() for x in (j.a, j.b, j.c)]
    assert np.array_equal(BN254_FR.to_limbs(j.h).get(), fast.groth16_quotient(*plain, j.shift))
    scalars = limbs_to_ints(j.scalars[0].get())
    e = workloads.closed_form_scalar(keys[: j.n], scalars)
    assert aff(j.msm_results()[0]) == (None if e == 0 else (G * e).affine())


def test_poisson_serving_mixes_batches_and_matches_alone(served):
    _, srs, pools = served
    jobs = served_workload(pools, 40, 4000.0, lo=8, hi=14, seed=12)
    server = ProofServer(srs, hbm_budget=1 << 31, slots=3, max_batch_jobs=16)
    rep = server.serve(jobs)
    s = rep.summary()
    assert s["jobs"] == 40 and s["p50_ms"] <= s["p99_ms"]
    assert all(j.arrival <= j.admitted <= j.completed for j in jobs)
    check_against_alone(jobs, srs, pools)


def test_budget_limits_inflight_work(served):
    _, srs, pools = served
    jobs = served_workload(pools, 12, None, lo=13, hi=14, seed=13)
    server = ProofServer(srs, hbm_budget=1, slots=2)
    with pytest.raises(ValueError, match="more device memory"):
        server.serve(jobs)
    one = max(server.executor.footprint(j) for j in jobs)
    server = ProofServer(srs, hbm_budget=2 * one, slots=4, max_batch_jobs=8)
    rep = server.serve(jobs)
    assert rep.peak_bytes <= 2 * one
    check_against_alone(jobs[:4], srs, pools)


def test_served_over_window_tables_equal_alone():
    """A server over an SRS with window tables; each job checked against its per-call run."""
    keys = workloads.srs_keys(1 << 13, 0x5EED0A)
    srs = SRS.from_keys(ints_to_limbs(keys)).precompute([12])
    pools = WitnessPools((1 << 13) + 77, seed=0x5EED0B)
    jobs = served_workload(pools, 10, None, lo=8, hi=13, seed=14)
    rep = ProofServer(srs, hbm_budget=1 << 31, slots=2, max_batch_jobs=4).serve(jobs)
    assert rep.batches >= 3
    check_against_alone(jobs, srs, pools)


def test_footprint_covers_what_a_batch_allocates():
    """Admission charges each job at least what it adds to any batch: over an SRS with and
    without window tables, for every accumulation chunk a batch may use, the jobs' footprints
    sum to no less than the batch's MSM plan, quotient workspace and h rows."""
    keys = workloads.srs_keys(1 << 13, 0x5EED0C)
    for tables in ([], [12]):
        srs = SRS.from_keys(ints_to_limbs(keys)).precompute(tables)
        ex = GpuExecutor(srs, slots=1)
        pools = WitnessPools((1 << 13) + 5, seed=0x5EED0D)
        for sizes in ([8, 13], [8, 8, 9, 13, 10], [11], [13, 13, 13]):
            jobs = [pools.job(i, k, 0.0) for i, k in enumerate(sizes)]
            charged = sum(ex.footprint(j) for j in jobs)
            for chunk in tuned_chunks():
                m = MsmBatch(srs, [s for j in jobs for s in j.scalars], windows=[0] * (3 * len(jobs)), chunk_len=chunk)
                need = m.nbytes + quotient_workspace_bytes(sizes) + sum(j.n * 32 for j in jobs)
                assert charged >= need, (tables, sizes, chunk, charged, need)
                del m
