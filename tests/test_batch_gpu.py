"""
Batched MSMs over a shared SRS and batched NTTs (zkgpu.batch), plain and graph-replayed.

Every batched result must equal the same computation run alone (the oracle's
closed form for MSMs over k_i G, the oracle's ntt for transforms); a HIP
graph of the batch is replayed several times with eager work in between and
each replay is checked.
"""

import random

import numpy as np
import pytest

from zkgpu import gpu_backend as gb, workloads
from zkgpu.batch import MsmBatch, NttBatch, SRS, capture
from zkgpu.field256_gpu import BN254_FR
from zkref import fast
from zkref.curve import G
from zkref.field import R_SCALAR as R, ints_to_limbs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def srs():
    keys = workloads.srs_keys(1 << 13, 0x5EED04)
    return keys, SRS.from_keys(ints_to_limbs(keys))


def expected(keys, scalars):
    e = workloads.closed_form_scalar(keys, scalars)
    return None if e == 0 else (G * e).affine()


def aff(p):
    return None if p.is_identity() else p.affine()


def test_msm_batch_matches_closed_forms(srs):
    keys, basis = srs
    sizes = [1, 2, 7, 100, 1000, 4096, 8192, 3000]
    windows = [0, 1, 4, 0, 9, 13, 16, 0]
    scalars = [workloads.witness_scalars(n, 1000 + j) for j, n in enumerate(sizes)]
    batch = MsmBatch(basis, [gb.asarray(ints_to_limbs(s)) for s in scalars], windows=windows)
    batch.launch()
    got = batch.results()
    assert [aff(p) for p in got] == [expected(keys[:n], s) for n, s in zip(sizes, scalars)]


def test_msm_batch_edge_cases(srs):
    keys, basis = srs
    cases = [[0, 0, 0], [1], [R - 1, 1], [2**256 - 1, R, R + 5]]
    batch = MsmBatch(basis, [gb.asarray(ints_to_limbs(c)) for c in cases], chunk_len=2)
    batch.launch()
    res = batch.results()
    assert res[0].is_identity()
    assert res[1].affine() == (G * keys[0]).affine()
    e = (keys[0] * (R - 1) + keys[1]) % R
    assert res[2].is_identity() if e == 0 else res[2].affine() == (G * e).affine()
    assert aff(res[3]) == expected(keys[:3], [c % R for c in cases[3]])


def ntt_inputs(sizes, seed):
    rng = random.Random(seed)
    vals = [[rng.randrange(R) for _ in range(n)] for n in sizes]
    return vals, [BN254_FR.from_ints(v) for v in vals]


def test_ntt_batch_matches_oracle():
    sizes = [2, 16, 1024, 2048, 4096, 1 << 13, 1 << 16]
    inverse = [False, True, False, True, False, True, False]
    vals, rows = ntt_inputs(sizes, 5)
    outs = [gb.empty((n, 4)) for n in sizes]
    NttBatch(rows, outs, inverse).launch()
    for v, o, inv in zip(vals, outs, inverse):
        assert np.array_equal(BN254_FR.to_limbs(o).get(), fast.ntt_rows(ints_to_limbs(v), invert=inv))


def test_graph_replays_match_eager(srs):
    keys, basis = srs
    s = gb.Stream()
    sizes = [100, 4096, 8192]
    scalars = [workloads.witness_scalars(n, 77 + n) for n in sizes]
    with gb.stream(s.handle):
        msm = MsmBatch(basis, [gb.asarray(ints_to_limbs(x)) for x in scalars])
        vals, rows = ntt_inputs([1024, 1 << 14], 9)
        outs = [gb.empty((len(v), 4)) for v in vals]
        ntt = NttBatch(rows, outs, [False, True])
    s.synchronize()
    graph = capture([msm, ntt], s)
    want_msm = [expected(keys[:n], x) for n, x in zip(sizes, scalars)]
    want_ntt = [fast.ntt_rows(ints_to_limbs(vals[0])), fast.ntt_rows(ints_to_limbs(vals[1]), invert=True)]
    for rep in range(3):
        with gb.stream(s.handle):
            gb.zeros(outs[0].shape).copy()  # unrelated eager work between replays
            graph.replay()
        s.synchronize()
        assert [aff(p) for p in msm.results()] == want_msm, rep
        for o, w in zip(outs, want_ntt):
            assert np.array_equal(BN254_FR.to_limbs(o).get(), w), rep


@pytest.fixture(scope="module")
def tabled():
    keys = workloads.srs_keys(1 << 13, 0x5EED09)
    basis = SRS.from_keys(ints_to_limbs(keys)).precompute([8, 12])
    return keys, basis


def test_window_tables_match_closed_forms(tabled):
    """MSMs over an SRS with window tables (every window's digits in one set of buckets)."""
    keys, basis = tabled
    sizes = [1, 2, 7, 100, 1000, 4096, 8192, 3000]
    scalars = [workloads.witness_scalars(n, 2000 + j) for j, n in enumerate(sizes)]
    for windows in (None, [8, 12, 8, 12, 12, 8, 12, 0]):
        batch = MsmBatch(basis, [gb.asarray(ints_to_limbs(s)) for s in scalars], windows=windows)
        batch.launch()
        assert [aff(p) for p in batch.results()] == [expected(keys[:n], s) for n, s in zip(sizes, scalars)]


def test_window_tables_edge_scalars_and_mixed_jobs(tabled):
    keys, basis = tabled
    cases = [[0, 0, 0], [1], [R - 1, 1], [2**256 - 1, R, R + 5], [R - 1] * 500]
    # window 9 has no table: that job uses its own windows beside the table jobs
    batch = MsmBatch(basis, [gb.asarray(ints_to_limbs(c)) for c in cases], windows=[0, 12, 9, 8, 0], chunk_len=2)
    batch.launch()
    got = [aff(p) for p in batch.results()]
    assert got == [expected(keys[: len(c)], [x % R for x in c]) for c in cases]


def test_window_table_needs_no_live_plan(tabled):
    _, basis = tabled
    plan = MsmBatch(basis, [gb.asarray(ints_to_limbs([1, 2]))])
    with pytest.raises(gb.ZkError, match="before plans"):
        basis.precompute([16])
    del plan
    import gc

    gc.collect()


def cache_stats():
    import ctypes

    t, b, p = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    gb.call("zk_ntt_cache_stats", ctypes.byref(t), ctypes.byref(b), ctypes.byref(p))
    return t.value, b.value, p.value


def test_cache_clear_keeps_the_tables_of_live_plans():
    """Clearing the table cache (the reference's _ntt_tables.cache_clear) frees only tables no plan
    holds: a plan, and a graph recorded from it, still find theirs after a clear."""
    import gc

    from zkgpu.batch import QuotientBatch

    gb.call("zk_ntt_cache_clear")
    sizes = [1 << 12, 1 << 13]
    vals, rows = ntt_inputs(sizes, 31)
    outs = [gb.empty((n, 4)) for n in sizes]
    s = gb.Stream()
    with gb.stream(s.handle):
        plan = NttBatch(rows, outs, [False, True])
        qrows = [BN254_FR.from_limbs(workloads.uniform_rows(1 << 11, 90 + k)) for k in range(3)]
        qb = QuotientBatch([qrows[0]], [qrows[1]], [qrows[2]])
    s.synchronize()
    graph = capture([plan, qb], s)
    tables, _, pinned = cache_stats()
    assert pinned >= 7 and tables >= pinned
    gb.call("zk_ntt_cache_clear")
    assert cache_stats()[0] == cache_stats()[2] == pinned
    want = [fast.ntt_rows(ints_to_limbs(vals[0])), fast.ntt_rows(ints_to_limbs(vals[1]), invert=True)]
    plain = [BN254_FR.to_limbs(x).get() for x in qrows]
    want_h = fast.groth16_quotient(*plain, 7)
    for rep in range(2):
        with gb.stream(s.handle):
            graph.replay()
        s.synchronize()
        for o, w in zip(outs, want):
            assert np.array_equal(BN254_FR.to_limbs(o).get(), w), rep
        assert np.array_equal(BN254_FR.to_limbs(qb.h[0]).get(), want_h), rep
    del graph, plan, qb
    gc.collect()
    assert cache_stats()[2] == 0
    gb.call("zk_ntt_cache_clear")
    assert cache_stats()[:2] == (0, 0)
    # a one-shot transform after the clear builds its tables again
    got = BN254_FR.to_limbs(BN254_FR.ntt(rows[0])).get()
    assert np.array_equal(got, want[0])
