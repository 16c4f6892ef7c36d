"""
zkgpu.msm_gpu on the GPU: the MSM equals the oracle's (a unique group element).

Mirrors the reference's tests/test_msm.py GPU cases (windows 1..16, full-size
scalars, buckets where P and -P cancel or P doubles, reduction mod r, heavy
buckets, chunked sums with chunk lengths 2, 3 and 32, boolean scalars) and
adds the benchmark shapes: the witness-like 2^16 and 2^20 MSMs over a basis
P_i = k_i G, checked against the closed form (sum s_i k_i mod r) G.
"""

import random
import sys

import numpy as np
import pytest

from zkgpu import gpu_backend, gpu_backend as gb, msm_gpu, workloads
from zkgpu.curve import Point as gPoint
from zkgpu.field256_gpu import BN254_FQ
from zkgpu.msm_gpu import msm_g1, msm_pippenger_gpu
from zkref import fast
from zkref.curve import G, P, Point
from zkref.field import R_SCALAR as R, ints_to_limbs, limbs_to_ints
from zkref.msm import msm_double_and_add, msm_naive, msm_pippenger

pytestmark = pytest.mark.gpu


def pippenger(points, scalars, window_size=None):
    px = np.array([p.x for p in points], dtype=object)
    py = np.array([p.y for p in points], dtype=object)
    got = msm_pippenger_gpu(px, py, np.array(scalars, dtype=object), window_size=window_size)
    return None if got.is_identity() else got.affine()


def aff(p):
    return None if p.is_identity() else p.affine()


@pytest.fixture(scope="module")
def msm_inputs():
    rng = random.Random(29)
    points = [G * rng.randint(1, 2**32) for _ in range(30)]
    scalars = [rng.randint(0, 2**32) for _ in range(30)]
    return points, scalars, msm_naive(points, scalars)


@pytest.fixture(scope="module")
def large_inputs():
    rng = random.Random(37)
    step = G * rng.randrange(1, R)
    points = [G * rng.randrange(1, R)]
    for _ in range(399):
        points.append(points[-1] + step)
    scalars = [rng.randrange(R) for _ in points]
    return points, scalars, msm_double_and_add(points, scalars)


@pytest.mark.parametrize("w", [None, 1, 3, 8, 11, 16])
def test_matches_naive(w, msm_inputs):
    points, scalars, expected = msm_inputs
    assert pippenger(points, scalars, w) == aff(expected)


@pytest.mark.parametrize("w", [3, 5, 8])
def test_full_size_scalars(w):
    rng = random.Random(31)
    points = [G * rng.randint(1, 2**32) for _ in range(12)]
    scalars = [rng.randrange(R) for _ in range(12)]
    assert pippenger(points, scalars, w) == aff(msm_double_and_add(points, scalars))


def test_bucket_edge_cases():
    assert pippenger([G, -G], [7, 7]) is None
    assert pippenger([G, G], [5, 5]) == (G * 10).affine()
    assert pippenger([G, G], [0, 0]) is None
    point = G * 3
    for s in (-5, R, R + 7, 2**300 + 1, R - 1, (R - 1) // 2, (R + 1) // 2):
        assert pippenger([point], [s]) == aff(point * (s % R))


def test_returns_a_point():
    assert msm_pippenger_gpu([], [], []) == gPoint.identity()
    res = msm_pippenger_gpu([G.x], [G.y], [R])
    assert isinstance(res, gPoint) and res.is_identity()
    assert msm_pippenger_gpu([G.x, G.x], [G.y, G.y], [2, 3]).affine() == (G * 5).affine()


@pytest.mark.parametrize("w", [2, 6, 9, 13, 16])
def test_at_scale(w, large_inputs):
    points, scalars, expected = large_inputs
    assert pippenger(points, scalars, w) == aff(expected)


def test_one_heavy_bucket_per_window():
    points = [G * k for k in range(1, 201)]
    scalar = 2**200 + 12345
    assert pippenger(points, [scalar] * 200) == (G * (sum(range(1, 201)) * scalar % R)).affine()


@pytest.mark.parametrize("chunk_len", [2, 3, 32])
@pytest.mark.parametrize("kind", ["boolean", "uniform"])
def test_chunked_sums(kind, chunk_len, large_inputs, monkeypatch):
    points, scalars, expected = large_inputs
    if kind == "boolean":
        rng = random.Random(41)
        scalars = [rng.randrange(2) for _ in points]
        expected = msm_double_and_add(points, scalars)
    monkeypatch.setattr(msm_gpu, "_CHUNK_LEN", chunk_len)
    assert pippenger(points, scalars, 9) == aff(expected)


def test_off_curve_points_are_rejected_on_the_device():
    with pytest.raises(ValueError, match=r"\(1, 3\) is not on the BN254 curve"):
        msm_pippenger_gpu([G.x, 1], [G.y, 3], [5, 6])


def test_self_test_failure_is_not_cached(monkeypatch):
    monkeypatch.setattr(msm_gpu, "_compiled", {})
    real = msm_gpu._result
    monkeypatch.setattr(msm_gpu, "_result", lambda rows: -real(rows))
    with pytest.raises(RuntimeError, match="failed their self-test"):
        msm_gpu.compile_kernels()
    assert msm_gpu._compiled == {}


def basis(n, seed):
    keys = workloads.srs_keys(n, seed)
    pts = workloads.srs_points(ints_to_limbs(keys))
    return keys, pts


def test_fixed_base_points_match_the_oracle():
    keys, pts = basis(64, 7)
    host = limbs_to_ints(BN254_FQ.to_limbs(pts.reshape(-1, 4)).get())
    assert [(host[2 * i], host[2 * i + 1]) for i in range(64)] == [(G * k).affine() for k in keys]


@pytest.mark.parametrize("w", [None, 10, 12, 13])
def test_config2_witness_msm_2_16(w):
    """2^16 points k_i G and the witness-like scalar mix, against the closed form and the oracle's C Pippenger."""
    from zkref import fast

    n = 1 << 16
    keys, pts = basis(n, 0x5EED02)
    scalars = workloads.witness_scalars(n, 0x5EED02)
    expected = (G * workloads.closed_form_scalar(keys, scalars)).affine()
    got = msm_g1(pts, gpu_backend.asarray(ints_to_limbs(scalars)), window_size=w)
    assert got.affine() == expected
    if w is None:
        plain = BN254_FQ.to_limbs(pts.reshape(-1, 4)).get().reshape(n, 2, 4)
        assert fast.msm_pippenger(plain, ints_to_limbs(scalars), 12) == expected


def test_msm_2_20_closed_form():
    n = 1 << 20
    keys, pts = basis(n, 0x5EED04)
    scalars = workloads.witness_scalars(n, 0x5EED04)
    got = msm_g1(pts, gpu_backend.asarray(ints_to_limbs(scalars)))
    assert got.affine() == (G * workloads.closed_form_scalar(keys, scalars)).affine()


@pytest.mark.parametrize("glv", [1, 0])
def test_glv_split_on_and_off(glv, large_inputs):
    """The same MSMs with the GLV split and with plain 254-bit digits: scalars that sit on the
    split's edges (lambda, r - lambda, r - 1, >= r, 2^127 +- 1, small values) and random ones."""
    import ctypes

    lam = 0xB3C4D79D41A917585BFC41088D8DAAA78B17EA66B99C90DD
    points, _, _ = large_inputs
    rng = random.Random(41)
    edge = [lam, R - lam, R - 1, R, R + 7, 2**256 - 1, 2**127, 2**127 - 1, 2**127 + 1, 1, 0, 2**32 - 1, lam + 1]
    scalars = edge + [rng.randrange(R) for _ in range(len(points) - len(edge))]
    expected = msm_double_and_add(points, [s % R for s in scalars])
    old = ctypes.c_int()
    gpu_backend.call("zk_msm_get_glv", ctypes.byref(old))
    gpu_backend.call("zk_msm_set_glv", glv)
    try:
        for w in (None, 4, 9, 13):
            assert pippenger(points, scalars, w) == aff(expected), (glv, w)
        # all scalars on one point: every digit stream lands in a handful of buckets
        same = [points[0]] * 64
        s2 = [rng.randrange(R) for _ in same]
        assert pippenger(same, s2) == aff(points[0] * (sum(s2) % R))
    finally:
        gpu_backend.call("zk_msm_set_glv", old.value)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_accumulation_variants(variant):
    """Every bucket accumulation kernel (Jacobian / XYZZ, serial / low-latency products)."""
    import ctypes

    n = 3000
    keys = workloads.srs_keys(n, 0x5EED0C)
    scalars = workloads.witness_scalars(n, 0x5EED0D) + []
    scalars[:4] = [R - 1, 1, 1, 0]  # repeated small digits: long runs in few buckets
    pts = workloads.srs_points(ints_to_limbs(keys))
    old = ctypes.c_int()
    gpu_backend.call("zk_msm_get_accumulate", ctypes.byref(old))
    gpu_backend.call("zk_msm_set_accumulate", variant)
    try:
        e = workloads.closed_form_scalar(keys, scalars)
        for chunk in (0, 2, 7):
            out = gpu_backend.empty((3, 4))
            msm_gpu.msm_g1_jacobian(pts, gpu_backend.asarray(ints_to_limbs(scalars)), out, chunk_len=chunk)
            got = msm_gpu._result(out)
            assert (got.is_identity() and e == 0) or got.affine() == (G * e).affine(), (variant, chunk)
    finally:
        gpu_backend.call("zk_msm_set_accumulate", old.value)


@pytest.mark.parametrize("log_seg", [1, 3, 4])
def test_reduction_segment_lengths(log_seg):
    """The bucket reduction with other segment lengths (fewer or more bit rows per window), over
    windows small enough that a segment covers a window's buckets, and over tuned ones."""
    import ctypes

    n = 3000
    keys = workloads.srs_keys(n, 0x5EED10)
    scalars = workloads.witness_scalars(n, 0x5EED11)
    scalars[:3] = [R - 1, 2**127 - 1, 0]
    pts = workloads.srs_points(ints_to_limbs(keys))
    e = workloads.closed_form_scalar(keys, scalars)
    old = ctypes.c_int()
    gpu_backend.call("zk_msm_get_segment", ctypes.byref(old))
    gpu_backend.call("zk_msm_set_segment", log_seg)
    try:
        for w in (0, 3, 6, 12):
            out = gpu_backend.empty((3, 4))
            msm_gpu.msm_g1_jacobian(pts, gpu_backend.asarray(ints_to_limbs(scalars)), out, window_size=w)
            got = msm_gpu._result(out)
            assert (got.is_identity() and e == 0) or got.affine() == (G * e).affine(), (log_seg, w)
    finally:
        gpu_backend.call("zk_msm_set_segment", old.value)


@pytest.mark.parametrize("w", [16, 17, 20])
def test_windows_past_16_bits(w):
    """Windows up to 16 bits store 16-bit digit codes for the sort, wider ones (the device entry
    point allows 20) 32-bit codes: both widths at their edges, with the largest digits a scalar can
    give in the top window (2^127 - 1 and r - 1 split to halves with every window digit set)."""
    n = 3000
    keys = workloads.srs_keys(n, 0x5EED12)
    scalars = workloads.witness_scalars(n, 0x5EED13)
    scalars[:4] = [R - 1, 2**127 - 1, 0, 2**126 + 2**110 - 1]
    pts = workloads.srs_points(ints_to_limbs(keys))
    e = workloads.closed_form_scalar(keys, scalars)
    out = gpu_backend.empty((3, 4))
    msm_gpu.msm_g1_jacobian(pts, gpu_backend.asarray(ints_to_limbs(scalars)), out, window_size=w)
    assert msm_gpu._result(out).affine() == (G * e).affine()


@pytest.mark.parametrize("chunk", [0, 2, 5])
def test_one_bucket_spans_every_range(chunk):
    """Scalars that send (almost) every entry to one bucket: the bucket's run crosses every range
    boundary of the accumulation, so its fragments go through both boundary passes and every
    level of the run-summary trees (with short ranges, several levels of them)."""
    n = 1 << 15
    keys = workloads.srs_keys(n, 0x5EED0E)
    pts = workloads.srs_points(ints_to_limbs(keys))
    rng = random.Random(43)
    cases = [[1] * n, [1 if rng.random() < 0.9 else rng.randrange(R) for _ in range(n)]]
    for scalars in cases:
        e = workloads.closed_form_scalar(keys, scalars)
        out = gpu_backend.empty((3, 4))
        msm_gpu.msm_g1_jacobian(pts, gpu_backend.asarray(ints_to_limbs(scalars)), out, chunk_len=chunk)
        got = msm_gpu._result(out)
        assert (got.is_identity() and e == 0) or got.affine() == (G * e).affine(), chunk


@pytest.mark.parametrize("chunk", [0, 2])
def test_run_sums_that_cancel_across_ranges(chunk):
    """A basis whose second half negates its first (P_{i+h} = -P_i) under equal scalars: each
    window's bucket sums to the identity, so the last merges of the run-summary trees add a partial
    sum X to -X, and the identity they produce must read as the identity to every later stage."""
    n = 1 << 14
    rng = random.Random(47)
    half = [rng.randrange(1, R) for _ in range(n // 2)]
    keys = half + [R - k for k in half]
    pts = workloads.srs_points(ints_to_limbs(keys))
    s = rng.randrange(1, R)
    cases = [[1] * n, [s] * n, [s] * (n - 1) + [rng.randrange(1, R)]]
    for scalars in cases:
        e = workloads.closed_form_scalar(keys, scalars)
        out = gpu_backend.empty((3, 4))
        msm_gpu.msm_g1_jacobian(pts, gpu_backend.asarray(ints_to_limbs(scalars)), out, chunk_len=chunk)
        got = msm_gpu._result(out)
        assert (got.is_identity() and e == 0) or got.affine() == (G * e).affine(), (chunk, scalars[-1])


@pytest.fixture(scope="module")
def basis20():
    n = 1 << 20
    keys = workloads.key_rows(n, 0x5EED04)
    points = workloads.srs_points(keys)
    scalars = workloads.witness_rows_fast(n, 0x5EED31)
    e = fast.dot_mod_r(keys, scalars)
    return points, scalars, (None if e == 0 else (G * e).affine())


def test_public_call_at_2p20_with_python_ints(basis20):
    """msm_pippenger_gpu with Python ints at 2^20 (bulk conversion, device reduction and on-curve
    check) equals the closed form (sum k_i s_i mod r) G."""
    import time

    points, scalars, want = basis20
    coords = BN254_FQ.to_ints(points.reshape(-1, 4))
    xs, ys = coords[0::2], coords[1::2]
    sc = limbs_to_ints(scalars)
    t0 = time.perf_counter()
    got = msm_pippenger_gpu(xs, ys, sc)
    print(f"public call, 2^20 Python ints: {(time.perf_counter() - t0) * 1e3:.1f} ms", file=sys.stderr, flush=True)
    assert got.affine() == want


def test_public_call_with_resident_points(basis20):
    """The same signature over device-resident inputs: an SRS (with and without window tables) or
    (n, 2, 4) rows in HBM as points_x, scalars as device or NumPy rows."""
    from zkgpu.batch import SRS

    points, scalars, want = basis20
    dev_sc = gb.asarray(scalars)
    assert msm_pippenger_gpu(points, None, dev_sc).affine() == want
    assert msm_pippenger_gpu(points, None, scalars).affine() == want
    srs = SRS(points)
    assert msm_pippenger_gpu(srs, None, dev_sc).affine() == want
    srs.precompute((16,))
    assert msm_pippenger_gpu(srs, None, dev_sc).affine() == want
    # an MSM over a prefix of the SRS
    keys = workloads.key_rows(1000, 0x5EED04)
    e = fast.dot_mod_r(keys, scalars[:1000])
    assert msm_pippenger_gpu(srs, None, dev_sc[:1000]).affine() == (G * e).affine()
    with pytest.raises(ValueError, match="Length mismatch"):
        msm_pippenger_gpu(points[:10], None, dev_sc[:11])
    with pytest.raises(TypeError, match="points_y must be None"):
        msm_pippenger_gpu(points, [1], dev_sc)
    assert msm_pippenger_gpu(srs, None, dev_sc[:0]).is_identity()


def test_resident_points_off_the_curve_are_refused():
    bad = np.zeros((3, 2, 4), dtype=np.uint64)
    pts = BN254_FQ.from_limbs(bad.reshape(-1, 4)).reshape(3, 2, 4)  # (0, 0) is not on the curve
    with pytest.raises(ValueError, match=r"\(0, 0\) is not on the BN254 curve"):
        msm_pippenger_gpu(pts, None, gb.asarray(np.ones((3, 4), dtype=np.uint64)))


def test_public_int_path_checks_in_the_reference_order():
    # integer types of x, then y, then the scalars, then the lengths (msm_gpu.py:432-439)
    with pytest.raises(TypeError, match="points_x must contain integers"):
        msm_pippenger_gpu([G.x, 1.5], [G.y], [1, 2, 3])
    with pytest.raises(TypeError, match="points_y must contain integers"):
        msm_pippenger_gpu([G.x], [G.y, "a"], [1, 2, 3])
    with pytest.raises(TypeError, match="scalars must contain integers"):
        msm_pippenger_gpu([G.x], [G.y, G.y], [1.0])
    with pytest.raises(ValueError, match="Length mismatch: 1 x-coordinates, 2 y-coordinates, 3 scalars"):
        msm_pippenger_gpu([G.x], [G.y, G.y], [1, 2, 3])
    with pytest.raises(ValueError, match="Length mismatch: 2 x-coordinates, 2 y-coordinates, 1 scalars"):
        msm_pippenger_gpu([G.x, G.x], [G.y, G.y], np.zeros((1, 4), dtype=np.uint64))


def test_public_int_path_order_across_conversion_parts():
    """Long inputs are converted and uploaded in parts of at least 2^16 points, x and y part by
    part: a bad y in the first part and a bad x in the second still report x first, as the
    reference checks all of x before y; then a bad scalar after good points, and a clean call."""
    n = (1 << 17) + 5
    xs, ys = [G.x] * n, [G.y] * n
    bad_x, bad_y = list(xs), list(ys)
    bad_x[100000] = 2.5
    bad_y[10] = "y"
    with pytest.raises(TypeError, match="points_x must contain integers"):
        msm_pippenger_gpu(bad_x, bad_y, [1] * n)
    with pytest.raises(TypeError, match="points_y must contain integers"):
        msm_pippenger_gpu(xs, bad_y, [1] * n)
    sc = [1] * n
    sc[n - 1] = None
    with pytest.raises(TypeError, match="scalars must contain integers"):
        msm_pippenger_gpu(xs, ys, sc)
    sc[n - 1] = 3
    assert msm_pippenger_gpu(xs, ys, sc).affine() == (G * (n + 2)).affine()


def test_public_int_path_input_kinds():
    # generators, object arrays, unreduced and negative coordinates, scalars of every range, and
    # scalars as plain uint64 limbs: always (sum s_i k_i mod r) G
    keys = [3, 5, 7, 11, 13]
    pts = [G * k for k in keys]
    scalars = [2**300 + 1, -9, R + 4, True, 2**256 - 1]
    expected = (G * (sum(s * k for s, k in zip(scalars, keys)) % R)).affine()
    xs = [p.x + (P if i % 2 else 0) for i, p in enumerate(pts)]  # below 2^256, reduced on the device
    ys = [p.y - (P if i % 3 == 0 else 0) for i, p in enumerate(pts)]  # negative: reduced on the host
    got = msm_pippenger_gpu((x for x in xs), np.array(ys, dtype=object), iter(scalars))
    assert got.affine() == expected
    limbs = ints_to_limbs([s % R for s in scalars])
    assert msm_pippenger_gpu(xs, ys, limbs).affine() == expected
    assert msm_pippenger_gpu(xs, ys, gb.asarray(limbs)).affine() == expected

