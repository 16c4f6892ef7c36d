"""
The drop-in API's host-side behaviour, checked without a GPU.

Validation order and messages follow the reference (field256_gpu.py:250-291,
347-368; curve_gpu.py:354-439; msm_gpu.py:385-446; curve.py:52-179). Without a
GPU the compute entry points must raise RuntimeError: the package has no CPU
fallback.
"""

import numpy as np
import pytest

from zkgpu import curve, field256_gpu, gpu_backend, msm_gpu
from zkgpu.curve import JACOBIAN_IDENTITY, G, P, Point, from_jacobian, to_jacobian
from zkgpu.field import FIELD_MODULUS as R
from zkgpu.field256_gpu import BN254_FQ, BN254_FR, Field256
from zkgpu.msm import _check_window_size, pippenger_window_size
from zkgpu.ntt import get_ntt_root, quadratic_non_residue

no_gpu = pytest.mark.skipif(gpu_backend.GPU_AVAILABLE, reason="checks the no-GPU behaviour")


def test_constants_match_published_values():
    assert BN254_FR._n_inv == 0xC2E1F593EFFFFFFF
    assert BN254_FQ._n_inv == 0x87D20782E4866389
    for f in (BN254_FR, BN254_FQ):
        assert f._n_inv * f.modulus % 2**64 == 2**64 - 1
        assert f._r2 == pow(2, 512, f.modulus)
    assert repr(Field256(3)) == "Field256(0x3)"


def test_source_carries_the_constants():
    src = BN254_FR.cuda_source()
    assert "#define ZK_NINV 0xc2e1f593efffffffULL" in src
    assert "#define ZK_P0 0x43e1f593f0000001ULL" in src
    assert "#define ZK_ONE0 0xac96341c4ffffffbULL" in src
    assert src.index("#define") < src.index("mont_raw")


@pytest.mark.parametrize("modulus", [2, 1, 0, -7, 2**64, 2**255 + 1, 2**256 - 189])
def test_invalid_moduli(modulus):
    with pytest.raises(ValueError, match="odd and in"):
        Field256(modulus)


def test_float_modulus():
    with pytest.raises(TypeError):
        Field256(7.0)


@pytest.mark.parametrize(
    "bad",
    [np.zeros((3, 4), dtype=np.int64), np.zeros((3, 3), dtype=np.uint64), np.zeros(12, dtype=np.uint64), [[0, 0, 0, 0]]],
    ids=["int64", "3-limbs", "flat", "list"],
)
def test_malformed_rows(bad):
    with pytest.raises(TypeError, match=r"\(n, 4\) uint64"):
        BN254_FR.mul(bad, np.zeros((1, 4), dtype=np.uint64))


def test_broadcast_mismatch():
    with pytest.raises(ValueError, match="broadcast"):
        BN254_FR.mul(np.zeros((2, 4), dtype=np.uint64), np.zeros((3, 4), dtype=np.uint64))


@pytest.mark.parametrize(
    ("field", "n", "message"),
    [(BN254_FR, 6, "power of 2"), (BN254_FQ, 4, "not divisible"), (Field256(2**255 - 1), 2, "odd prime")],
)
def test_ntt_validation_precedes_any_device_work(field, n, message):
    with pytest.raises(ValueError, match=message):
        field.ntt(np.zeros((n, 4), dtype=np.uint64))


def test_roots():
    assert quadratic_non_residue(R) == 5
    assert get_ntt_root(2**28, R) == 19103219067921713944291392827692070036145651957329286315305642004821462161904
    with pytest.raises(ValueError, match="power of 2"):
        get_ntt_root(12, R)


def test_window_model():
    assert [pippenger_window_size(n) for n in (0, 100, 1_000, 10_000, 100_000)] == [1, 5, 7, 10, 13]
    for bad in (0, 17):
        with pytest.raises(ValueError, match="window_size"):
            _check_window_size(bad)
    with pytest.raises(TypeError):
        _check_window_size(2.0)


def test_msm_argument_validation():
    with pytest.raises(ValueError, match="Length mismatch"):
        msm_gpu.msm_pippenger_gpu([G.x], [G.y], [1, 2])
    for bad in (0, 17):
        with pytest.raises(ValueError, match="window_size"):
            msm_gpu.msm_pippenger_gpu([G.x], [G.y], [1], window_size=bad)
    with pytest.raises(TypeError, match="scalars must contain integers"):
        msm_gpu.msm_pippenger_gpu([G.x], [G.y], [1.5])
    with pytest.raises(TypeError):
        msm_gpu.msm_pippenger_gpu([G.x], [G.y], [1], P)  # window_size is keyword-only
    assert msm_gpu.msm_pippenger_gpu([], [], []) == Point.identity()


def test_point_type():
    assert Point(1, 2 + P) == G
    assert -G == Point(1, P - 2)
    with pytest.raises(ValueError, match="not on the BN254 curve"):
        Point(1, 3)
    with pytest.raises(ValueError, match="identity"):
        Point(1, None)
    with pytest.raises(TypeError):
        Point(1.0, 2.0)
    assert from_jacobian(JACOBIAN_IDENTITY).is_identity()
    assert from_jacobian(to_jacobian(G)) == G
    lam = 12345
    assert from_jacobian((1 * lam**2 % P, 2 * lam**3 % P, lam)) == G
    with pytest.raises(ValueError, match="no affine coordinates"):
        Point.identity().affine()


def test_affine_rows_reject_the_identity():
    from zkgpu import curve_gpu

    with pytest.raises(ValueError, match="point at infinity"):
        curve_gpu.affine_to_device([G, Point.identity()])
    with pytest.raises(ValueError, match="triples"):
        curve_gpu.jacobian_to_device([(1, 2)])


@no_gpu
def test_no_cpu_fallback():
    for call in (
        lambda: BN254_FR.from_ints([1, 2]),
        lambda: BN254_FR.mul(np.zeros((1, 4), dtype=np.uint64), np.zeros((1, 4), dtype=np.uint64)),
        lambda: BN254_FR.ntt(np.zeros((4, 4), dtype=np.uint64)),
        lambda: msm_gpu.msm_pippenger_gpu([G.x], [G.y], [3]),
        lambda: G * 3,
    ):
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            call()


def test_bulk_int_rows_match_python_conversion():
    """The C ingest helper (zkgpu._ingest) gives the rows int.to_bytes gives, reduces only values
    outside [0, 2^256) on the host, accepts what operator.index accepts, and raises the
    reference's TypeError (msm_gpu.py:385-390)."""
    import random

    import numpy as np

    from zkgpu import field256_gpu as fg

    assert fg._ingest is not None, "zkgpu._ingest is not built"
    r = fg.BN254_SCALAR_MODULUS
    rng = random.Random(3)
    vals = [rng.getrandbits(256) for _ in range(500)] + [0, 1, 2**256 - 1, 2**256, -1, -(2**300), r, True]
    vals.append(np.int64(-7))
    rows, items = fg._rows_from_ints(iter(vals), r, "scalars")
    want = [v if 0 <= v < 2**256 else int(v) % r for v in (int(x) for x in vals)]
    assert fg._limbs_to_ints(rows) == want
    assert list(items) == vals
    assert rows.shape == (len(vals), 4) and rows.dtype == np.dtype("<u8")
    back = b"".join(v.to_bytes(32, "little") for v in want)
    assert rows.tobytes() == back
    with pytest.raises(TypeError, match=r"^scalars must contain integers: 'float' object cannot be interpreted"):
        fg._rows_from_ints([1, 2.5], r, "scalars")
    with pytest.raises(TypeError, match=r"^'str' object cannot be interpreted as an integer"):
        fg._rows_from_ints([1, "2"], r)
    assert fg._rows_from_ints([], r)[0].shape == (0, 4)


def test_bulk_conversion_on_worker_threads_matches_the_serial_one():
    """Long inputs are converted on several threads: the rows equal the serial conversion's for
    every kind of element wherever it sits, and the first bad element in index order raises."""
    import random

    import numpy as np

    from zkgpu import field256_gpu as fg

    r = fg.BN254_SCALAR_MODULUS
    rng = random.Random(11)
    n = (1 << 18) + 5
    vals = [rng.getrandbits(rng.choice((1, 64, 254, 256))) for _ in range(n)]
    odd = [-1, -(2**300), 2**256, 2**400 + 3, True, False, np.int64(-9), np.uint64(2**63 + 1)]
    for pos in rng.sample(range(n), 200):
        vals[pos] = rng.choice(odd)
    rows, _ = fg._rows_from_ints(vals, r, "scalars")
    want = [v if 0 <= v < 2**256 else v % r for v in (int(x) for x in vals)]
    assert fg._limbs_to_ints(rows) == want
    bad = list(vals)
    bad[n - 3], bad[n // 2] = "late", 1.5
    with pytest.raises(TypeError, match="'float' object"):
        fg._rows_from_ints(bad, r, "scalars")


def test_rows_into_writes_strided_rows_and_parts():
    """zkgpu._ingest.rows_into (the public MSM's staging conversion) writes row i at 32-byte row
    offset + i * stride of the destination, converts [start, stop) when given, leaves every other
    row alone, and refuses a destination that is too small."""
    import random

    from zkgpu import field256_gpu as fg

    r = fg.BN254_SCALAR_MODULUS
    rng = random.Random(5)
    vals = [rng.getrandbits(rng.choice((1, 30, 31, 60, 240, 241, 255, 256))) for _ in range(3000)]
    vals[7], vals[9], vals[11] = -3, 2**256 + 1, 2**270
    want = [(v if 0 <= v < 2**256 else v % r).to_bytes(32, "little") for v in vals]
    dest = bytearray(32 * (2 * len(vals) + 3))
    count, items = fg._ingest.rows_into(vals, r, "points_y", dest, 2, 3)
    assert count == len(vals) and items is vals
    assert bytes(dest[:96]) == bytes(96)
    for i, w in enumerate(want):
        assert dest[32 * (3 + 2 * i) : 32 * (4 + 2 * i)] == w
        assert dest[32 * (4 + 2 * i) : 32 * (5 + 2 * i)] == bytes(32)
    parts = bytearray(32 * len(vals))
    for lo in range(0, len(vals), 700):
        fg._ingest.rows_into(vals, r, "scalars", parts, 1, 0, lo, lo + 700)
    assert bytes(parts) == b"".join(want)
    with pytest.raises(ValueError, match="too small"):
        fg._ingest.rows_into(vals, r, "scalars", bytearray(32 * len(vals)), 2, 0)
    with pytest.raises(TypeError, match=r"^points_x must contain integers: 'float' object"):
        fg._ingest.rows_into([1, 2.5], r, "points_x", bytearray(64), 1, 0)


def test_ingest_worker_pool_across_jobs_and_a_fork():
    """The conversion's worker threads are kept between calls: jobs of every size in a row (one
    thread, a few, all sixteen), values the workers leave to the calling thread mixed in, and a
    conversion in a forked child, where the parent's workers do not exist, all give every row."""
    import os
    import random

    from zkgpu import field256_gpu as fg

    r = fg.BN254_SCALAR_MODULUS
    rng = random.Random(11)

    def check(n):
        vals = [rng.getrandbits(rng.choice((5, 200, 256))) for _ in range(n)]
        for i in range(0, n, 997):
            vals[i] = -vals[i] - 1  # reduced by the calling thread
        buf, _ = fg._ingest.rows_from_ints(vals, r, "v")
        return bytes(buf) == b"".join((v if v >= 0 else v % r).to_bytes(32, "little") for v in vals)

    for n in (1, 8191, 8192, 40000, 3, 140000, 9000, 0, 140000):
        assert check(n), n
    pid = os.fork()
    if pid == 0:
        os._exit(0 if check(140000) and check(20000) else 1)
    _, status = os.waitpid(pid, 0)
    assert os.WIFEXITED(status) and os.WEXITSTATUS(status) == 0
    assert check(140000)
