"""
The committed fixtures (tests/golden/vectors.json and the package's
_selftest_vectors.py) are reproduced by the oracle, so they cannot drift from it.
"""

import json
import random
from pathlib import Path

from zkgpu import _selftest_vectors as sv
from zkref.curve import G, P, Point, jacobian_add, jacobian_add_affine, jacobian_double
from zkref.field import R_SCALAR as R
from zkref.msm import msm_double_and_add, msm_pippenger
from zkref.ntt import ntt

GOLDEN = Path(__file__).resolve().parent / "golden" / "vectors.json"


def test_vectors_json():
    v = json.loads(GOLDEN.read_text())
    assert int(v["r"]) == R and int(v["p"]) == P
    for case in v["ntt"]:
        a = [int(x) for x in case["input"]]
        assert [str(x) for x in ntt(a, R)] == case["forward"]
        assert [str(x) for x in ntt(a, R, True)] == case["inverse"]
    m = v["msm"]
    pts = [Point(int(x), int(y)) for x, y in m["points"]]
    sc = [int(s) for s in m["scalars"]]
    assert [str(c) for c in msm_pippenger(pts, sc, window_size=5).affine()] == m["result"]


def test_selftest_vectors():
    assert [list(jacobian_add(tuple(a), tuple(b))) for a, b in zip(sv.ADD_A, sv.ADD_B)] == sv.ADD_OUT
    assert [list(jacobian_add_affine(tuple(a), *q)) for a, q in zip(sv.MADD_A, sv.MADD_B)] == sv.MADD_OUT
    assert [list(jacobian_double(tuple(a))) for a in sv.DBL_A] == sv.DBL_OUT
    pts = [Point(x, y) for x, y in sv.MSM_POINTS]
    assert list(msm_double_and_add(pts, sv.MSM_SCALARS).affine()) == sv.MSM_OUT
    # every branch is present: equal points, inverse points, the identity on either side
    assert any(a == b for a, b in zip(sv.ADD_A, sv.ADD_B))
    assert [1, 1, 0] in sv.ADD_A and [1, 1, 0] in sv.ADD_B and [1, 1, 0] in sv.MADD_A


def test_seeded_streams_are_stable():
    """The workload generators depend on random.Random staying the same sequence."""
    assert random.Random(0x5EED01).randrange(R) == random.Random(0x5EED01).randrange(R)
