"""
Writes the committed fixtures of this test suite from the oracle (oracle/zkref).

Outputs
-------
tests/golden/vectors.json : seeded inputs and the oracle's outputs for the
    parity tests (field ops, NTTs, point formulas, MSMs), small enough to
    commit and to check on a machine without the oracle's C library.
zk-gpu-accelerator_amd/zkgpu/_selftest_vectors.py : the exact Jacobian
    triples and the small MSM the package checks its kernels against when
    they are first used (the cases of the reference's load-time self-tests,
    src/zkgpu/curve_gpu.py:313-319 and msm_gpu.py:250-252).

Run from the repository root: python tests/golden/make_vectors.py
"""

from __future__ import annotations

import json
import random
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))

from zkref.curve import G, P, JACOBIAN_IDENTITY, from_jacobian, jacobian_add, jacobian_add_affine, jacobian_double, jacobian_mul  # noqa: E402
from zkref.field import R_SCALAR  # noqa: E402
from zkref.msm import msm_double_and_add, msm_pippenger  # noqa: E402
from zkref.ntt import ntt  # noqa: E402


def selftest_module() -> str:
    j5, j7, j11 = (jacobian_mul(G, k) for k in (5, 7, 11))
    neg5 = (j5[0], -j5[1] % P, j5[2])
    p7, p11 = from_jacobian(j7).affine(), from_jacobian(j11).affine()
    pairs = [(j5, j7), (j5, j5), (j5, neg5), (JACOBIAN_IDENTITY, j7), (j5, JACOBIAN_IDENTITY)]
    mixed = [(j5, p7), (j11, p11), (jacobian_add(j11, jacobian_double(j11)), p11), (JACOBIAN_IDENTITY, p7),
             ((j11[0], -j11[1] % P, j11[2]), p11)]
    singles = [j5, j11, JACOBIAN_IDENTITY]
    pts = [G, G * 2, G * 3, -G, G, G * 5, G * 6, G * 7]
    scalars = [5, 1, 2, 5, 5, R_SCALAR - 1, 12345, 2**200 + 7]
    expected = msm_double_and_add(pts, scalars)
    lines = [
        '"""',
        "Fixed self-test data for the kernels, checked when they are first used.",
        "",
        "Exact Jacobian triples (BN254 G1, plain coordinates) for every branch of",
        "add-2007-bl, madd-2007-bl and dbl-2009-l, and one small MSM that reaches",
        "several points per bucket, P and -P in a bucket, P twice in a bucket and",
        "digit-0 entries. Generated by tests/golden/make_vectors.py; do not edit.",
        '"""',
        "",
        f"ADD_A = {[list(a) for a, _ in pairs]!r}",
        f"ADD_B = {[list(b) for _, b in pairs]!r}",
        f"ADD_OUT = {[list(jacobian_add(a, b)) for a, b in pairs]!r}",
        f"MADD_A = {[list(a) for a, _ in mixed]!r}",
        f"MADD_B = {[list(q) for _, q in mixed]!r}",
        f"MADD_OUT = {[list(jacobian_add_affine(a, *q)) for a, q in mixed]!r}",
        f"DBL_A = {[list(a) for a in singles]!r}",
        f"DBL_OUT = {[list(jacobian_double(a)) for a in singles]!r}",
        f"MSM_POINTS = {[list(p.affine()) for p in pts]!r}",
        f"MSM_SCALARS = {scalars!r}",
        f"MSM_WINDOW = 6",
        f"MSM_OUT = {list(expected.affine())!r}",
        "",
    ]
    return "\n".join(lines)


def vectors() -> dict:
    rng = random.Random(0x5EED)
    out: dict = {"r": str(R_SCALAR), "p": str(P)}
    # NTT over r: n = 2^k for k in 0..6, forward and inverse
    ntts = []
    for k in range(0, 7):
        a = [rng.randrange(R_SCALAR) for _ in range(1 << k)]
        ntts.append({"input": [str(v) for v in a], "forward": [str(v) for v in ntt(a, R_SCALAR)],
                     "inverse": [str(v) for v in ntt(a, R_SCALAR, True)]})
    out["ntt"] = ntts
    # MSM: 24 random multiples of G with full-size scalars, window 5
    ks = [rng.randrange(1, R_SCALAR) for _ in range(24)]
    pts = [G * k for k in ks]
    sc = [rng.randrange(R_SCALAR) for _ in range(24)]
    res = msm_pippenger(pts, sc, window_size=5)
    out["msm"] = {"points": [[str(c) for c in p.affine()] for p in pts], "scalars": [str(s) for s in sc],
                  "result": [str(c) for c in res.affine()]}
    return out


if __name__ == "__main__":
    (ROOT / "zk-gpu-accelerator_amd" / "zkgpu" / "_selftest_vectors.py").write_text(selftest_module())
    (ROOT / "tests" / "golden" / "vectors.json").write_text(json.dumps(vectors(), indent=1))
    print("wrote _selftest_vectors.py and vectors.json")
