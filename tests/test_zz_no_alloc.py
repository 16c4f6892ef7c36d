"""
zk_msm_g1_noalloc, the MSM entry that allocates no device memory, against zk_msm_g1.

The same arguments must give the same group element, compared bit for bit in affine coordinates
(its Jacobian coordinates vary from run to run of either entry, since the sort places a bucket's
entries in LDS-atomic order), on two calls in a row, the second with a smaller input; the result
is also pinned to the closed form (sum k_i s_i mod r) G over the basis P_i = k_i G. Sizes whose
workspace exceeds the fixed scratch array are refused with an error before anything runs.
"""

import ctypes
import sys

import numpy as np
import pytest

from zkgpu import gpu_backend as gb, workloads
from zkref import fast
from zkref.curve import G
from zkref.field import limbs_to_ints

pytestmark = pytest.mark.gpu

P = 21888242871839275222246405745257275088696311157297823662689037894645226208583


def jacobian(name, pts, sc, n, window=0, chunk=0):
    out = gb.zeros((3, 4))
    gb.call(name, pts.ptr, sc.ptr, n, window, chunk, out.ptr, gb.current_stream())
    return out.get()


def affine(jac_mont):
    x, y, z = (v * pow(2, -256, P) % P for v in limbs_to_ints(np.ascontiguousarray(jac_mont)))
    if z == 0:
        return None
    zi = pow(z, -1, P)
    return x * zi * zi % P, y * zi * zi * zi % P


def test_noalloc_equals_msm_g1_on_two_calls():
    n = 1 << 12
    keys = workloads.key_rows(n, 0x5EED21)
    pts = workloads.srs_points(keys)
    host = workloads.witness_rows(n, 0x5EED21)
    sc = gb.asarray(host)
    for m, window, chunk in ((n, 0, 0), (700, 0, 0), (n, 9, 4), (300, 5, 0)):
        print(f"noalloc: n={m} window={window} chunk={chunk}", file=sys.stderr, flush=True)
        want = jacobian("zk_msm_g1", pts, sc, m, window, chunk)
        got = jacobian("zk_msm_g1_noalloc", pts, sc, m, window, chunk)
        assert affine(got) == affine(want), (m, window, chunk)
        e = fast.dot_mod_r(keys[:m], host[:m])
        assert affine(got) == (None if e == 0 else (G * e).affine()), (m, window, chunk)


def test_noalloc_refuses_msms_beyond_its_scratch():
    cap = ctypes.c_uint64()
    gb.call("zk_msm_noalloc_max_bytes", ctypes.byref(cap))
    assert cap.value == 64 << 20
    n = 1 << 20
    pts, sc = gb.zeros((n, 2, 4)), gb.zeros((n, 4))
    out = gb.zeros((3, 4))
    with pytest.raises(gb.ZkError, match="bytes of scratch"):
        gb.call("zk_msm_g1_noalloc", pts.ptr, sc.ptr, n, 0, 0, out.ptr, gb.current_stream())
    # the largest power of two accepted with the automatic window and chunk (all-zero inputs: the
    # identity if accepted, nothing run if refused)
    largest = 0
    for k in range(10, 20):
        try:
            gb.call("zk_msm_g1_noalloc", pts.ptr, sc.ptr, 1 << k, 0, 0, out.ptr, gb.current_stream())
        except gb.ZkError:
            break
        largest = k
    gb.synchronize()
    print(f"noalloc: largest accepted 2^{largest}", file=sys.stderr, flush=True)
    assert largest >= 14
