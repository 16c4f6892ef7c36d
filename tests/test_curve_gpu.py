"""
zkgpu.curve_gpu on the GPU: exact Jacobian coordinates against the oracle.

Mirrors the reference's tests/test_curve_gpu.py: every special case of
add-2007-bl, madd-2007-bl and dbl-2009-l, inputs with Z != 1, chained
operations, conversions, broadcasting and errors.
"""

import random

import pytest

from zkgpu import curve_gpu
from zkgpu.curve import G as gG, Point as gPoint
from zkgpu.field256_gpu import BN254_FQ
from zkref.curve import JACOBIAN_IDENTITY, G, P, from_jacobian, jacobian_add, jacobian_add_affine, jacobian_double, jacobian_mul
from zkref.field import R_SCALAR as R

pytestmark = pytest.mark.gpu


def jacobians(n, seed):
    rng = random.Random(seed)
    return [jacobian_mul(G * rng.randrange(1, R), rng.randrange(2, 2**64)) for _ in range(n)]


def negate(j):
    return (j[0], -j[1] % P, j[2])


def test_add_matches_oracle():
    js = jacobians(12, 1)
    a = js + js[:4] + js[4:8] + [JACOBIAN_IDENTITY] * 2 + js[8:10]
    b = js[::-1] + js[:4] + [negate(j) for j in js[4:8]] + js[10:12] + [JACOBIAN_IDENTITY] * 2
    got = curve_gpu.add(curve_gpu.jacobian_to_device(a), curve_gpu.jacobian_to_device(b))
    assert curve_gpu.jacobian_ints(got) == [jacobian_add(x, y) for x, y in zip(a, b)]


def test_add_affine_matches_oracle():
    js = jacobians(10, 2)
    pts = [from_jacobian(j) for j in jacobians(10, 3)]
    same = [from_jacobian(j) for j in js[:4]]
    acc = js + js[:4] + [negate(j) for j in js[4:8]] + [JACOBIAN_IDENTITY] * 2
    q = pts + same + [from_jacobian(j) for j in js[4:8]] + pts[:2]
    got = curve_gpu.add_affine(curve_gpu.jacobian_to_device(acc), curve_gpu._affine_rows([p.affine() for p in q]))
    assert curve_gpu.jacobian_ints(got) == [jacobian_add_affine(j, *p.affine()) for j, p in zip(acc, q)]


def test_double_matches_oracle():
    js = [*jacobians(10, 4), JACOBIAN_IDENTITY, (1, 2, 1), (5, 0, 1)]
    got = curve_gpu.double(curve_gpu.jacobian_to_device(js))
    assert curve_gpu.jacobian_ints(got) == [jacobian_double(j) for j in js]


def test_many_random_points():
    js = jacobians(300, 7)
    rng = random.Random(8)
    other = [js[rng.randrange(300)] for _ in range(300)]
    got = curve_gpu.add(curve_gpu.jacobian_to_device(js), curve_gpu.jacobian_to_device(other))
    assert curve_gpu.jacobian_ints(got) == [jacobian_add(x, y) for x, y in zip(js, other)]


def test_chained_doublings_and_additions():
    acc = curve_gpu.to_device([gG])
    for _ in range(10):
        acc = curve_gpu.double(acc)
    acc = curve_gpu.add_affine(acc, curve_gpu.affine_to_device([gG * 3]))
    assert curve_gpu.from_device(acc)[0].affine() == (G * 1027).affine()


def test_points_round_trip_and_revalidation():
    pts = [gG, gG * 5, -gG, gPoint.identity()]
    rows = curve_gpu.to_device(pts)
    assert rows.shape == (4, 3, 4)
    assert curve_gpu.from_device(rows) == pts
    assert curve_gpu.jacobian_ints(rows)[3] == JACOBIAN_IDENTITY
    with pytest.raises(ValueError, match="not on the BN254 curve"):
        curve_gpu.from_device(curve_gpu.jacobian_to_device([(1, 3, 1)]))


def test_point_operators_run_on_the_gpu():
    assert (gG + gG).affine() == (G * 2).affine()
    assert gG.double().affine() == (G * 2).affine()
    assert (gG * 12345).affine() == (G * 12345).affine()
    assert (gG * R).is_identity()


def test_broadcasting_empty_and_errors():
    js = jacobians(3, 5)
    rows, one = curve_gpu.jacobian_to_device(js), curve_gpu.to_device([gG])
    g = (1, 2, 1)
    assert curve_gpu.jacobian_ints(curve_gpu.add(rows, one)) == [jacobian_add(j, g) for j in js]
    assert curve_gpu.jacobian_ints(curve_gpu.add(one, rows)) == [jacobian_add(g, j) for j in js]
    got = curve_gpu.add_affine(rows, curve_gpu.affine_to_device([gG]))
    assert curve_gpu.jacobian_ints(got) == [jacobian_add_affine(j, 1, 2) for j in js]
    empty = curve_gpu.to_device([])
    assert empty.shape == (0, 3, 4)
    assert curve_gpu.from_device(curve_gpu.add(empty, empty)) == []
    with pytest.raises(ValueError, match="broadcast"):
        curve_gpu.add(curve_gpu.to_device([gG, gG]), curve_gpu.to_device([gG, gG, gG]))
    affine = curve_gpu.affine_to_device([gG])
    with pytest.raises(TypeError, match=r"\(n, 3, 4\) uint64 array of Jacobian points"):
        curve_gpu.add(affine, affine)
    with pytest.raises(TypeError, match=r"\(n, 2, 4\) uint64 array of affine points"):
        curve_gpu.add_affine(curve_gpu.to_device([gG]), curve_gpu.to_device([gG]))
    with pytest.raises(TypeError, match="Jacobian points"):
        curve_gpu.double(BN254_FQ.from_ints([1, 2, 3]))


def test_self_test_failure_is_not_cached(monkeypatch):
    monkeypatch.setattr(curve_gpu, "_compiled", {})
    real = curve_gpu.jacobian_ints
    monkeypatch.setattr(curve_gpu, "jacobian_ints", lambda x: [(a, b, (c + 1) % P) for a, b, c in real(x)])
    with pytest.raises(RuntimeError, match="failed its self-test"):
        curve_gpu.compile_kernels()
    assert curve_gpu._compiled == {}
