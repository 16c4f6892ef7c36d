"""
The 64-bit path (zkgpu.field_gpu, zkgpu.ntt_gpu): host logic on the CPU, kernels on the GPU.

The expected values are Python's exact integers (*, +, pow, %) for the elementwise functions
and the oracle's NTT (oracle/zkref/ntt.py, the reference's CPU transform for any prime modulus)
for ntt_gpu; at sizes the oracle cannot finish quickly the transform is checked through its
inverse and through direct evaluations X_k = sum_j a_j w^(jk) at sampled k. The cases follow
the reference's own tests (tests/test_field_gpu.py:56-175, tests/test_ntt.py:223-290).
"""

import random

import numpy as np
import pytest

from zkgpu import field_gpu, gpu_backend as gb
from zkgpu.field_gpu import _launch_shape
from zkgpu.ntt import get_ntt_root
from zkgpu.ntt_gpu import ntt_gpu
from zkref.ntt import ntt as ntt_oracle

NTT_MOD = 998_244_353
INT64_SAFE_LIMIT = 3_037_000_500
LARGEST_INT64_PRIME = 2**63 - 25
LARGE_INT64_NTT_PRIME = 2_147_483_641 * 2**32 + 1
INT64_MIN, INT64_MAX = -(2**63), 2**63 - 1
MODULI = [7, NTT_MOD, INT64_SAFE_LIMIT, INT64_SAFE_LIMIT + 1, 2**61 - 1, LARGEST_INT64_PRIME]
NTT_MODULI = [NTT_MOD, 12_289, LARGE_INT64_NTT_PRIME, 7_340_033]


def elements(n, mod, seed):
    rng = random.Random(seed)
    return [rng.randrange(mod) for _ in range(n)]


def dev(values, dtype=np.int64):
    return gb.asarray(np.array(values, dtype=dtype))


# ---- host logic (no GPU) ----------------------------------------------------------------------


class _Shaped:
    def __init__(self, shape):
        self.shape = tuple(shape)


@pytest.mark.parametrize(
    "sa,sb",
    [((5,), (5,)), ((), (7,)), ((3, 4), (4,)), ((3, 1), (1, 4)), ((2, 3, 4), (3, 1)), ((2, 1, 4, 1, 3), (1, 5, 1, 2, 1)),
     ((1, 1, 6), (6,)), ((4, 1), ()), ((2, 3, 4, 5, 6), (2, 3, 4, 5, 6)), ((0, 3), (3,)),
     ((2, 1) * 4, (1, 3) * 4)],
)
def test_broadcast_strides_address_what_numpy_broadcasts(sa, sb):
    """Every output index i reads operand elements at sum_d c_d(i) stride_d: the element
    np.broadcast_to puts there."""
    out = np.broadcast_shapes(sa, sb)
    shape, (ta, tb) = _launch_shape(out, [_Shaped(sa), _Shaped(sb)])
    assert 1 <= len(shape) <= 8
    n = int(np.prod(out))
    assert int(np.prod(shape)) == max(n, 1) or n == 0
    ia = np.broadcast_to(np.arange(int(np.prod(sa))).reshape(sa), out).ravel()
    ib = np.broadcast_to(np.arange(int(np.prod(sb))).reshape(sb), out).ravel()
    for i in range(n):
        rem, oa, ob = i, 0, 0
        for d in range(len(shape) - 1, -1, -1):
            c = rem % shape[d] if d else rem
            rem //= shape[d]
            oa += c * ta[d]
            ob += c * tb[d]
        assert (oa, ob) == (ia[i], ib[i]), i


def test_nine_unmergeable_dimensions_are_refused():
    sa, sb = (2, 1) * 4 + (2,), (1, 3) * 4 + (1,)
    with pytest.raises(ValueError, match="at most 8"):
        _launch_shape(np.broadcast_shapes(sa, sb), [_Shaped(sa), _Shaped(sb)])


@pytest.mark.parametrize("mod", [0, 1, -7, 2**63])
def test_modulus_outside_int64_is_refused_before_any_gpu_work(mod):
    x = np.array([1, 2])
    for call in (lambda: field_gpu.gpu_batch_mul(x, x, mod), lambda: field_gpu.gpu_batch_add(x, x, mod),
                 lambda: field_gpu.gpu_batch_pow(x, 2, mod)):
        with pytest.raises(ValueError, match="2 <= mod < 2"):
            call()


def test_non_integer_arguments_are_refused_before_any_gpu_work():
    ints, floats = np.array([1, 2]), np.array([1.5, 2.0])
    with pytest.raises(TypeError):
        field_gpu.gpu_batch_mul(ints, ints, 7.0)
    with pytest.raises(TypeError, match="integer array"):
        field_gpu.gpu_batch_mul(floats, ints, 7)
    with pytest.raises(TypeError, match="integer array"):
        field_gpu.gpu_batch_add(ints, floats, 7)
    with pytest.raises(TypeError, match="integer array"):
        field_gpu.gpu_batch_pow(floats, 2, 7)
    with pytest.raises(TypeError):
        field_gpu.gpu_batch_pow(ints, 2.0, 7)


def test_ntt_modulus_beyond_int64_is_refused_before_any_gpu_work():
    with pytest.raises(ValueError, match="below 2\\^63"):
        ntt_gpu(np.arange(4, dtype=np.int64), 21888242871839275222246405745257275088548364400416034343698204186575808495617)


@pytest.mark.skipif(gb.GPU_AVAILABLE, reason="checks the no-GPU error")
def test_without_a_gpu_the_calls_fail_loudly():
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        field_gpu.gpu_batch_mul(np.array([1]), np.array([2]), 7)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        ntt_gpu(np.array([1, 2]), NTT_MOD)


# ---- elementwise kernels ----------------------------------------------------------------------


@pytest.mark.gpu
@pytest.mark.parametrize("mod", MODULI)
def test_batch_mul_and_add_match_python(mod):
    a, b = elements(1000, mod, 1), elements(1000, mod, 2)
    got = field_gpu.gpu_batch_mul(dev(a), dev(b), mod)
    assert got.dtype == np.int64 and got.shape == (1000,)
    assert got.tolist() == [x * y % mod for x, y in zip(a, b)]
    assert field_gpu.gpu_batch_add(dev(a), dev(b), mod).tolist() == [(x + y) % mod for x, y in zip(a, b)]


@pytest.mark.gpu
@pytest.mark.parametrize("mod", [INT64_SAFE_LIMIT, INT64_SAFE_LIMIT + 1, LARGEST_INT64_PRIME, 2**62, 2**62 + 1, 2, 3])
def test_extreme_residues(mod):
    a = [mod - 1, mod - 1, mod - 2, 0, 1]
    b = [mod - 1, 1, mod - 1, mod - 1, mod - 1]
    assert field_gpu.gpu_batch_mul(dev(a), dev(b), mod).tolist() == [x * y % mod for x, y in zip(a, b)]
    assert field_gpu.gpu_batch_add(dev(a), dev(b), mod).tolist() == [(x + y) % mod for x, y in zip(a, b)]


@pytest.mark.gpu
@pytest.mark.parametrize("exp", [0, 1, 2, 65537, NTT_MOD - 2])
def test_batch_pow_matches_python(exp):
    bases = elements(256, NTT_MOD, 5)
    assert field_gpu.gpu_batch_pow(dev(bases), exp, NTT_MOD).tolist() == [pow(x, exp, NTT_MOD) for x in bases]


@pytest.mark.gpu
@pytest.mark.parametrize("exp", [3, 2**61 - 3, 3**3000, 7**5000 + 1], ids=["3", "fermat", "3^3000", "7^5000+1"])
def test_batch_pow_large_modulus_and_long_exponents(exp):
    """2^61 - 3 is the Fermat inverse; 3^3000 and 7^5000 + 1 take several 4096-bit chunks."""
    mod = 2**61 - 1
    bases = elements(64, mod, 6) + [0, 1, mod - 1]
    assert field_gpu.gpu_batch_pow(dev(bases), exp, mod).tolist() == [pow(x, exp, mod) for x in bases]


@pytest.mark.gpu
@pytest.mark.parametrize("mod", [7, LARGEST_INT64_PRIME, 10, 2**40])
def test_unreduced_and_negative_inputs_reduce_into_the_field(mod):
    a = [-3, -10, INT64_MAX, INT64_MIN, 12, 0, -1, mod, -mod]
    b = [5, 1, INT64_MAX, INT64_MIN, -12, -1, INT64_MIN, min(3 * mod, INT64_MAX - 2), -1]
    assert field_gpu.gpu_batch_mul(dev(a), dev(b), mod).tolist() == [x * y % mod for x, y in zip(a, b)]
    assert field_gpu.gpu_batch_add(dev(a), dev(b), mod).tolist() == [(x + y) % mod for x, y in zip(a, b)]
    assert field_gpu.gpu_batch_pow(dev(a), 3, mod).tolist() == [pow(x, 3, mod) for x in a]


@pytest.mark.gpu
def test_integer_dtypes_are_widened_like_astype_int64():
    """(NTT_MOD - 1)^2 overflows int32 but not the widened product; uint64 keeps its bit pattern."""
    values = dev([NTT_MOD - 1, 3], np.int32)
    assert field_gpu.gpu_batch_mul(values, values, NTT_MOD).tolist() == [pow(NTT_MOD - 1, 2, NTT_MOD), 9]
    assert field_gpu.gpu_batch_pow(values, 2, 2**61 - 1).tolist() == [(NTT_MOD - 1) ** 2, 9]
    mod = 1_000_003
    for dtype in (np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64):
        info = np.iinfo(dtype)
        host = np.array([info.min, info.max, 0, 1, info.max // 3], dtype=dtype)
        wide = [int(v) for v in host.astype(np.int64)]
        got = field_gpu.gpu_batch_add(gb.asarray(host), dev([0] * 5), mod).tolist()
        assert got == [v % mod for v in wide], dtype


@pytest.mark.gpu
def test_host_operands_are_uploaded():
    assert field_gpu.gpu_batch_mul([3, 4], np.array([5, 6]), 7).tolist() == [1, 3]


@pytest.mark.gpu
def test_negative_exponent_is_the_inverse_and_rejects_zero():
    bases = [3, 5, 6]
    for exp in (-1, -3):
        assert field_gpu.gpu_batch_pow(dev(bases), exp, 7).tolist() == [pow(x, exp, 7) for x in bases]
    big = elements(300, LARGEST_INT64_PRIME, 9)
    assert field_gpu.gpu_batch_pow(dev(big), -5, LARGEST_INT64_PRIME).tolist() == [pow(x, -5, LARGEST_INT64_PRIME) for x in big]
    with pytest.raises(ValueError, match="not invertible"):
        field_gpu.gpu_batch_pow(dev([3, 7]), -1, 7)
    with pytest.raises(ValueError, match="not invertible"):
        field_gpu.gpu_batch_pow(dev([3, 14, 5]).reshape(3, 1), -2, 7)


@pytest.mark.gpu
def test_broadcasting():
    mod = LARGEST_INT64_PRIME
    a = np.array(elements(12, mod, 11), dtype=np.int64).reshape(3, 4)
    row = np.array(elements(4, mod, 12), dtype=np.int64)
    col = np.array(elements(3, mod, 13), dtype=np.int64).reshape(3, 1)
    scalar = np.array(elements(1, mod, 14)[0], dtype=np.int64)

    def want(x, y, f):
        x, y = np.broadcast_arrays(x.astype(object), y.astype(object))
        return np.vectorize(f, otypes=[object])(x, y).tolist()

    for x, y in ((a, row), (row, a), (a, col), (col, row), (scalar, a), (a, scalar), (scalar, scalar)):
        got = field_gpu.gpu_batch_mul(gb.asarray(x), gb.asarray(y), mod)
        assert got.shape == np.broadcast_shapes(x.shape, y.shape)
        assert got.tolist() == want(x, y, lambda u, v: int(u) * int(v) % mod)
        assert field_gpu.gpu_batch_add(gb.asarray(x), gb.asarray(y), mod).tolist() == want(x, y, lambda u, v: (int(u) + int(v)) % mod)
    assert field_gpu.gpu_batch_mul(gb.asarray(np.zeros((0, 4), np.int64)), gb.asarray(row), mod).shape == (0, 4)
    with pytest.raises(ValueError):
        field_gpu.gpu_batch_mul(gb.asarray(a), gb.asarray(np.zeros(3, np.int64)), mod)


@pytest.mark.gpu
def test_ten_million_elements_match_numpy():
    """The reference's largest published field size (10^7), against NumPy's exact int64 product."""
    n, mod = 10_000_000, NTT_MOD
    rng = np.random.default_rng(3)
    a = rng.integers(INT64_MIN, INT64_MAX, n, dtype=np.int64, endpoint=True)
    b = rng.integers(0, mod, n, dtype=np.int64)
    ar, br = np.mod(a, mod), b
    got = field_gpu.gpu_batch_mul(gb.asarray(a), gb.asarray(b), mod).get()
    assert np.array_equal(got, ar * br % mod)
    assert np.array_equal(field_gpu.gpu_batch_add(gb.asarray(a), gb.asarray(b), mod).get(), (ar + br) % mod)


# ---- NTT ---------------------------------------------------------------------------------------


@pytest.mark.gpu
@pytest.mark.parametrize("mod", [NTT_MOD, 12_289, LARGE_INT64_NTT_PRIME])
@pytest.mark.parametrize("n", [1, 2, 8, 256, 1 << 11, 1 << 12])
def test_ntt_matches_the_oracle(n, mod):
    a = elements(n, mod, n)
    assert ntt_gpu(dev(a), mod).tolist() == ntt_oracle(a, mod)


@pytest.mark.gpu
@pytest.mark.parametrize("mod", NTT_MODULI)
def test_ntt_two_passes_match_the_oracle(mod):
    for k in (13, 16):
        if (mod - 1) % (1 << k):
            continue
        a = elements(1 << k, mod, k)
        assert ntt_gpu(dev(a), mod).tolist() == ntt_oracle(a, mod), k
        assert ntt_gpu(dev(a), mod, invert=True).tolist() == ntt_oracle(a, mod, invert=True), k


@pytest.mark.gpu
@pytest.mark.parametrize("mod", [NTT_MOD, 12_289, LARGE_INT64_NTT_PRIME])
def test_ntt_inverse_round_trip(mod):
    a = elements(1024, mod, 55)
    assert ntt_gpu(ntt_gpu(dev(a), mod), mod, invert=True).tolist() == a


@pytest.mark.gpu
def test_ntt_repeated_calls_agree_and_leave_the_input_alone():
    a = gb.asarray(np.arange(4096, dtype=np.int64))
    first = ntt_gpu(a, NTT_MOD).tolist()
    assert ntt_gpu(a, NTT_MOD).tolist() == first
    assert a.tolist() == list(range(4096))


@pytest.mark.gpu
@pytest.mark.parametrize("mod", [NTT_MOD, 12_289, LARGE_INT64_NTT_PRIME])
def test_ntt_reduces_unreduced_input(mod):
    a = [mod + 5, 2 * mod + 1 if 2 * mod + 1 <= INT64_MAX else -1, INT64_MAX, INT64_MIN] + [0] * 4
    assert ntt_gpu(dev(a), mod).tolist() == ntt_oracle(a, mod)
    assert ntt_gpu(a, mod).tolist() == ntt_oracle(a, mod)  # host input


@pytest.mark.gpu
def test_ntt_single_element_is_reduced_and_empty_stays_empty():
    assert ntt_gpu(dev([NTT_MOD + 5]), NTT_MOD).tolist() == [5]
    assert ntt_gpu(dev([], np.int64), NTT_MOD).shape == (0,)


@pytest.mark.gpu
def test_ntt_validates_like_the_cpu_ntt():
    with pytest.raises(ValueError, match="odd prime"):
        ntt_gpu(dev(list(range(4))), 21)
    with pytest.raises(ValueError, match="power of 2"):
        ntt_gpu(dev(list(range(3))), NTT_MOD)
    with pytest.raises(ValueError, match="not supported"):
        ntt_gpu(dev(list(range(1 << 13))), 12_289)
    with pytest.raises(TypeError, match="integer array"):
        ntt_gpu(gb.asarray(np.zeros(4)), NTT_MOD)


def _evaluations(a: np.ndarray, mod: int, omega: int, ks: list[int]) -> list[int]:
    """X_k = sum_j a_j omega^(jk) mod mod for a 30-bit modulus, vectorised in uint64."""
    n, m = len(a), np.uint64(mod)
    half = 1 << ((n.bit_length() - 1) // 2)  # j = half * u + v
    u = np.arange(n, dtype=np.uint64) // np.uint64(half)
    v = np.arange(n, dtype=np.uint64) % np.uint64(half)
    out = []
    for k in ks:
        w = pow(omega, k, mod)
        wv = np.array([pow(w, i, mod) for i in range(half)], dtype=np.uint64)
        wu = np.array([pow(w, half * i, mod) for i in range(n // half)], dtype=np.uint64)
        pw = wu[u] * wv[v] % m
        out.append(int((a.astype(np.uint64) * pw % m).sum() % m))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("log_n", [20, 22, 23])
def test_ntt_at_size_through_evaluations_and_its_inverse(log_n):
    """2^20 is two passes, 2^22 and 2^23 three: sampled outputs equal direct evaluations of the
    polynomial and the inverse transform returns the input exactly."""
    n, mod = 1 << log_n, NTT_MOD
    rng = np.random.default_rng(log_n)
    a = rng.integers(0, mod, n, dtype=np.int64)
    x = gb.asarray(a)
    got = ntt_gpu(x, mod)
    ks = [0, 1, n - 1, n // 2, *[int(v) for v in rng.integers(0, n, 4)]]
    host = got.get()
    assert [int(host[k]) for k in ks] == _evaluations(a, mod, get_ntt_root(n, mod), ks)
    assert np.array_equal(ntt_gpu(got, mod, invert=True).get(), a)


@pytest.mark.gpu
def test_ntt_three_passes_with_a_63_bit_prime():
    """2^23 points modulo 2147483641 * 2^32 + 1: the round trip, linearity and X_0 = sum a_j."""
    n, mod = 1 << 23, LARGE_INT64_NTT_PRIME
    rng = np.random.default_rng(23)
    a = rng.integers(INT64_MIN, INT64_MAX, n, dtype=np.int64, endpoint=True)
    b = rng.integers(0, mod, n, dtype=np.int64)
    fa, fb = ntt_gpu(gb.asarray(a), mod), ntt_gpu(gb.asarray(b), mod)
    assert np.array_equal(ntt_gpu(fa, mod, invert=True).get(), np.mod(a, mod) if mod < 2**63 else None)
    f_sum = ntt_gpu(field_gpu.gpu_batch_add(gb.asarray(a), gb.asarray(b), mod), mod)
    assert np.array_equal(f_sum.get(), field_gpu.gpu_batch_add(fa, fb, mod).get())
    assert int(fa.get()[0]) == sum(int(v) for v in np.mod(a, mod).tolist()) % mod
