"""
The C ABI and the kernels' constants, checked without a GPU.

- libzkgpu_hip.so loads and exports every entry point include/zkgpu.h declares;
- zk_constants.hpp holds exactly what the 29-bit-limb arithmetic needs for
  BN254's r and p (recomputed here from the moduli);
- the Montgomery product's column accumulator cannot overflow 64 bits and its
  output stays below 2p: a Python replay of the kernel's product-scanning loop
  at the extreme operands (the reference's five-word bound test,
  test_field256_gpu.py:97-118, restated for nine 29-bit limbs).
"""

import ctypes
import random
import re
from pathlib import Path

import pytest

from zkgpu import gpu_backend

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "zkgpu.h"
CONSTS = ROOT / "zk-gpu-accelerator_amd" / "csrc" / "zk_constants.hpp"
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P = 21888242871839275222246405745257275088696311157297823662689037894645226208583


def declared() -> list[str]:
    return re.findall(r"^(?:int|const char\*)\s+(zk_\w+)\(", HEADER.read_text(), re.M)


def test_library_exports_every_declared_entry_point():
    assert gpu_backend.LIB_PATH.exists(), "build libzkgpu_hip.so first (__graft_entry__.build())"
    lib = ctypes.CDLL(str(gpu_backend.LIB_PATH))
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_match_the_header():
    names = set(declared())
    bound = set(gpu_backend.SIGNATURES) | {"zk_last_error"}
    assert bound == names


def limbs29(v):
    return [(v >> (29 * i)) & ((1 << 29) - 1) for i in range(9)]


def parse_consts():
    text = CONSTS.read_text()
    out = {}
    for struct in ("FrConst", "FqConst"):
        body = text.split(f"struct {struct}")[1].split("\n};")[0]
        d = {}
        for name, vals in re.findall(r"(\w+)\[\d\] = \{([^}]*)\}", body):
            d[name] = [int(v.strip().rstrip("u"), 16) for v in vals.split(",")]
        d["ninv"] = int(re.search(r"ninv = (0x[0-9a-f]+)u", body).group(1), 16)
        out[struct] = d
    return out


@pytest.mark.parametrize("struct,m", [("FrConst", R), ("FqConst", P)])
def test_constants(struct, m):
    c = parse_consts()[struct]
    assert c["p"] == limbs29(m)
    assert c["ninv"] == (-pow(m, -1, 2**29)) % 2**29
    for name, e in (("c256", 256), ("c266", 266), ("c517", 517), ("c522", 522), ("one", 261)):
        assert c[name] == limbs29(pow(2, e, m)), name
    assert c["mone"] == limbs29(-pow(2, 261, m) % m)  # -1 in the 2^261 form
    assert c["p32"] == [(m >> (32 * i)) & 0xFFFFFFFF for i in range(8)]
    # redundant 4p and 8p: the value, and limbs 0..7 large enough that a + kp - b keeps every
    # limb non-negative for a normalized b (4p) or a carry-free sum of two normalized values (8p)
    for name, k, floor in (("p4r", 4, 2**29 - 1), ("p8r", 8, 2**30 - 2)):
        r = c[name]
        assert sum(v << (29 * i) for i, v in enumerate(r)) == k * m, name
        assert all(v >= floor for v in r[:8]) and r[8] >= 0, name
        assert all(v < 2**32 for v in r), name


def mont29(a, b, m):
    """Replay of zk_field.hpp's mont_raw: returns (result, largest accumulator seen)."""
    p = limbs29(m)
    ninv = (-pow(m, -1, 2**29)) % 2**29
    A, B = limbs29(a), limbs29(b)
    mm, acc, top, out = [0] * 9, 0, 0, [0] * 9
    for k in range(9):
        for i in range(k + 1):
            acc += A[i] * B[k - i]
        for i in range(k):
            acc += mm[i] * p[k - i]
        mm[k] = (acc & 0xFFFFFFFF) * ninv & ((1 << 29) - 1)
        acc += mm[k] * p[0]
        top = max(top, acc)
        assert acc % (1 << 29) == 0
        acc >>= 29
    for k in range(9, 17):
        for i in range(k - 8, 9):
            acc += A[i] * B[k - i] + mm[i] * p[k - i]
        top = max(top, acc)
        out[k - 9] = acc & ((1 << 29) - 1)
        acc >>= 29
    out[8] = acc
    return sum(v << (29 * i) for i, v in enumerate(out)), top


def mont29_limbs(A, B, m):
    """mont29 on raw limb vectors (limbs beyond 29 bits, as the carry-free operands have)."""
    p = limbs29(m)
    ninv = (-pow(m, -1, 2**29)) % 2**29
    mm, acc, top, out = [0] * 9, 0, 0, [0] * 9
    for k in range(9):
        acc += sum(A[i] * B[k - i] for i in range(k + 1)) + sum(mm[i] * p[k - i] for i in range(k))
        mm[k] = (acc & 0xFFFFFFFF) * ninv & ((1 << 29) - 1)
        acc += mm[k] * p[0]
        top = max(top, acc)
        acc >>= 29
    for k in range(9, 17):
        acc += sum(A[i] * B[k - i] + mm[i] * p[k - i] for i in range(k - 8, 9))
        top = max(top, acc)
        out[k - 9] = acc & ((1 << 29) - 1)
        acc >>= 29
    out[8] = acc
    return sum(v << (29 * i) for i, v in enumerate(out)), top


def lc_finish29(L, m):
    """zk_field.hpp's lc_finish on signed limbs: the quotient from the top limb, then carries."""
    p = limbs29(m)
    t = L[8] - 4096
    q = t // (p[8] + 1) if t > 0 else 0
    a = [x - q * y for x, y in zip(L, p)]
    for i in range(8):
        a[i + 1] += a[i] >> 29
        a[i] &= (1 << 29) - 1
    return a


@pytest.mark.parametrize("struct,m", [("FrConst", R), ("FqConst", P)])
def test_wide_butterfly_sums(struct, m):
    """The NTT's carry-free sums: x0 + x2 and x1 + x3 (each below 4p), their sum reduced by the
    top-limb quotient, their difference + 4p likewise, and their difference + 8p (p8r) as a
    product operand, at the largest limbs and values the stage pair can hand over."""
    c = parse_consts()[struct]
    rng = random.Random(m % 997)
    hi = 2 * m - 1  # the largest value below 2p, and the largest normalized limbs
    vals = [0, 1, m - 1, m, hi, (1 << 253) - 1] + [rng.randrange(2 * m) for _ in range(200)]
    maxlimbs = [2**29 - 1] * 8 + [limbs29(hi)[8]]
    tw = [limbs29(m - 1), limbs29(hi), [2**29 - 1] * 8 + [limbs29(hi)[8]]]
    for _ in range(300):
        xs = [limbs29(rng.choice(vals)) for _ in range(4)]
        if rng.random() < 0.2:
            xs[rng.randrange(4)] = maxlimbs  # limbs at their bound (a value up to ~2^254)
        y0 = [a + b for a, b in zip(xs[0], xs[2])]
        y1 = [a + b for a, b in zip(xs[1], xs[3])]
        v0, v1 = (sum(v << (29 * i) for i, v in enumerate(y)) for y in (y0, y1))
        s = lc_finish29([a + b for a, b in zip(y0, y1)], m)
        sv = sum(v << (29 * i) for i, v in enumerate(s))
        assert sv < 2 * m and sv % m == (v0 + v1) % m and all(0 <= x < 2**29 for x in s[:8])
        d = lc_finish29([4 * pi + a - b for pi, a, b in zip(limbs29(m), y0, y1)], m)
        dv = sum(v << (29 * i) for i, v in enumerate(d))
        assert dv < 2 * m and dv % m == (v0 - v1) % m
        op = [a + r - b for a, r, b in zip(y0, c["p8r"], y1)]
        assert all(0 <= x < 5 * 2**29 for x in op)
        w = rng.choice(tw)
        t, top = mont29_limbs(op, w, m)
        wv = sum(v << (29 * i) for i, v in enumerate(w))
        assert top < 2**64 and t < 2 * m
        assert t % m == (v0 - v1) * wv * pow(2, -261, m) % m


@pytest.mark.parametrize("m", [R, P, 2**255 - 19, 2**255 - 1, 3, 2**64 + 13])
def test_montgomery_product_bounds(m):
    rng = random.Random(m % 1000)
    edge = [0, 1, m - 1, m - 2, (2**256 - 1) % m]
    wide = [2 * m - 1] if m < 2**255 else []  # inputs below 2p (lazy reduction) stay safe
    pairs = [(x, y) for x in edge + wide for y in edge + wide] + [(rng.randrange(m), rng.randrange(m)) for _ in range(500)]
    for a, b in pairs:
        t, top = mont29(a, b, m)
        assert top < 2**64
        assert t < 2 * m
        assert t % m == a * b * pow(2, -261, m) % m
