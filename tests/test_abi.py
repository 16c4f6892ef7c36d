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
    assert c["p32"] == [(m >> (32 * i)) & 0xFFFFFFFF for i in range(8)]


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
