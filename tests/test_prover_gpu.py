"""
The Groth16 prover's work on the device (zkgpu.prover) against the oracle, as the reference's demo
checks its GPU proof against its CPU proof (demo/prover_simulation.py:235-245): the same quotient
coefficients and the same four G1 MSM results. The oracle's quotient is the restated
Radix2Domain composition; the MSMs are checked against the oracle's Pippenger (A) and the
closed form (sum k_i s_i mod r) G over the basis P_i = k_i G (all four).
"""

import numpy as np
import pytest

from zkgpu.field import FIELD_MODULUS as R
from zkgpu.prover import MSM_NAMES, NUM_PUBLIC, check_quotient, make_inputs, prove_gpu
from zkref import fast
from zkref.curve import G
from zkref.domain import groth16_quotient
from zkref.field import ints_to_limbs
from zkref.msm import msm_pippenger


def test_inputs_follow_the_demo():
    inp = make_inputs(1000, seed=0)
    n = inp.domain.size
    assert n == 1024 and len(inp.a) == len(inp.b) == len(inp.c) == n
    assert all(z == x * y % R for x, y, z in zip(inp.a, inp.b, inp.c))
    assert inp.a[1002:] == [0] * 22 and inp.witness[0] == 1 and len(inp.witness) == 1000
    step = (inp.keys[1] - inp.keys[0]) % R
    assert all((inp.keys[i + 1] - inp.keys[i]) % R == step for i in range(len(inp.keys) - 1))
    assert len(inp.keys) == max(1000, n - 1)
    assert inp.msm_sizes() == {"A": 1000, "B": 1000, "L": 1000 - NUM_PUBLIC - 1, "H": n - 1}


@pytest.mark.gpu
@pytest.mark.parametrize("constraints", [100, 3000])
def test_gpu_proof_agrees_with_the_oracle(constraints):
    inp = make_inputs(constraints, seed=7)
    run = prove_gpu(inp)
    shift = inp.domain.coset_offset
    want_h = groth16_quotient(inp.a, inp.b, inp.c, shift)
    assert run.h == want_h
    check_quotient(run.h, inp.domain.size)
    scalars = inp.msm_scalars(want_h)
    for name in MSM_NAMES:
        s = scalars[name]
        e = fast.dot_mod_r(ints_to_limbs(inp.keys[: len(s)]), ints_to_limbs(s))
        assert run.msm_results[name].affine() == (G * e).affine(), name
    a_points = [G * k for k in inp.keys[: len(scalars["A"])]]
    assert run.msm_results["A"].affine() == msm_pippenger(a_points, scalars["A"]).affine()
    assert run.total_ms > 0 and set(run.msm_ms) == set(MSM_NAMES)
