"""
zkgpu.prover: one Groth16 proof's polynomial and G1 work on the device, timed per step.

The reference's demo (demo/prover_simulation.py:1-40, 114-223) models a prover for an R1CS of m
constraints: the quotient h = (A B - C) / Z_H on a domain H of n >= m + 2 points (three inverse
NTTs, three coset NTTs, a pointwise step, one inverse coset NTT: seven NTTs) and four G1 MSMs,
sum z_i [A_i], sum z_i [B_i] (N = m witness values each), sum z_i [L_i] over the private witness
values, and sum h_i [H_i] over h's first n - 1 coefficients. make_inputs builds the same stand-ins
(random A z and B z on H with C z = (A z)(B z), so h has degree at most n - 2; a random witness
with z_0 = 1; one arithmetic progression of points P_0 + i S shared by the four queries) from the
same seeded generator, so the inputs equal the reference's for a given seed.

GpuProver keeps what a prover builds once per circuit on the device (the basis as an SRS with
its converted records, the domain's tables inside the fused quotient) and runs a proof as the
reference's prove_gpu does: Python ints in (the constraint evaluations and the witness), h's
coefficients and four affine points out. The quotient is zk_groth16_quotient (seven transforms,
the pointwise step fused into their passes); h feeds the H MSM on the device, without a round
trip through Python ints.
"""

from __future__ import annotations

import random
import time
from collections.abc import Callable
from dataclasses import dataclass
from typing import Any, TypeVar

import numpy as np

from zkgpu import gpu_backend as _gb
from zkgpu.curve import Point
from zkgpu.domain import Radix2Domain
from zkgpu.field import FIELD_MODULUS as R

NUM_PUBLIC = 1  # public inputs besides the constant 1 every witness starts with
MSM_NAMES = ("A", "B", "L", "H")
T = TypeVar("T")


@dataclass(frozen=True)
class ProofInputs:
    """Everything one proof needs, for a circuit of a given size (demo/prover_simulation.py:66-85)."""

    constraints: int
    domain: Radix2Domain
    a: list[int]  # A z on H (zero on padding rows), likewise b and c
    b: list[int]
    c: list[int]
    witness: list[int]  # z: N = constraints values, z[0] = 1
    keys: list[int]  # the basis is P_i = keys[i] G: max(N, n - 1) points

    def msm_sizes(self) -> dict[str, int]:
        n = self.domain.size
        return {"A": self.constraints, "B": self.constraints, "L": self.constraints - NUM_PUBLIC - 1, "H": n - 1}

    def msm_scalars(self, h: list[int]) -> dict[str, list[int]]:
        """The scalars of the four G1 MSMs, given the quotient's coefficients h (:83-85)."""
        private = self.witness[NUM_PUBLIC + 1 :]
        return {"A": self.witness, "B": self.witness, "L": private, "H": h[: self.domain.size - 1]}


def make_inputs(constraints: int, seed: int = 0) -> ProofInputs:
    """The demo's inputs for `constraints` constraints (demo/prover_simulation.py:88-104): the same
    draws from random.Random(seed); the point progression P_0 + i S is kept as its multiples of G,
    k_0 + i s mod r, so the basis is generated on the device (workloads.srs_points)."""
    if constraints < NUM_PUBLIC + 2:
        raise ValueError(f"need at least {NUM_PUBLIC + 2} constraints, got {constraints}")
    rng = random.Random(seed)
    rows = constraints + NUM_PUBLIC + 1  # a row per public input and one for z[0]
    domain = Radix2Domain.at_least(rows)
    padding = [0] * (domain.size - rows)
    a = [rng.randrange(R) for _ in range(rows)] + padding
    b = [rng.randrange(R) for _ in range(rows)] + padding
    c = [x * y % R for x, y in zip(a, b, strict=True)]
    witness = [1] + [rng.randrange(R) for _ in range(constraints - 1)]
    step = rng.randrange(1, R)  # S = step G
    first = rng.randrange(1, R)  # P_0 = first G
    count = max(constraints, domain.size - 1)
    keys = [(first + i * step) % R for i in range(count)]
    return ProofInputs(constraints, domain, a, b, c, witness, keys)


@dataclass(frozen=True)
class Run:
    """One proof's timings (ms) and results (demo/prover_simulation.py:157-173)."""

    quotient_ms: float
    msm_ms: dict[str, float]
    h: list[int]
    msm_results: dict[str, Point]

    @property
    def msm_total_ms(self) -> float:
        return sum(self.msm_ms.values())

    @property
    def total_ms(self) -> float:
        return self.quotient_ms + self.msm_total_ms


def _timed(fn: Callable[[], T]) -> tuple[float, T]:
    t0 = time.perf_counter()
    out = fn()  # every step returns host data (ints or a Point), so the GPU is done by now
    return (time.perf_counter() - t0) * 1e3, out


def check_quotient(h: list[int], n: int) -> None:
    """A satisfied circuit makes A B - C divisible by Z_H, so h has degree at most n - 2 (:181-186)."""
    if len(h) != n or h[-1] != 0:
        raise RuntimeError("the quotient is not a polynomial of degree <= n - 2")


class GpuProver:
    """Per-circuit device state (the basis as an SRS, built once, untimed) and one proof's work."""

    def __init__(self, inputs: ProofInputs, *, window_tables: tuple[int, ...] = ()) -> None:
        from zkgpu.batch import SRS
        from zkgpu.field256_gpu import BN254_FR, _ints_to_limbs
        from zkgpu.workloads import srs_points

        _gb.require_gpu("GpuProver")
        self.inputs = inputs
        self.fr = BN254_FR
        self.srs = SRS(srs_points(_ints_to_limbs(inputs.keys)))
        if window_tables:
            self.srs.precompute(window_tables)
        self.shift = inputs.domain.coset_offset
        if self.shift is None:
            raise ValueError("the domain has no coset offset")

    def quotient(self) -> tuple[Any, list[int]]:
        """h's n coefficients: Montgomery rows on the device and Python ints (ints in and out, as
        the reference's quotient_gpu, :142-150)."""
        fr = self.fr
        d = self.inputs.domain
        a, b, c = (fr.from_ints(v) for v in (self.inputs.a, self.inputs.b, self.inputs.c))
        h = _gb.empty((d.size, 4))
        shift = _gb.limbs_of(self.shift)
        _gb.call("zk_groth16_quotient", a.ptr, b.ptr, c.ptr, h.ptr, d.size.bit_length() - 1, shift, _gb.current_stream())
        return h, fr.to_ints(h)

    def prove(self) -> Run:
        from zkgpu.msm_gpu import msm_pippenger_gpu

        fr = self.fr
        inputs = self.inputs
        n = inputs.domain.size
        quotient_ms, (h_rows, h) = _timed(self.quotient)
        check_quotient(h, n)
        witness = None
        msm_ms: dict[str, float] = {}
        results: dict[str, Point] = {}
        for name, size in inputs.msm_sizes().items():
            def step(name: str = name, size: int = size) -> Point:
                nonlocal witness
                if name == "H":
                    scalars = fr.to_limbs(h_rows[:size])  # plain limbs, straight from the device
                else:
                    if witness is None:
                        witness = _gb.asarray(_witness_rows(inputs.witness))
                    scalars = witness[: inputs.constraints] if name != "L" else witness[NUM_PUBLIC + 1 :]
                return msm_pippenger_gpu(self.srs, None, scalars)

            msm_ms[name], results[name] = _timed(step)
        return Run(quotient_ms, msm_ms, h, results)


def _witness_rows(witness: list[int]) -> np.ndarray:
    from zkgpu.field256_gpu import _rows_from_ints

    rows, _ = _rows_from_ints(witness, R, "witness")
    return np.ascontiguousarray(rows, dtype=np.uint64)


def prove_gpu(inputs: ProofInputs, *, window_tables: tuple[int, ...] = ()) -> Run:
    """The GPU proof after one untimed warm-up proof, as a prover that runs repeatedly sees it
    (demo/prover_simulation.py:199-223)."""
    prover = GpuProver(inputs, window_tables=window_tables)
    prover.prove()
    _gb.synchronize()
    return prover.prove()
