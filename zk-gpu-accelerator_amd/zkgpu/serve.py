"""
zkgpu.serve: a proof-job server with continuous batching on one GPU.

A proof job is the Groth16 prover's polynomial step at circuit size n (the
quotient h = (A B - C) / Z_H on a coset, Radix2Domain.quotient) plus three G1
MSMs of size n over a shared, device-resident SRS (the reference's single-job
flow, demo/prover_simulation.py:114-150 with msm_pippenger_gpu for the
commitments). The reference proves one job at a time, one call per transform
and per MSM; this server keeps the basis and every job's witness on the device
and runs jobs as they arrive:

- admission: a job enters only when its device workspace (quotient scratch,
  MSM plan buffers, its h vector) fits the HBM budget beside the work already
  in flight, in arrival order (no job overtakes an earlier one);
- continuous batching: every free stream slot takes all admitted jobs that are
  waiting (up to max_batch_jobs) as one batch: one QuotientBatch launch
  sequence for all their quotients and one MsmBatch for all their 3 x count
  MSMs. Batches on different slots overlap on the device, and jobs arriving
  while others run are batched as soon as a slot frees up;
- completion: a batch's slot is polled through its event; its MSM results are
  copied to the host, its workspace is released, and each job's latency is its
  completion time minus its arrival time.

Results do not depend on the batch a job lands in: every kernel computes exact
field and group values, and the MSM result, a group element, is compared in
affine form (its Jacobian coordinates may differ between runs).

The scheduling policy (Scheduler) is independent of the GPU: GpuExecutor
launches real batches, and the tests drive the same policy with a simulated
executor on the CPU.
"""

from __future__ import annotations

import math
import random
import time
from collections import deque
from collections.abc import Callable, Sequence
from dataclasses import dataclass, field
from typing import Any, Protocol

import numpy as np


@dataclass
class ProofJob:
    """One proof job: quotient inputs a, b, c ((n, 4) Montgomery rows) and three scalar vectors."""

    job_id: int
    log_n: int
    arrival: float  # seconds after the start of serving
    a: Any = None
    b: Any = None
    c: Any = None
    scalars: tuple[Any, ...] = ()
    shift: int = 7
    # filled in by the server
    h: Any = None
    msm_jacobian: np.ndarray | None = None  # (3, 3, 4) Montgomery Jacobian limbs
    admitted: float | None = None
    completed: float | None = None
    batch: int | None = None

    @property
    def n(self) -> int:
        return 1 << self.log_n

    @property
    def latency(self) -> float:
        if self.completed is None:
            raise ValueError(f"job {self.job_id} has not completed")
        return self.completed - self.arrival

    def msm_results(self) -> list[Any]:
        """The three MSM results as curve Points (None for the identity)."""
        from zkgpu.curve import from_jacobian
        from zkgpu.field256_gpu import BN254_FQ, _limbs_to_ints

        if self.msm_jacobian is None:
            raise ValueError(f"job {self.job_id} has no MSM results")
        rows = BN254_FQ.to_limbs(np.ascontiguousarray(self.msm_jacobian.reshape(-1, 4))).get()
        flat = _limbs_to_ints(rows)
        return [from_jacobian((flat[3 * k], flat[3 * k + 1], flat[3 * k + 2])) for k in range(3)]


class Executor(Protocol):
    """What the scheduler needs from a device: sizes, launches and completion checks."""

    slots: int

    def footprint(self, job: ProofJob) -> int: ...

    def launch(self, slot: int, jobs: list[ProofJob], batch_id: int) -> Any: ...

    def done(self, handle: Any) -> bool: ...

    def finish(self, handle: Any) -> None: ...


@dataclass
class ServeReport:
    jobs: list[ProofJob]
    wall_s: float
    batches: int
    batch_sizes: list[int] = field(default_factory=list)
    peak_bytes: int = 0
    budget: int = 0

    def latency_ms(self, q: float) -> float:
        lat = sorted(j.latency * 1e3 for j in self.jobs)
        if not lat:
            return float("nan")
        # nearest-rank percentile
        k = max(0, min(len(lat) - 1, math.ceil(q / 100.0 * len(lat)) - 1))
        return lat[k]

    @property
    def jobs_per_s(self) -> float:
        if not self.jobs:
            return 0.0
        first = min(j.arrival for j in self.jobs)
        last = max(j.completed for j in self.jobs if j.completed is not None)
        return len(self.jobs) / max(last - first, 1e-9)

    def summary(self) -> dict[str, Any]:
        return {
            "jobs": len(self.jobs),
            "jobs_per_s": round(self.jobs_per_s, 2),
            "p50_ms": round(self.latency_ms(50), 3),
            "p99_ms": round(self.latency_ms(99), 3),
            "mean_ms": round(sum(j.latency for j in self.jobs) * 1e3 / max(len(self.jobs), 1), 3),
            "batches": self.batches,
            "mean_batch_jobs": round(sum(self.batch_sizes) / max(len(self.batch_sizes), 1), 2),
            "peak_workspace_gib": round(self.peak_bytes / 2**30, 3),
            "hbm_budget_gib": round(self.budget / 2**30, 3),
            "wall_s": round(self.wall_s, 3),
        }


class Scheduler:
    """
    Open-loop arrivals, admission by memory budget in arrival order, and continuous
    batching over the executor's stream slots.
    """

    def __init__(self, executor: Executor, *, budget: int, max_batch_jobs: int = 64,
                 clock: Callable[[], float] = time.perf_counter, sleep: Callable[[float], None] = time.sleep,
                 poll_s: float = 50e-6) -> None:
        if budget <= 0:
            raise ValueError("the HBM budget must be positive")
        if max_batch_jobs < 1:
            raise ValueError("max_batch_jobs must be at least 1")
        self.executor = executor
        self.budget = int(budget)
        self.max_batch_jobs = int(max_batch_jobs)
        self.clock = clock
        self.sleep = sleep
        self.poll_s = poll_s

    def run(self, jobs: Sequence[ProofJob]) -> ServeReport:
        ex = self.executor
        for j in jobs:
            if ex.footprint(j) > self.budget:
                raise ValueError(f"job {j.job_id} (n = 2^{j.log_n}) needs more device memory than the whole budget")
        pending = deque(sorted(jobs, key=lambda j: (j.arrival, j.job_id)))
        waiting: deque[ProofJob] = deque()
        free_slots = list(range(ex.slots))
        inflight: list[tuple[int, Any, list[ProofJob], int]] = []
        used = peak = 0
        batches = 0
        sizes: list[int] = []
        t0 = self.clock()
        while pending or waiting or inflight:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            progressed = False
            # completions free slots and memory
            for entry in list(inflight):
                slot, handle, members, nbytes = entry
                if ex.done(handle):
                    ex.finish(handle)
                    t = self.clock() - t0
                    for j in members:
                        j.completed = t
                    used -= nbytes
                    free_slots.append(slot)
                    inflight.remove(entry)
                    progressed = True
            # admission: each free slot takes every waiting job that fits, in arrival order
            while free_slots and waiting:
                batch: list[ProofJob] = []
                nbytes = 0
                while waiting and len(batch) < self.max_batch_jobs:
                    need = ex.footprint(waiting[0])
                    if used + nbytes + need > self.budget:
                        break
                    batch.append(waiting.popleft())
                    nbytes += need
                if not batch:
                    break
                slot = free_slots.pop(0)
                t = self.clock() - t0
                for j in batch:
                    j.admitted = t
                    j.batch = batches
                handle = ex.launch(slot, batch, batches)
                inflight.append((slot, handle, batch, nbytes))
                used += nbytes
                peak = max(peak, used)
                batches += 1
                sizes.append(len(batch))
                progressed = True
            if not progressed:
                wait = self.poll_s
                if not inflight and not waiting and pending:
                    wait = max(0.0, min(pending[0].arrival - (self.clock() - t0), 0.01))
                if wait > 0:
                    self.sleep(wait)
        wall = self.clock() - t0
        return ServeReport(list(jobs), wall, batches, sizes, peak, self.budget)


# ---- the GPU executor -------------------------------------------------------------------


class GpuExecutor:
    """Launches batches on `slots` HIP streams: one QuotientBatch and one MsmBatch per batch."""

    def __init__(self, srs: Any, *, slots: int = 2, windows: Callable[[int], int] | None = None) -> None:
        from zkgpu import gpu_backend as gb

        gb.require_gpu("GpuExecutor")
        if slots < 1:
            raise ValueError("at least one stream slot is needed")
        self._gb = gb
        self.srs = srs
        self.slots = slots
        self.streams = [gb.Stream() for _ in range(slots)]
        self.windows = windows
        self._footprints: dict[int, int] = {}

    def _window(self, n: int) -> int:
        return self.windows(n) if self.windows else 0

    def footprint(self, job: ProofJob) -> int:
        """Device bytes the job adds to a batch: its h rows, its quotient workspace and its three
        MSMs' plan buffers, the latter sized over the SRS (whose window tables change the entry and
        bucket counts) for the accumulation chunk that needs the most memory among those a batch
        may use (the chunk is the tuned one of the batch's largest MSM)."""
        from zkgpu.batch import msm_workspace_bytes, quotient_workspace_bytes, tuned_chunks

        fp = self._footprints.get(job.log_n)
        if fp is None:
            n = job.n
            msm = max(msm_workspace_bytes([n] * 3, [self._window(n)] * 3, srs=self.srs, chunk_len=ch)
                      for ch in tuned_chunks())
            fp = n * 32 + quotient_workspace_bytes([job.log_n]) + msm + 3 * 96 + 4096
            self._footprints[job.log_n] = fp
        return fp

    def launch(self, slot: int, jobs: list[ProofJob], batch_id: int) -> Any:
        from zkgpu.batch import MsmBatch, QuotientBatch

        gb = self._gb
        with gb.stream(self.streams[slot].handle):
            for j in jobs:
                j.h = gb.empty((j.n, 4))
            q = QuotientBatch([j.a for j in jobs], [j.b for j in jobs], [j.c for j in jobs],
                              shifts=[j.shift for j in jobs], h=[j.h for j in jobs])
            scalars = [s for j in jobs for s in j.scalars]
            m = MsmBatch(self.srs, scalars, windows=[self._window(int(s.shape[0])) for s in scalars])
            q.launch()
            m.launch()
            ev = gb.Event().record()
        return (slot, ev, jobs, q, m)

    def done(self, handle: Any) -> bool:
        return handle[1].done()

    def finish(self, handle: Any) -> None:
        slot, _, jobs, q, m = handle
        with self._gb.stream(self.streams[slot].handle):
            host = m.out.get()
        for k, j in enumerate(jobs):
            j.msm_jacobian = host[3 * k : 3 * k + 3].copy()

    def synchronize(self) -> None:
        for s in self.streams:
            s.synchronize()


class ProofServer:
    """A GpuExecutor under the Scheduler: serve(jobs) runs them as they arrive and reports latencies."""

    def __init__(self, srs: Any, *, hbm_budget: int | None = None, slots: int = 2, max_batch_jobs: int = 64,
                 windows: Callable[[int], int] | None = None) -> None:
        self.executor = GpuExecutor(srs, slots=slots, windows=windows)
        if hbm_budget is None:
            free, _ = self.executor._gb.mem_info()
            hbm_budget = int(free * 0.5)
        self.scheduler = Scheduler(self.executor, budget=hbm_budget, max_batch_jobs=max_batch_jobs)

    def serve(self, jobs: Sequence[ProofJob]) -> ServeReport:
        return self.scheduler.run(jobs)


# ---- workloads ------------------------------------------------------------------------------


def poisson_arrivals(count: int, rate: float, seed: int) -> list[float]:
    """Arrival times (seconds) of a Poisson process of the given rate (jobs per second)."""
    if rate <= 0:
        raise ValueError("the arrival rate must be positive")
    rng = random.Random(seed)
    t, out = 0.0, []
    for _ in range(count):
        t += rng.expovariate(rate)
        out.append(t)
    return out


class WitnessPools:
    """
    Device-resident random inputs that jobs take windows of: Montgomery rows for the quotient
    inputs and witness-like plain scalars for the MSMs (each job reads its own offsets, so no
    two jobs of a run see the same inputs unless their offsets coincide).
    """

    def __init__(self, rows: int, seed: int = 0x5EED05) -> None:
        from zkgpu import gpu_backend as gb
        from zkgpu.workloads import uniform_rows, witness_rows_fast

        gb.require_gpu("WitnessPools")
        self.rows = rows
        self.evals = gb.asarray(uniform_rows(rows, seed))
        self.scalars = gb.asarray(witness_rows_fast(rows, seed + 1))

    def job(self, job_id: int, log_n: int, arrival: float, shift: int = 7) -> ProofJob:
        n = 1 << log_n
        if n > self.rows:
            raise ValueError(f"a job of 2^{log_n} rows does not fit pools of {self.rows} rows")
        rng = random.Random(job_id * 7919 + log_n)
        span = self.rows - n
        offs = [rng.randint(0, span) for _ in range(6)]
        ev = [self.evals[o : o + n] for o in offs[:3]]
        sc = tuple(self.scalars[o : o + n] for o in offs[3:])
        return ProofJob(job_id, log_n, arrival, ev[0], ev[1], ev[2], sc, shift)


def served_workload(pools: WitnessPools, count: int, rate: float | None, *, lo: int = 12, hi: int = 20,
                    seed: int = 0x5EED06) -> list[ProofJob]:
    """count jobs with log2 n uniform in [lo, hi]; Poisson arrivals at `rate` jobs/s, or all at 0 if None."""
    rng = random.Random(seed)
    sizes = [rng.randint(lo, hi) for _ in range(count)]
    times = poisson_arrivals(count, rate, seed + 1) if rate else [0.0] * count
    return [pools.job(i, k, t) for i, (k, t) in enumerate(zip(sizes, times))]


def run_plain(job: ProofJob, srs_points: Any, *, srs: Any = None) -> None:
    """
    The per-call path for one job (the reference's structure): one quotient call and three
    separate MSM calls, each waiting for its result before the next starts. The MSMs run over
    the raw points (zk_msm_g1), or, given srs, as one-job batches over the SRS and its window
    tables (zk_msm_g1_batch: the same precomputation the batched form uses, without batching).
    """
    import ctypes

    from zkgpu import gpu_backend as gb
    from zkgpu.domain import Radix2Domain

    dom = Radix2Domain(job.n)
    job.h = dom.quotient(job.a, job.b, job.c, offset=job.shift)
    outs = []
    for s in job.scalars:
        out = gb.empty((3, 4))
        if srs is None:
            gb.call("zk_msm_g1", srs_points.ptr, s.ptr, job.n, 0, 0, out.ptr, gb.current_stream())
        else:
            ptrs, sizes = (ctypes.c_void_p * 1)(s.ptr), (ctypes.c_int64 * 1)(job.n)
            gb.call("zk_msm_g1_batch", srs.handle, 1, ptrs, sizes, None, out.ptr, gb.current_stream())
        outs.append(out.get())
    job.msm_jacobian = np.stack(outs)
