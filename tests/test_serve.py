"""
The serving policy on the CPU: a simulated executor on a virtual clock stands in for the GPU,
so admission by memory budget, arrival order, continuous batching and the latency report are
checked without a device (zkgpu/serve.py's Scheduler is the same object the GPU server runs).
"""

from __future__ import annotations

import math

import pytest

from zkgpu.serve import ProofJob, Scheduler, ServeReport, poisson_arrivals


class VirtualClock:
    def __init__(self) -> None:
        self.t = 0.0

    def __call__(self) -> float:
        return self.t

    def sleep(self, dt: float) -> None:
        self.t += max(dt, 1e-6)


class SimExecutor:
    """Batches take a fixed time per job plus a launch cost; memory is n bytes per job."""

    def __init__(self, clock: VirtualClock, slots: int, per_row: float = 1e-6, launch_cost: float = 1e-3) -> None:
        self.clock = clock
        self.slots = slots
        self.per_row = per_row
        self.launch_cost = launch_cost
        self.inflight_bytes = 0
        self.peak_inflight = 0
        self.busy_slots: set[int] = set()
        self.launches: list[tuple[int, list[int], float]] = []

    def footprint(self, job: ProofJob) -> int:
        return job.n

    def launch(self, slot: int, jobs: list[ProofJob], batch_id: int):
        assert slot not in self.busy_slots, "a slot was given two batches at once"
        self.busy_slots.add(slot)
        nbytes = sum(self.footprint(j) for j in jobs)
        self.inflight_bytes += nbytes
        self.peak_inflight = max(self.peak_inflight, self.inflight_bytes)
        end = self.clock() + self.launch_cost + self.per_row * sum(j.n for j in jobs)
        self.launches.append((slot, [j.job_id for j in jobs], self.clock()))
        return (slot, end, nbytes, jobs)

    def done(self, handle) -> bool:
        return self.clock() >= handle[1]

    def finish(self, handle) -> None:
        slot, _, nbytes, jobs = handle
        self.busy_slots.discard(slot)
        self.inflight_bytes -= nbytes
        for j in jobs:
            j.msm_jacobian = "done"


def make_jobs(sizes: list[int], arrivals: list[float]) -> list[ProofJob]:
    return [ProofJob(i, k, t) for i, (k, t) in enumerate(zip(sizes, arrivals))]


def run(jobs, *, slots=2, budget=1 << 30, max_batch=64, per_row=1e-6):
    clock = VirtualClock()
    ex = SimExecutor(clock, slots, per_row=per_row)
    sched = Scheduler(ex, budget=budget, max_batch_jobs=max_batch, clock=clock, sleep=clock.sleep, poll_s=1e-5)
    return sched.run(jobs), ex


def test_every_job_completes_once_after_arriving():
    jobs = make_jobs([12 + (i % 9) for i in range(200)], poisson_arrivals(200, 500.0, seed=1))
    rep, ex = run(jobs)
    ids = sorted(i for _, members, _ in ex.launches for i in members)
    assert ids == list(range(200))
    for j in jobs:
        assert j.msm_jacobian == "done"
        assert j.arrival <= j.admitted <= j.completed
    assert rep.batches == len(ex.launches)


def test_admission_keeps_arrival_order():
    jobs = make_jobs([12 + (i * 5) % 9 for i in range(100)], poisson_arrivals(100, 2000.0, seed=2))
    run(jobs)
    order = sorted(jobs, key=lambda j: (j.arrival, j.job_id))
    admitted = [j.admitted for j in order]
    assert admitted == sorted(admitted)
    batches = [j.batch for j in order]
    assert batches == sorted(batches)


def test_memory_budget_is_never_exceeded():
    sizes = [20, 20, 19, 20, 18, 20, 20, 17] * 4
    budget = 3 * (1 << 20)  # three of the largest jobs at once
    jobs = make_jobs(sizes, [0.0] * len(sizes))
    rep, ex = run(jobs, slots=4, budget=budget)
    assert ex.peak_inflight <= budget
    assert rep.peak_bytes <= budget
    assert all(j.completed is not None for j in jobs)


def test_job_larger_than_budget_is_rejected():
    jobs = make_jobs([20], [0.0])
    clock = VirtualClock()
    sched = Scheduler(SimExecutor(clock, 1), budget=(1 << 20) - 1, clock=clock, sleep=clock.sleep)
    with pytest.raises(ValueError, match="more device memory"):
        sched.run(jobs)


def test_jobs_arriving_during_a_batch_join_the_next_one():
    # one slot; a long first job; ten small jobs arrive while it runs
    sizes = [20] + [12] * 10
    arrivals = [0.0] + [0.01 + 0.001 * i for i in range(10)]
    jobs = make_jobs(sizes, arrivals)
    rep, ex = run(jobs, slots=1, per_row=1e-7)  # first job takes ~0.1 s
    assert [len(m) for _, m, _ in ex.launches] == [1, 10]
    assert rep.batch_sizes == [1, 10]


def test_batches_respect_max_batch_jobs():
    jobs = make_jobs([12] * 50, [0.0] * 50)
    rep, ex = run(jobs, slots=2, max_batch=8)
    assert max(len(m) for _, m, _ in ex.launches) <= 8
    assert sum(rep.batch_sizes) == 50


def test_slots_overlap():
    jobs = make_jobs([16] * 4, [0.0, 0.0, 0.0, 0.0])
    _, ex = run(jobs, slots=2, max_batch=1)
    starts = [t for _, _, t in ex.launches]
    assert starts[0] == starts[1]  # two slots start together
    assert starts[2] > starts[0]


def test_idle_server_waits_for_arrivals():
    jobs = make_jobs([12, 12], [0.5, 1.0])
    rep, _ = run(jobs)
    assert jobs[0].admitted >= 0.5 and jobs[1].admitted >= 1.0
    assert rep.wall_s >= 1.0


def test_percentiles_are_nearest_rank():
    jobs = make_jobs([12] * 100, [0.0] * 100)
    for i, j in enumerate(jobs):
        j.completed = (i + 1) * 1e-3  # latencies 1..100 ms
    rep = ServeReport(jobs, 0.1, 1)
    assert rep.latency_ms(50) == pytest.approx(50.0)
    assert rep.latency_ms(99) == pytest.approx(99.0)
    assert rep.latency_ms(100) == pytest.approx(100.0)
    assert rep.jobs_per_s == pytest.approx(100 / 0.1)


def test_poisson_arrivals_have_the_requested_rate():
    t = poisson_arrivals(20000, 250.0, seed=3)
    assert all(b > a for a, b in zip(t, t[1:]))
    assert 20000 / t[-1] == pytest.approx(250.0, rel=0.03)
    gaps = [b - a for a, b in zip([0.0] + t, t)]
    mean = sum(gaps) / len(gaps)
    var = sum((g - mean) ** 2 for g in gaps) / len(gaps)
    assert math.sqrt(var) / mean == pytest.approx(1.0, rel=0.05)  # exponential gaps: cv = 1


def test_invalid_scheduler_arguments():
    clock = VirtualClock()
    with pytest.raises(ValueError):
        Scheduler(SimExecutor(clock, 1), budget=0)
    with pytest.raises(ValueError):
        Scheduler(SimExecutor(clock, 1), budget=1, max_batch_jobs=0)
    with pytest.raises(ValueError):
        poisson_arrivals(3, 0.0, seed=1)
