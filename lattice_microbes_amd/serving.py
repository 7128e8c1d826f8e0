"""Continuous-batching server for MPD-RDME simulation requests.

No reference counterpart (the reference runs one simulation per process).
Requests (a model, a seed, a step count) arrive over time; the server keeps
every admitted job's lattice, tables and step counter resident on the GPU
and advances all admitted jobs together through an Ensemble, in chunks of
steps.  Between chunks it retires finished jobs and admits waiting ones while
their device footprint fits the HBM budget, so requests that arrive while
others run join the running batch.  A job's result is the same as running it
alone: each member of an Ensemble is bit-identical to its plain run.
"""
from __future__ import annotations

import hashlib
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from .ensemble import Ensemble
from .models import ModelSpec
from .solver import MpdRdmeSolver


@dataclass
class Job:
    id: int
    spec: ModelSpec
    seed: int
    steps: int
    arrival: float                     # seconds after the start of the run
    replicate: int = 0
    # filled in by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    done_steps: int = 0
    counts: Optional[np.ndarray] = None
    digest: Optional[str] = None
    solver: Optional[MpdRdmeSolver] = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def site_steps(self) -> int:
        return self.spec.N * self.steps


def footprint(spec: ModelSpec) -> int:
    """Device bytes of a job: two lattice buffers, site types, the reaction
    list (N/8 entries of index, P particle bytes and type), tables."""
    tables = 4 * (spec.nST * spec.nST * spec.nS + spec.nST * spec.nS * (spec.nS + 3)) + 64 * spec.R
    rlist = max(64 * 1024, spec.N // 8) * (spec.P + 5)
    return 2 * spec.N * spec.P + spec.N + rlist + tables + (1 << 20)


def lattice_digest(solver: MpdRdmeSolver) -> str:
    return hashlib.sha256(solver.getLatticeNative().tobytes()).hexdigest()


def run_alone(job: Job, device: int = 0) -> str:
    """The job's final-lattice digest when run by itself (the plain path)."""
    s = MpdRdmeSolver(device)
    job.spec.load_into(s, replicate=job.replicate, seed=job.seed)
    s.runTimesteps(job.steps)
    d = lattice_digest(s)
    s.close()
    return d


class Server:
    """Admits, batches and retires jobs; see the module docstring."""

    def __init__(self, device: int = 0, hbm_budget: int = 64 << 30, chunk: int = 250, graph: bool = True,
                 keep_digest: bool = True, max_active: Optional[int] = None, max_sites: Optional[int] = None,
                 policy: str = "fcfs", starvation_s: float = 5.0,
                 clock: Callable[[], float] = time.perf_counter):
        """hbm_budget: device bytes the admitted jobs may hold (footprint());
        max_active: at most this many jobs at once; max_sites: at most this
        many lattice sites stepped per batched step (beyond a few whole-cell
        lattices the GPU is saturated and more jobs only slow every one);
        policy: "fcfs" admits in arrival order, "sjf" the waiting job with the
        least work (sites x steps) first, except that a job waiting longer
        than starvation_s goes first."""
        if policy not in ("fcfs", "sjf"):
            raise ValueError(f"unknown admission policy {policy!r}")
        self.device = device
        self.hbm_budget = int(hbm_budget)
        self.chunk = int(chunk)
        self.graph = graph
        self.keep_digest = keep_digest
        self.max_active = max_active
        self.max_sites = max_sites
        self.policy = policy
        self.starvation_s = float(starvation_s)
        self.clock = clock

    def _candidates(self, waiting: List[Job], now: float) -> List[Job]:
        """Waiting jobs in the order admission considers them."""
        if self.policy == "fcfs":
            return list(waiting)
        return sorted(waiting, key=lambda j: (now - j.arrival <= self.starvation_s, j.site_steps, j.arrival))

    def _admit(self, job: Job, ens: Ensemble, now: float) -> None:
        s = MpdRdmeSolver(self.device)
        job.spec.load_into(s, replicate=job.replicate, seed=job.seed)
        job.solver = s
        job.admitted = now
        ens.add(s)

    def _retire(self, job: Job, ens: Ensemble, now: float) -> None:
        s = job.solver
        ens.remove(s)
        job.counts = s.getSpeciesCounts()
        if self.keep_digest:
            job.digest = lattice_digest(s)
        job.finished = now
        s.close()
        job.solver = None

    def run(self, jobs: List[Job]) -> Dict[str, float]:
        """Serves jobs (arrival times relative to the call); returns summary
        statistics.  Every job gets finished, counts and digest set."""
        order = sorted(jobs, key=lambda j: j.arrival)
        waiting: List[Job] = []
        active: List[Job] = []
        used = 0
        sites = 0
        nxt = 0
        ens = Ensemble(self.device, graph=self.graph)
        busy = 0.0
        t0 = self.clock()
        try:
            while nxt < len(order) or waiting or active:
                # Launch the next chunk for the running jobs, then do the host
                # work (arrivals, uploads of admitted jobs) while it runs.
                k = min([self.chunk] + [j.steps - j.done_steps for j in active]) if active else 0
                tb = self.clock()
                if k:
                    ens.run(k, synchronize=False)
                now = self.clock() - t0
                while nxt < len(order) and order[nxt].arrival <= now:
                    waiting.append(order[nxt])
                    nxt += 1
                # Admission, in the policy's order, of every candidate whose
                # footprint and sites fit (a job larger than the limits runs
                # when nothing else does; under fcfs a job that does not fit
                # holds back the ones behind it); admitted jobs join from the
                # next chunk on.
                joining = []
                for job in self._candidates(waiting, now):
                    if self.max_active is not None and len(active) + len(joining) >= self.max_active:
                        break
                    alone = not (active or joining)
                    fits = used + footprint(job.spec) <= self.hbm_budget and \
                        (self.max_sites is None or sites + job.spec.N <= self.max_sites)
                    if not (fits or alone):
                        if self.policy == "fcfs":
                            break
                        continue
                    waiting.remove(job)
                    self._admit(job, ens, self.clock() - t0)
                    used += footprint(job.spec)
                    sites += job.spec.N
                    joining.append(job)
                if k:
                    ens.synchronize()
                    busy += self.clock() - tb
                    for j in active:
                        j.done_steps += k
                    for j in [j for j in active if j.done_steps >= j.steps]:
                        self._retire(j, ens, self.clock() - t0)
                        used -= footprint(j.spec)
                        sites -= j.spec.N
                        active.remove(j)
                active.extend(joining)
                if not active and not waiting and nxt < len(order):
                    time.sleep(max(0.0, min(order[nxt].arrival - (self.clock() - t0), 0.01)))
        finally:
            for j in active:
                if j.solver is not None:
                    j.solver.close()
            ens.close()
        wall = self.clock() - t0
        return summarize(order, wall, busy)


def serve_plain(jobs: List[Job], device: int = 0, clock: Callable[[], float] = time.perf_counter,
                keep_digest: bool = True, step_sync: bool = False) -> Dict[str, float]:
    """The plain baseline: one simulation at a time, in arrival order.  With
    step_sync the host waits for every step, as the reference's
    runTimestep does (MpdRdmeSolver.cu:1091)."""
    order = sorted(jobs, key=lambda j: j.arrival)
    t0 = clock()
    busy = 0.0
    for job in order:
        now = clock() - t0
        if job.arrival > now:
            time.sleep(job.arrival - now)
        job.admitted = clock() - t0
        s = MpdRdmeSolver(device)
        job.spec.load_into(s, replicate=job.replicate, seed=job.seed)
        tb = clock()
        if step_sync:
            for _ in range(job.steps):
                s.runTimesteps(1)
        else:
            s.runTimesteps(job.steps)
        busy += clock() - tb
        job.done_steps = job.steps
        job.counts = s.getSpeciesCounts()
        if keep_digest:
            job.digest = lattice_digest(s)
        job.finished = clock() - t0
        s.close()
    wall = clock() - t0
    return summarize(order, wall, busy)


def summarize(jobs: List[Job], wall: float, busy: float) -> Dict[str, float]:
    lat = np.array([j.latency for j in jobs])
    site_steps = float(sum(j.site_steps for j in jobs))
    return {
        "jobs": len(jobs),
        "wall_s": wall,
        "gpu_busy_s": busy,
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "site_steps": site_steps,
        "site_steps_per_s": site_steps / wall,
    }


def poisson_jobs(templates: List[ModelSpec], n: int, rate_hz: float, steps_range=(1000, 10000),
                 seed: int = 1) -> List[Job]:
    """n requests with exponential inter-arrival times (mean 1/rate_hz), a
    uniformly drawn template, step count and seed each."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for k in range(n):
        t += float(rng.exponential(1.0 / rate_hz))
        spec = templates[int(rng.integers(len(templates)))]
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        jobs.append(Job(id=k, spec=spec, seed=int(rng.integers(1, 2**31)), steps=steps, arrival=t, replicate=k))
    return jobs
