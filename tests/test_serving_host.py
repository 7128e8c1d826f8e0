"""Admission logic of the continuous-batching server on the CPU: the
Ensemble and the solvers are replaced by test doubles that record what the
server does (which jobs are resident, in which order they were admitted), so
the policies and limits are checked without a GPU.  The GPU tests
(test_gpu_parity.py, test_gpu_configs.py) check that served results equal
solo runs."""
from dataclasses import dataclass

import numpy as np
import pytest

from lattice_microbes_amd import serving


@dataclass(eq=False)
class FakeSpec:
    name: str
    N: int
    P: int = 8
    nS: int = 3
    nST: int = 1
    R: int = 2

    def load_into(self, solver, replicate, seed):
        solver.spec = self


class FakeSolver:
    def __init__(self, device=0):
        self.spec = None
        self.closed = False

    def getSpeciesCounts(self):
        return np.zeros(self.spec.nS, np.uint32)

    def getLatticeNative(self):
        return np.zeros(4, np.uint32)

    def close(self):
        self.closed = True


class FakeEnsemble:
    log = []

    def __init__(self, device=0, graph=True):
        self.members = []

    def add(self, s):
        self.members.append(s)
        FakeEnsemble.log.append(("add", s.spec.name))

    def remove(self, s):
        self.members.remove(s)

    def run(self, k, synchronize=True):
        FakeEnsemble.log.append(("run", k, sum(m.spec.N for m in self.members), len(self.members)))

    def synchronize(self):
        pass

    def close(self):
        pass


@pytest.fixture
def fakes(monkeypatch):
    monkeypatch.setattr(serving, "Ensemble", FakeEnsemble)
    monkeypatch.setattr(serving, "MpdRdmeSolver", FakeSolver)
    FakeEnsemble.log = []
    return FakeEnsemble


def jobs_at_zero(sizes_steps):
    return [serving.Job(id=k, spec=FakeSpec(f"j{k}", n), seed=k, steps=st, arrival=0.0, replicate=k)
            for k, (n, st) in enumerate(sizes_steps)]


def test_candidates_order():
    jobs = jobs_at_zero([(1000, 50), (10, 50), (100, 50), (10, 10)])
    jobs[2].arrival = -10.0  # has waited 10 s at now = 0
    fcfs = serving.Server(policy="fcfs")._candidates(jobs, 0.0)
    assert [j.id for j in fcfs] == [0, 1, 2, 3]
    sjf = serving.Server(policy="sjf", starvation_s=5.0)._candidates(jobs, 0.0)
    # the starving job first, then the least work (sites x steps)
    assert [j.id for j in sjf] == [2, 3, 1, 0]
    with pytest.raises(ValueError):
        serving.Server(policy="lifo")


def test_sjf_admits_smallest_work_within_site_limit(fakes):
    jobs = jobs_at_zero([(1000, 20), (300, 20), (300, 5), (100, 40), (600, 10)])
    serving.Server(max_sites=1000, chunk=5, policy="sjf", clock=iter(np.arange(0, 100, 0.01)).__next__).run(jobs)
    adds = [e[1] for e in fakes.log if e[0] == "add"]
    # work: j0 20000, j1 6000, j2 1500, j3 4000, j4 6000; under 1000 sites:
    # j2 (300) + j3 (100) + j1 (300) fit (700), j4 (600) does not, j0 waits
    assert adds[:3] == ["j2", "j3", "j1"]
    for e in fakes.log:
        if e[0] == "run":
            assert e[2] <= 1000 or e[3] == 1, f"batch over the site limit: {e}"
    assert all(j.done_steps == j.steps and j.finished is not None for j in jobs)
    assert all(j.solver is None for j in jobs)


def test_fcfs_holds_back_behind_a_job_that_does_not_fit(fakes):
    jobs = jobs_at_zero([(600, 10), (600, 10), (100, 10)])
    serving.Server(max_sites=1000, chunk=10, policy="fcfs", clock=iter(np.arange(0, 100, 0.01)).__next__).run(jobs)
    runs = [e for e in fakes.log if e[0] == "run"]
    # j1 does not fit beside j0, and j2 may not overtake it
    assert runs[0][3] == 1 and runs[0][2] == 600
    adds = [e[1] for e in fakes.log if e[0] == "add"]
    assert adds == ["j0", "j1", "j2"]


def test_oversized_job_runs_alone(fakes):
    jobs = jobs_at_zero([(5000, 10), (10, 10)])
    serving.Server(max_sites=1000, chunk=10, policy="sjf", clock=iter(np.arange(0, 100, 0.01)).__next__).run(jobs)
    runs = [e for e in fakes.log if e[0] == "run"]
    assert any(e[2] == 5000 and e[3] == 1 for e in runs)
    assert all(j.done_steps == j.steps for j in jobs)
