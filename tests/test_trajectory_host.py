"""The host side of generateTrajectory, runSolver and the hook protocol, on the
CPU: the solver's device calls are replaced by the oracle (a test double, so
that the event schedule, the hook return codes, the file records and the
max-count bookkeeping are checked without a GPU).  The same cases run on the
GPU, through the HIP library, in test_gpu_trajectory.py."""
import warnings

import numpy as np
import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd.models import ModelSpec
from lattice_microbes_amd.runner import runSolver
from lattice_microbes_amd.simfile import SimulationFile
from lattice_microbes_amd.solver import _last_step
import oracle
from trajectory_cases import (HOOK_HI, HOOK_LWI, HOOK_STEPS, HOOK_WI, SEED, EditingHooks, expected_hooked,
                              expected_plain, oracle_with_rate, small_spec, type_counts)


class OracleSolver(lmx.MpdRdmeSolver):
    """MpdRdmeSolver whose device calls go to the oracle instead of the HIP
    library: a test double for the host logic only."""

    def __init__(self, device=0):  # no lmx_create: nothing touches a device
        self._h = None
        self.device = device
        self.parameters, self.replicate = {}, 0
        self.numberSpecies = self.numberReactions = self.numberSiteTypes = 0
        self.shape, self.particlesPerSite, self.tau, self.latticeSpacing = (0, 0, 0), 0, 0.0, 0.0
        self.steps_run = []

    def close(self):
        pass

    def initialize(self, replicate, parameters, resources=None, boundary="periodic"):
        self.parameters = {k: str(v) for k, v in parameters.items()}
        self.replicate = int(replicate)
        self.tau = float(self.parameters["timestep"])
        self.seed_top = int(float(self.parameters.get("seed", "0")))

    def buildModel(self, nS, R, init, types, K, S, D, kCols=1):
        self.numberSpecies, self.numberReactions = int(nS), int(R)
        self._rx = (np.asarray(types, np.uint32), np.asarray(K, np.float64).reshape(R, kCols)[:, 0],
                    np.asarray(S, np.int32).reshape(nS, R), np.asarray(D, np.uint32).reshape(nS, R))

    def buildDiffusionModel(self, nST, DF, RL, X, Y, Z, P, bpp, spacing, lattice, sites, rowMajorData=True):
        types, K, S, D = self._rx
        self.spec = ModelSpec("file", (X, Y, Z), spacing, self.tau, P, self.numberSpecies, nST, types, K, S, D,
                              np.asarray(DF, np.float64).reshape(nST, nST, self.numberSpecies),
                              np.asarray(RL, np.uint32).reshape(self.numberReactions, nST),
                              np.asarray(lattice, np.uint8).reshape(X, Y, Z, P),
                              np.asarray(sites, np.uint8).reshape(X, Y, Z))
        self.numberSiteTypes, self.shape, self.particlesPerSite, self.latticeSpacing = nST, (X, Y, Z), P, spacing
        self.orc = oracle.Oracle(self.spec, seed_top=self.seed_top, replicate=self.replicate)

    def setReactionRate(self, rxid, rate):
        self.spec, self.orc = oracle_with_rate(self.spec, self.orc, rxid, rate, self.replicate)

    def runTimesteps(self, nsteps, synchronize=True):
        self.steps_run.append(int(nsteps))
        assert self.orc.run(int(nsteps)) == 0

    def getLatticeNative(self):
        return self.orc.native()

    def setLatticeNative(self, native):
        self.orc.lattice[:] = np.asarray(native, np.uint32).reshape(-1)

    def getSites(self):
        X, Y, Z = self.shape
        return np.ascontiguousarray(self.orc.sites.reshape(Z, Y, X).transpose(2, 1, 0))

    def setSites(self, sites):
        self.orc.sites[:] = np.ascontiguousarray(np.asarray(sites, np.uint8).transpose(2, 1, 0)).reshape(-1)

    def getSpeciesCounts(self):
        return self.orc.species_counts()

    def getOverflowStats(self):
        return self.orc.stats.overflow_entries, self.orc.stats.overflow_steps


class EditingOracleSolver(EditingHooks, OracleSolver):
    pass


def test_last_step_matches_the_reference_loop():
    # while (time < maxTime) { ++t; time = t * tau; } (MpdRdmeSolver.cu:572-612)
    for tau, mt in [(1e-4, 0.1), (0.1, 0.3), (3e-3, 1.0), (1.0, 0.0), (0.5, 1.0), (2e-5, 0.0123), (7e-3, 0.7)]:
        t, time = 0, 0.0
        while time < mt:
            t += 1
            time = t * tau
        assert _last_step(tau, mt) == t


def test_plain_trajectory_records(tmp_path):
    spec = small_spec()
    path = str(tmp_path / "traj.lm")
    steps, wi, lwi = 23, 5, 7
    spec.to_file(path, {"seed": SEED, "maxTime": repr(steps * spec.tau), "writeInterval": wi,
                        "latticeWriteInterval": lwi})
    solver = OracleSolver()
    assert runSolver(path, 3, solver)
    # steps between events run as one batch
    assert solver.steps_run == [5, 2, 3, 4, 1, 5, 1, 2]
    counts, frames, maxp = expected_plain(spec, 3, steps, wi, lwi)
    with SimulationFile(path, writable=False) as f:
        np.testing.assert_array_equal(f.getSpeciesCountTimes(3), [t * spec.tau for t in sorted(counts)])
        np.testing.assert_array_equal(f.getSpeciesCounts(3), np.array([counts[t] for t in sorted(counts)]))
        np.testing.assert_array_equal(f.getLatticeTimes(3), [t * spec.tau for t in sorted(frames)])
        for k, t in enumerate(sorted(frames)):
            np.testing.assert_array_equal(f.getLattice(3, k), frames[t])
        np.testing.assert_array_equal(f.getReplicateDataset(3, "MaxParticleCounts"), maxp)
        np.testing.assert_array_equal(f.getReplicateDataset(3, "MaxSiteCounts"), type_counts(spec.sites, spec.nST))


def test_zero_intervals_fire_every_step(tmp_path):
    spec = small_spec()
    path = str(tmp_path / "every.lm")
    spec.to_file(path, {"seed": SEED, "maxTime": repr(4 * spec.tau), "writeInterval": 0,
                        "latticeWriteInterval": 0})
    solver = OracleSolver()
    assert runSolver(path, 0, solver)
    assert solver.steps_run == [1, 1, 1, 1]
    with SimulationFile(path, writable=False) as f:
        assert len(f.getSpeciesCountTimes(0)) == 5 and len(f.getLatticeTimes(0)) == 5


def test_hook_protocol(tmp_path):
    spec = small_spec()
    path = str(tmp_path / "hook.lm")
    spec.to_file(path, {"seed": SEED, "maxTime": repr(HOOK_STEPS * spec.tau), "writeInterval": HOOK_WI,
                        "latticeWriteInterval": HOOK_LWI, "hookInterval": HOOK_HI})
    solver = EditingOracleSolver()
    assert runSolver(path, 5, solver)
    assert solver.calls == [0, 4, 8, 12, 16]
    frames, counts, site_steps, final = expected_hooked(spec, 5)
    with SimulationFile(path, writable=False) as f:
        assert len(f.getLatticeTimes(5)) == len(frames)
        for k, fr in enumerate(frames):
            np.testing.assert_array_equal(f.getLattice(5, k), fr, err_msg=f"frame {k}")
        np.testing.assert_array_equal(f.getSpeciesCounts(5), counts)
        # the site type changed at step 8 shows in the max site counts
        X, Y, Z = spec.shape
        sites = spec.sites.copy()
        st0 = type_counts(sites, spec.nST)
        sites[9, 1, 2] = 2
        np.testing.assert_array_equal(f.getReplicateDataset(5, "MaxSiteCounts"),
                                      np.maximum(st0, type_counts(sites, spec.nST)))
    np.testing.assert_array_equal(solver.ended.reshape(-1), final)
    assert site_steps == [8, 16]


def test_hook_code_3_does_not_upload(tmp_path):
    """Return code 3 records the sites but does not copy the host lattice
    back (MpdRdmeSolver.cu:975-978): an edit returned with 3 never reaches
    the device."""

    class Discarding(OracleSolver):
        def hookSimulation(self, time, lattice):
            if time > 0:
                lattice.getParticleLatticeView()[:] = 0
                return 3
            return 0

    spec = small_spec()
    path = str(tmp_path / "three.lm")
    spec.to_file(path, {"seed": SEED, "maxTime": repr(4 * spec.tau), "writeInterval": 4,
                        "latticeWriteInterval": 100, "hookInterval": 2})
    solver = Discarding()
    assert runSolver(path, 1, solver)
    orc = oracle.Oracle(spec, seed_top=SEED, replicate=1)
    assert orc.run(4) == 0
    np.testing.assert_array_equal(solver.orc.lattice, orc.lattice)
    with SimulationFile(path, writable=False) as f:
        # counts at step 4 come from the (emptied) host copy, as the
        # reference counts the host lattice after the hook
        assert f.getSpeciesCounts(1)[-1].sum() == 0


class StoppingSolver(EditingOracleSolver):
    """The hook of step 8 asks every trajectory to stop (globalAbort)."""

    def hookSimulation(self, time, lattice):
        code = super().hookSimulation(time, lattice)
        if int(round(time / self.tau)) == 8:
            lmx.request_abort()
        return code


def test_stop_request_ends_with_the_end_of_trajectory(tmp_path, capsys):
    """A stop request (the reference's globalAbort, read before every step,
    MpdRdmeSolver.cu:561-565) ends the loop after the events of the step it
    was made at; the end of the trajectory still runs (:662-667): the final
    lattice goes to onEndTrajectory and the counts and max counts recorded so
    far are written."""
    spec = small_spec()
    path = str(tmp_path / "stop.lm")
    spec.to_file(path, {"seed": SEED, "maxTime": repr(HOOK_STEPS * spec.tau), "writeInterval": HOOK_WI,
                        "latticeWriteInterval": HOOK_LWI, "hookInterval": HOOK_HI})
    solver = StoppingSolver()
    try:
        assert runSolver(path, 5, solver)
    finally:
        lmx.clear_abort()
    assert "Global abort: terminating solver" in capsys.readouterr().err
    assert solver.calls == [0, 4, 8]
    frames, counts, site_steps, final = expected_hooked(spec, 5, stop=8)
    with SimulationFile(path, writable=False) as f:
        assert len(f.getLatticeTimes(5)) == len(frames) == 3
        for k, fr in enumerate(frames):
            np.testing.assert_array_equal(f.getLattice(5, k), fr, err_msg=f"frame {k}")
        np.testing.assert_array_equal(f.getSpeciesCounts(5), counts)
        np.testing.assert_array_equal(f.getSpeciesCountTimes(5), [t * spec.tau for t in range(0, 9, HOOK_WI)])
        assert f.getReplicateDataset(5, "MaxParticleCounts").sum() > 0
    np.testing.assert_array_equal(solver.ended.reshape(-1), final)
    assert site_steps == [8]


class CountingOverflows(OracleSolver):
    """Reports `per_window` steps with overflows in every window of 1000
    steps (the oracle runs the lattice; the statistics are scripted)."""

    per_window = 0

    def getOverflowStats(self):
        done = sum(self.steps_run)
        return done, (done // 1000) * self.per_window + min(done % 1000, self.per_window)


@pytest.mark.parametrize("per_window, warned", [(11, 2), (10, 0)])
def test_overflow_use_warning_per_1000_steps(tmp_path, per_window, warned):
    """More than 10 steps with overflows in a window of 1000 steps warns
    (overflowTimesteps / overflowListUses, MpdRdmeSolver.cu:1180-1186); the
    step batches end at the window boundaries."""
    spec = small_spec()
    path = str(tmp_path / "ovf.lm")
    spec.to_file(path, {"seed": SEED, "maxTime": repr(2500 * spec.tau), "writeInterval": 2500,
                        "latticeWriteInterval": 2500})
    solver = CountingOverflows()
    solver.per_window = per_window
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert runSolver(path, 0, solver)
    msgs = [str(x.message) for x in w if "uses of the particle overflow list" in str(x.message)]
    assert len(msgs) == warned
    assert all(m.startswith(f"{per_window} uses") for m in msgs)
    assert solver.steps_run == [1000, 1000, 500]


def test_progress_line(tmp_path, capsys):
    """perfPrintInterval (seconds; MpdRdmeSolver.cu:526-603): the progress
    line with the average wall time per step."""
    spec = small_spec()
    path = str(tmp_path / "perf.lm")
    spec.to_file(path, {"seed": SEED, "maxTime": repr(6 * spec.tau), "writeInterval": 3,
                        "latticeWriteInterval": 6, "perfPrintInterval": 0})
    assert runSolver(path, 0, OracleSolver())
    err = capsys.readouterr().err
    assert err.count("Average walltime per timestep") == 2 and "100% done" in err
