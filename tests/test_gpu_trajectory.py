"""Trajectories into Lattice Microbes HDF5 files, and the hook protocol, on
the GPU, checked bit for bit against the oracle stepped to the same steps.

runSimulation reads the model, parameters and initial lattice from the file
and writes species counts, lattice frames, site frames and the max-count
datasets back into it (MpdRdmeSolver.cu:485-668, 824-985); a solver subclass
overriding hookSimulation / onBeginTrajectory changes the lattice, the site
types and a rate constant mid-run, and the oracle applies the same changes at
the same steps.
"""
import numpy as np
import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd.runner import runSimulation, runSolver
from lattice_microbes_amd.simfile import SimulationFile
import oracle
from trajectory_cases import (HOOK_HI, HOOK_LWI, HOOK_STEPS, HOOK_WI, PLAN, SEED, EditingHooks, expected_hooked,
                              expected_plain, small_spec, type_counts)

pytestmark = pytest.mark.gpu


def test_run_simulation_writes_trajectory(tmp_path):
    spec = small_spec()
    path = str(tmp_path / "traj.lm")
    steps, wi, lwi = 23, 5, 7
    spec.to_file(path, {"seed": SEED, "maxTime": repr(steps * spec.tau), "writeInterval": wi,
                        "latticeWriteInterval": lwi})
    assert runSimulation(path, 3)
    counts, frames, maxp = expected_plain(spec, 3, steps, wi, lwi)
    with SimulationFile(path, writable=False) as f:
        assert f.replicateExists(3) and not f.replicateExists(0)
        np.testing.assert_array_equal(f.getSpeciesCountTimes(3), [t * spec.tau for t in sorted(counts)])
        np.testing.assert_array_equal(f.getSpeciesCounts(3), np.array([counts[t] for t in sorted(counts)]))
        np.testing.assert_array_equal(f.getLatticeTimes(3), [t * spec.tau for t in sorted(frames)])
        for k, t in enumerate(sorted(frames)):
            np.testing.assert_array_equal(f.getLattice(3, k), frames[t], err_msg=f"frame at step {t}")
        np.testing.assert_array_equal(f.getReplicateDataset(3, "MaxParticleCounts"), maxp)
        np.testing.assert_array_equal(f.getReplicateDataset(3, "MaxSiteCounts"), type_counts(spec.sites, spec.nST))


def test_second_replicate_and_bad_file(tmp_path, capsys):
    spec = small_spec()
    path = str(tmp_path / "two.lm")
    spec.to_file(path, {"seed": SEED, "maxTime": repr(6 * spec.tau), "writeInterval": 3,
                        "latticeWriteInterval": 6})
    assert runSimulation(path, 0) and runSimulation(path, 1)
    with SimulationFile(path, writable=False) as f:
        for r in (0, 1):
            orc = oracle.Oracle(spec, seed_top=SEED, replicate=r)
            assert orc.run(6) == 0
            np.testing.assert_array_equal(f.getLattice(r, 1), orc.row_major())
            np.testing.assert_array_equal(f.getSpeciesCounts(r)[-1], orc.species_counts())
        assert not np.array_equal(f.getLattice(0, 1), f.getLattice(1, 1))
    # an unsupported parameter is reported the way the reference does, not raised
    bad = str(tmp_path / "bad.lm")
    spec.to_file(bad, {"seed": SEED, "maxTime": repr(2 * spec.tau), "timestep": "0"})
    assert not runSimulation(bad, 0)
    assert "Exception during execution of replicate" in capsys.readouterr().err


class EditingSolver(EditingHooks, lmx.MpdRdmeSolver):
    pass


def test_hook_protocol_matches_oracle(tmp_path):
    spec = small_spec()
    path = str(tmp_path / "hook.lm")
    spec.to_file(path, {"seed": SEED, "maxTime": repr(HOOK_STEPS * spec.tau), "writeInterval": HOOK_WI,
                        "latticeWriteInterval": HOOK_LWI, "hookInterval": HOOK_HI})
    solver = EditingSolver(0)
    try:
        assert runSolver(path, 5, solver)
        assert solver.calls == [0, 4, 8, 12, 16]
        frames, counts, site_steps, final = expected_hooked(spec, 5)
        with SimulationFile(path, writable=False) as f:
            assert len(f.getLatticeTimes(5)) == len(frames)
            for k, fr in enumerate(frames):
                np.testing.assert_array_equal(f.getLattice(5, k), fr, err_msg=f"frame {k}")
            np.testing.assert_array_equal(f.getSpeciesCounts(5), counts)
        np.testing.assert_array_equal(solver.ended.reshape(-1), final)
        assert site_steps == [8, 16]
    finally:
        solver.close()


def test_lm_module_runs_a_jlm_style_file(tmp_path):
    """A file built through the lm compatibility module, run with lm.runSolver
    and a solver composed the way jLM's makeSolver does it."""
    from lattice_microbes_amd import compat
    from test_lm_compat import build_with_lm

    compat.install()
    import lm

    spec = small_spec()
    path = str(tmp_path / "jlm.h5")
    build_with_lm(spec, path)
    with lm.SimulationFile(path) as f:
        f.setParameter("seed", str(SEED))

    class Watch:
        def __init__(self):
            super().__init__()
            self.seen = []

        def hookSimulation(self, t, lattice):
            self.seen.append(round(t / self.tau))
            return 0

    solver = type("Watch_MpdRdmeSolver", (Watch, lm.MpdRdmeSolver), {})()
    try:
        lm.runSolver(path, 2, solver, [0], 0)
        assert solver.seen == [0]            # hookInterval "" disables all but the initial check
    finally:
        solver.close()
    counts, frames, _ = expected_plain(spec, 2, 20, 5, 10)
    with SimulationFile(path, writable=False) as f:
        for k, t in enumerate(sorted(frames)):
            np.testing.assert_array_equal(f.getLattice(2, k), frames[t])
        np.testing.assert_array_equal(f.getSpeciesCounts(2), np.array([counts[t] for t in sorted(counts)]))


class DeviceEditingSolver(EditingHooks, lmx.MpdRdmeSolver):
    """The same plan as a device-side hook: the edits are torch operations on
    the solver's device buffers (no host copies)."""
    deviceHooks = True

    def hookSimulation(self, time, lattice):
        step = int(round(time / self.tau))
        self.calls.append(step)
        edit, code, rate = PLAN[step]
        if edit is not None:
            edit.apply(lattice.getParticleLatticeView(), lattice.getSiteLatticeView(), self.species_offset)
        if rate is not None:
            self.setReactionRate(0, rate)
        assert lattice.getParticleLatticeView().is_cuda and lattice.getXSize() == 16
        return code

    def onEndTrajectory(self, lattice):
        self.ended = lattice.getParticleLatticeView().cpu().numpy()
        return 0


def need_torch_first():
    import torch

    if not torch.cuda.is_initialized():
        pytest.skip("torch must start before the library in this process (run with -m gpu)")


def test_device_hooks_match_oracle(tmp_path):
    need_torch_first()
    spec = small_spec()
    path = str(tmp_path / "dhook.lm")
    spec.to_file(path, {"seed": SEED, "maxTime": repr(HOOK_STEPS * spec.tau), "writeInterval": HOOK_WI,
                        "latticeWriteInterval": HOOK_LWI, "hookInterval": HOOK_HI})
    solver = DeviceEditingSolver(0)
    try:
        assert runSolver(path, 5, solver)
        assert solver.calls == [0, 4, 8, 12, 16]
        frames, counts, site_steps, final = expected_hooked(spec, 5)
        with SimulationFile(path, writable=False) as f:
            assert len(f.getLatticeTimes(5)) == len(frames)
            for k, fr in enumerate(frames):
                np.testing.assert_array_equal(f.getLattice(5, k), fr, err_msg=f"frame {k}")
            np.testing.assert_array_equal(f.getSpeciesCounts(5), counts)
        np.testing.assert_array_equal(solver.ended.reshape(-1), final)
    finally:
        solver.close()


def test_device_hook_invalid_species_is_reported(tmp_path, capsys):
    class Bad(lmx.MpdRdmeSolver):
        deviceHooks = True

        def hookSimulation(self, time, lattice):
            if time > 0:
                lattice.getParticleLatticeView()[0, 1, 1, 1, 0] = 200   # nS is 4
                return 1
            return 0

    need_torch_first()
    spec = small_spec()
    path = str(tmp_path / "bad.lm")
    spec.to_file(path, {"seed": SEED, "maxTime": repr(8 * spec.tau), "writeInterval": 4,
                        "latticeWriteInterval": 8, "hookInterval": 4})
    solver = Bad(0)
    try:
        assert not runSolver(path, 0, solver)
    finally:
        solver.close()
    assert "invalid species or site type" in capsys.readouterr().err


# ---------------------------------------------------------------------------
# The uint32 lattice (IntMpdRdmeSolver): the same trajectories with every
# species id raised by OFF (past a byte), against the byte lattice's expected
# records raised the same way.
# ---------------------------------------------------------------------------
OFF = 300


def raised(a, off=OFF):
    a = np.asarray(a).astype(np.uint32)
    a[a > 0] += off
    return a


def int_spec():
    from test_gpu_int_lattice import shifted

    return shifted(small_spec(), OFF)


def test_int_solver_run_simulation(tmp_path):
    spec, ispec = small_spec(), int_spec()
    path = str(tmp_path / "int.lm")
    steps, wi, lwi = 23, 5, 7
    ispec.to_file(path, {"seed": SEED, "maxTime": repr(steps * spec.tau), "writeInterval": wi,
                         "latticeWriteInterval": lwi})
    assert runSimulation(path, 3, solverClass=lmx.INT_SOLVER_NAME)
    counts, frames, maxp = expected_plain(spec, 3, steps, wi, lwi)
    with SimulationFile(path, writable=False) as f:
        c = f.getSpeciesCounts(3)
        assert c.shape[1] == spec.nS + OFF and not c[:, :OFF].any()
        np.testing.assert_array_equal(c[:, OFF:], np.array([counts[t] for t in sorted(counts)]))
        for k, t in enumerate(sorted(frames)):
            fr = f.getLattice(3, k)
            assert fr.dtype == np.uint32
            np.testing.assert_array_equal(fr, raised(frames[t]), err_msg=f"frame at step {t}")
        mp = f.getReplicateDataset(3, "MaxParticleCounts")
        np.testing.assert_array_equal(mp[OFF + 1:], maxp[1:])
        assert not mp[:OFF + 1].any()


class IntEditingSolver(EditingHooks, lmx.IntMpdRdmeSolver):
    species_offset = OFF


class IntDeviceEditingSolver(DeviceEditingSolver, lmx.IntMpdRdmeSolver):
    species_offset = OFF


@pytest.mark.parametrize("device", [False, True], ids=["host-hooks", "device-hooks"])
def test_int_solver_hooks(tmp_path, device):
    if device:
        need_torch_first()
    spec, ispec = small_spec(), int_spec()
    path = str(tmp_path / "ihook.lm")
    ispec.to_file(path, {"seed": SEED, "maxTime": repr(HOOK_STEPS * spec.tau), "writeInterval": HOOK_WI,
                         "latticeWriteInterval": HOOK_LWI, "hookInterval": HOOK_HI})
    solver = IntDeviceEditingSolver(0) if device else IntEditingSolver(0)
    try:
        assert runSolver(path, 5, solver)
        assert solver.calls == [0, 4, 8, 12, 16]
        frames, counts, site_steps, final = expected_hooked(spec, 5)
        with SimulationFile(path, writable=False) as f:
            assert len(f.getLatticeTimes(5)) == len(frames)
            for k, fr in enumerate(frames):
                np.testing.assert_array_equal(f.getLattice(5, k), raised(fr), err_msg=f"frame {k}")
            np.testing.assert_array_equal(f.getSpeciesCounts(5)[:, OFF:], counts)
        # final: the byte lattice's native bytes in slot order per word plane
        X, Y, Z = spec.shape
        fb = final.reshape(spec.P // 4, Z, Y, X, 4).transpose(0, 4, 1, 2, 3).reshape(spec.P, Z, Y, X)
        np.testing.assert_array_equal(np.asarray(solver.ended).reshape(spec.P, Z, Y, X).astype(np.uint32),
                                      raised(fb))
    finally:
        solver.close()
