"""The C ABI from a C++ caller with no Python in its process
(tests/capi/caller.cpp over include/lmx_mpd_solver.hpp, built by build()): a
model, the trajectory parameters and a hook plan written to a file, run
through lmx::MpdRdmeSolver::generateTrajectory, and every record it produced
(species counts and their times, lattice frames, site frames, max counts, the
lattice onEndTrajectory saw, the hook calls) compared bit for bit with the
oracle stepped to the same steps with the same edits
(MpdRdmeSolver.cu:485-668, 824-985)."""
import math
import os
import subprocess

import numpy as np
import pytest

from lattice_microbes_amd import models
import oracle
from trajectory_cases import (BEGIN_EDIT, HOOK_HI, HOOK_LWI, HOOK_STEPS, HOOK_WI, PLAN, SEED, expected_hooked,
                              expected_plain, small_spec, type_counts)

pytestmark = pytest.mark.gpu

CALLER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "capi", "_build", "capi_caller")


def write_input(path, spec, replicate, seed_top, max_steps, wi, lwi, hi=None, begin=None, plan=None,
                abort_at=None):
    X, Y, Z = spec.shape
    R = spec.R
    with open(path, "wb") as f:
        boundary = 1 if spec.boundary == "reflecting" else 0
        np.array([spec.nS, R, spec.nST, X, Y, Z, spec.P, 1, replicate, seed_top, boundary, 0,
                  0 if hi is None else 1], np.uint32).tofile(f)
        np.array([spec.tau, spec.spacing, max_steps * spec.tau, wi, lwi, hi or 0], np.float64).tofile(f)
        np.asarray(spec.initial_counts(), np.uint32).tofile(f)
        np.asarray(spec.types, np.uint32).tofile(f)
        np.asarray(spec.K, np.float64).reshape(-1).tofile(f)
        np.asarray(spec.S, np.int32).reshape(-1).tofile(f)
        np.asarray(spec.D, np.uint32).reshape(-1).tofile(f)
        np.asarray(spec.DF, np.float64).reshape(-1).tofile(f)
        np.asarray(spec.RL, np.uint32).reshape(-1).tofile(f)
        np.asarray(spec.lattice, np.uint8).reshape(-1).tofile(f)
        np.asarray(spec.sites, np.uint8).reshape(-1).tofile(f)
        b = begin
        np.array([b.x, b.y, b.z, b.s, -1] if b else [-1, 0, 0, 0, -1], np.int32).tofile(f)
        plan = plan or {}
        np.array([len(plan)], np.uint32).tofile(f)
        for step, (edit, code, rate) in sorted(plan.items()):
            np.array([step], np.uint32).tofile(f)
            e = [edit.x, edit.y, edit.z, edit.s, -1 if edit.st is None else edit.st] if edit else [-1, 0, 0, 0, -1]
            np.array(e + [code, 1 if step == abort_at else 0], np.int32).tofile(f)
            np.array([math.nan if rate is None else rate], np.float64).tofile(f)


def read_records(path, nS):
    raw = open(path, "rb").read()
    out = {"counts_t": [], "counts": [], "frames": [], "sites": [], "hooks": []}
    o = 0
    while o < len(raw):
        kind, step = np.frombuffer(raw, np.uint32, 2, o)
        time = float(np.frombuffer(raw, np.float64, 1, o + 8)[0])
        n = int(np.frombuffer(raw, np.uint64, 1, o + 16)[0])
        payload = raw[o + 24:o + 24 + n]
        o += 24 + n
        if kind == 1:
            out["counts_t"] += list(np.frombuffer(payload, np.float64, int(step)))
            out["counts"] += list(np.frombuffer(payload, np.uint32, int(step) * nS, 8 * int(step)).reshape(-1, nS))
        elif kind == 2:
            out["frames"].append((int(step), time, np.frombuffer(payload, np.uint8)))
        elif kind == 3:
            out["sites"].append((int(step), time, np.frombuffer(payload, np.uint8)))
        elif kind == 4:
            out["max_particles"] = np.frombuffer(payload, np.uint32)
        elif kind == 5:
            out["max_sites"] = np.frombuffer(payload, np.uint32)
        elif kind == 6:
            out["end"] = np.frombuffer(payload, np.uint8)
        elif kind == 7:
            out["hooks"].append(int(step))
    out["counts"] = np.array(out["counts"])
    return out


def run_caller(tmp_path, *args, **kw):
    if not os.path.exists(CALLER):
        pytest.fail("the C++ caller is not built (build() compiles tests/capi/caller.cpp)")
    model, out = str(tmp_path / "model.bin"), str(tmp_path / "out.bin")
    write_input(model, *args, **kw)
    r = subprocess.run([CALLER, model, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return read_records(out, args[0].nS), r.stderr


@pytest.mark.parametrize("case", [
    dict(shape=(16, 12, 10), P=8, nS=4, nST=3, R=6, density=0.3, boundary="periodic"),
    dict(shape=(32, 8, 6), P=16, nS=5, nST=2, R=5, density=0.2, boundary="reflecting"),
], ids=["P8", "P16-reflecting"])
def test_cpp_trajectory_matches_oracle(gpu, tmp_path, case):
    """Counts every 3 steps, frames every 5, 23 steps (maxTime 23 tau): the
    cadence counts intervals in steps and stops at the first step whose time
    reaches maxTime."""
    spec = models.random_model(seed=19, **case)
    steps, wi, lwi = 23, 3, 5
    rec, _ = run_caller(tmp_path, spec, 2, SEED, steps, wi, lwi)
    counts, frames, maxp = expected_plain(spec, 2, steps, wi, lwi)
    # without hookInterval only the initial hook check runs (MpdRdmeSolver.cu:556)
    assert rec["hooks"] == [0]
    np.testing.assert_array_equal(rec["counts_t"], [t * spec.tau for t in sorted(counts)])
    np.testing.assert_array_equal(rec["counts"], np.array([counts[t] for t in sorted(counts)]))
    assert [f[0] for f in rec["frames"]] == sorted(frames)
    for step, time, fr in rec["frames"]:
        assert time == step * spec.tau
        np.testing.assert_array_equal(fr, frames[step].reshape(-1), err_msg=f"frame at step {step}")
    np.testing.assert_array_equal(rec["max_particles"], maxp)
    np.testing.assert_array_equal(rec["max_sites"], type_counts(spec.sites, spec.nST))
    orc = oracle.Oracle(spec, seed_top=SEED, replicate=2)
    assert orc.run(steps) == 0
    np.testing.assert_array_equal(rec["end"], orc.lattice.view(np.uint8))


@pytest.mark.parametrize("abort_at", [None, 8], ids=["full", "stop-request-at-8"])
def test_cpp_hook_protocol_matches_oracle(gpu, tmp_path, abort_at):
    """onBeginTrajectory's edit, hooks every 4 steps with return codes 1, 1 (and
    a rate change), 2 (and a site-type change), 0 and 3, counts every 2 steps
    and frames every 4.  With a stop request made by the hook at step 8, the
    events of step 8 are recorded and the trajectory ends there: the lattice is
    copied back, onEndTrajectory sees it, the buffered counts and the max
    counts are written (MpdRdmeSolver.cu:561-565, 662-667)."""
    spec = small_spec()
    stop = HOOK_STEPS if abort_at is None else abort_at
    rec, err = run_caller(tmp_path, spec, 5, SEED, HOOK_STEPS, HOOK_WI, HOOK_LWI, HOOK_HI, BEGIN_EDIT, PLAN,
                          abort_at)
    frames, counts, site_steps, final = expected_hooked(spec, 5, stop)
    assert rec["hooks"] == list(range(0, stop + 1, HOOK_HI))
    assert [f[0] for f in rec["frames"]] == list(range(0, stop + 1, HOOK_LWI))
    for (step, _, fr), ref in zip(rec["frames"], frames):
        np.testing.assert_array_equal(fr, ref.reshape(-1), err_msg=f"frame at step {step}")
    np.testing.assert_array_equal(rec["counts_t"], [t * spec.tau for t in range(0, stop + 1, HOOK_WI)])
    np.testing.assert_array_equal(rec["counts"], counts)
    assert [s[0] for s in rec["sites"]] == site_steps
    np.testing.assert_array_equal(rec["end"], final)
    if abort_at is not None:
        assert "Global abort" in err
