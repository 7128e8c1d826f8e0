"""GPU parity of the relaxed overflow handler (rdme.mpd.overflowhandler =
"relaxed", correct_overflows, MpdRdmeSolver.cu:1677-1761): dense lattices
whose overflows are placed in their own sites or carried from step to step,
on every kernel path, the uint32 lattice and the batched ensemble, compared
word for word with the oracle, together with the carried count."""
import numpy as np
import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd import models
import oracle

pytestmark = pytest.mark.gpu

RELAXED = {"rdme.mpd.overflowhandler": "relaxed"}


class RelaxedIntSolver(lmx.IntMpdRdmeSolver):
    # IntMpdRdmeSolver ignores the handler as the reference's does; the C ABI
    # offers the relaxed handler on the uint32 lattice too.
    _parses_overflow_handler = True


def _run(solver, orc, steps, chunks):
    for _ in range(chunks):
        solver.runTimesteps(steps // chunks)
        assert orc.run(steps // chunks) == 0
        np.testing.assert_array_equal(solver.getLatticeNative(), orc.native())
        assert solver.getUnresolvedOverflows() == orc.unresolved
    e, s = solver.getOverflowStats()
    assert (e, s) == (orc.stats.overflow_entries, orc.stats.overflow_steps)


CASES = [
    # shape, P, nS, nST, R, density, path option
    ((12, 10, 8), 8, 3, 1, 0, 0.45, "default"),
    ((12, 10, 8), 4, 3, 2, 4, 0.7, "default"),
    ((64, 8, 8), 8, 4, 2, 6, 0.5, "default"),
    ((64, 8, 8), 16, 4, 2, 6, 0.35, "default"),
    ((64, 8, 8), 8, 4, 2, 6, 0.5, "march"),
    ((64, 8, 8), 8, 4, 2, 6, 0.5, "fused"),
    ((33, 7, 5), 8, 4, 2, 6, 0.5, "generic"),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-P{c[1]}-R{c[4]}-{c[6]}" for c in CASES])
def test_relaxed_matches_oracle(gpu, case):
    shape, P, nS, nST, R, density, path = case
    spec = models.random_model(shape=shape, P=P, nS=nS, nST=nST, R=R, density=density, seed=21, qmax=0.5)
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=23, extra=RELAXED)
    assert solver.overflowHandler == "relaxed"
    if path == "march":
        solver.setLaunchOption(lmx._lib.OPT_YZ_KERNEL, 1)
    elif path == "fused":
        solver.setPath(lmx._lib.PATH_FUSED)
    elif path == "generic":
        solver.setLaunchOption(lmx._lib.OPT_FORCE_GENERIC, 1)
    orc = oracle.Oracle(spec, seed_top=23, overflow="relaxed")
    _run(solver, orc, steps=12, chunks=4)
    assert orc.stats.overflow_entries > 0
    solver.close()


def test_relaxed_carries_entries_across_steps(gpu):
    spec = models.random_model(shape=(12, 10, 8), P=4, nS=3, nST=1, R=0, density=0.6, seed=11, qmax=0.5)
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=13, extra=RELAXED)
    orc = oracle.Oracle(spec, seed_top=13, overflow="relaxed")
    _run(solver, orc, steps=6, chunks=6)
    assert orc.unresolved > 100
    # the lattice and the carried particles together keep every species' count
    n = solver.getUnresolvedOverflows()
    carried = np.bincount(orc.carry_particle[:n], minlength=spec.nS + 1)[1:]
    np.testing.assert_array_equal(solver.getSpeciesCounts() + carried, spec.initial_counts())
    solver.close()


def test_relaxed_uint32_lattice(gpu):
    spec = models.random_model(shape=(16, 8, 6), P=8, nS=4, nST=2, R=6, density=0.5, seed=31, qmax=0.5)
    spec.lattice = spec.lattice.astype(np.uint32)
    solver = RelaxedIntSolver(0)
    spec.load_into(solver, seed=7, extra=RELAXED)
    orc = oracle.Oracle(spec, seed_top=7, overflow="relaxed")
    _run(solver, orc, steps=10, chunks=5)
    assert orc.stats.overflow_entries > 0
    solver.close()


def test_relaxed_ensemble_members(gpu):
    ens = lmx.Ensemble(0, graph=True)
    pairs = []
    for k, shape in enumerate([(12, 10, 8), (64, 8, 8), (12, 5, 7)]):
        spec = models.random_model(shape=shape, P=8, nS=4, nST=2, R=5, density=0.5, seed=50 + k, qmax=0.5)
        s = lmx.MpdRdmeSolver(0)
        extra = RELAXED if k != 1 else {}         # a classic member in the same batch
        spec.load_into(s, replicate=k, seed=3 + k, extra=extra)
        orc = oracle.Oracle(spec, seed_top=3 + k, replicate=k, overflow="relaxed" if k != 1 else "classic")
        ens.add(s)
        pairs.append((s, orc))
    for chunk in (2, 3):
        ens.run(chunk)
        for s, orc in pairs:
            assert orc.run(chunk) == 0
            np.testing.assert_array_equal(s.getLatticeNative(), orc.native())
            assert s.getUnresolvedOverflows() == orc.unresolved
    ens.close()


def test_handler_parameter(gpu):
    spec = models.random_model(shape=(8, 6, 5), P=8, nS=3, nST=2, R=3, density=0.3, seed=2)
    s = lmx.MpdRdmeSolver(0)
    with pytest.warns(UserWarning, match="Unknown overflow handler requested: 'other'"):
        spec.load_into(s, seed=5, extra={"rdme.mpd.overflowhandler": "other"})
    assert s.overflowHandler == "classic"
    s.close()
    spec.lattice = spec.lattice.astype(np.uint32)
    i = lmx.IntMpdRdmeSolver(0)
    spec.load_into(i, seed=5, extra=RELAXED)
    assert i.overflowHandler == "classic"      # IntMpdRdmeSolver has no relaxed handler
    assert i.getUnresolvedOverflows() == 0
    i.close()
