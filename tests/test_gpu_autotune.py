"""The launch-shape autotuner (lmx_autotune): it times every y/z kernel and
window option from the current state, so afterwards the lattice, step
counter, overflow statistics and carried overflow entries must be exactly
as before, and the following steps must still equal the oracle."""
import numpy as np
import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd import models
import oracle

pytestmark = pytest.mark.gpu


class RelaxedIntSolver(lmx.IntMpdRdmeSolver):
    _parses_overflow_handler = True


CASES = [
    # shape, P, density, relaxed, uint32
    ((64, 16, 12), 8, 0.45, True, False),     # carried entries across the tuning
    ((256, 8, 8), 8, 0.3, False, False),       # whole-chunk waves keep the plane flags
    ((32, 8, 8), 4, 0.3, False, False),
    ((16, 8, 6), 8, 0.4, True, True),          # uint32 lattice
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-P{c[1]}{'-relaxed' if c[3] else ''}{'-uint32' if c[4] else ''}"
                                             for c in CASES])
def test_autotune_keeps_state_and_results(gpu, case):
    shape, P, density, relaxed, words = case
    spec = models.random_model(shape=shape, P=P, nS=4, nST=2, R=6, density=density, seed=61, qmax=0.5)
    if words:
        spec.lattice = spec.lattice.astype(np.uint32)
    solver = RelaxedIntSolver(0) if words else lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=29, extra={"rdme.mpd.overflowhandler": "relaxed"} if relaxed else None)
    orc = oracle.Oracle(spec, seed_top=29, overflow="relaxed" if relaxed else "classic")
    solver.runTimesteps(3)
    assert orc.run(3) == 0
    before = solver.getLatticeNative()
    stats, unresolved, t = solver.getOverflowStats(), solver.getUnresolvedOverflows(), solver.timestep
    choice = solver.autotune(steps=3)
    assert choice["ms_per_step"] > 0 and choice["yz_kernel"] in (0, 1, 2, 3)
    np.testing.assert_array_equal(solver.getLatticeNative(), before)
    assert solver.getOverflowStats() == stats
    assert solver.getUnresolvedOverflows() == unresolved
    assert solver.timestep == t
    solver.runTimesteps(5)
    assert orc.run(5) == 0
    np.testing.assert_array_equal(solver.getLatticeNative(), orc.native())
    assert solver.getUnresolvedOverflows() == orc.unresolved
    if relaxed:
        assert orc.stats.overflow_entries > 0
    solver.close()


def test_autotune_refuses_ensemble_members(gpu):
    spec = models.random_model(shape=(16, 8, 8), P=8, nS=3, nST=2, R=3, density=0.2, seed=3)
    s = lmx.MpdRdmeSolver(0)
    spec.load_into(s, seed=4)
    ens = lmx.Ensemble(0)
    ens.add(s)
    with pytest.raises(Exception, match="ensemble member"):
        s.autotune(steps=2)
    ens.close()
