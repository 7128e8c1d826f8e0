"""lmx_run_steps_static, the stepping entry point that allocates nothing (its
reaction list lives in a static device array of the library), called
through the C ABI and compared bit for bit with lmx_run_steps on the same
inputs: two calls in a row, the second on a smaller lattice, so that the
shared array holds the first call's entries when the second starts."""
import ctypes

import numpy as np
import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd import _lib, models

pytestmark = pytest.mark.gpu

MAX_SITES = 1024 * 256 * 8  # the static list's limit, as include/lmx_mpd.h states


def crowded(shape, seed):
    """The whole-cell layout (10 site types, 30 species, 90 first/second-order
    reactions, no zero-order ones) with a fifth of the fluid sites occupied."""
    X, Y, Z = shape
    return models.whole_cell(shape=shape, nS=30, R=90, n_particles=X * Y * Z // 5, seed=seed)


def pair(spec, replicate, seed):
    a, b = lmx.MpdRdmeSolver(0), lmx.MpdRdmeSolver(0)
    spec.load_into(a, replicate=replicate, seed=seed)
    spec.load_into(b, replicate=replicate, seed=seed)
    return a, b


def static_steps(solver, n):
    _lib.check(_lib.load().lmx_run_steps_static(solver._h, ctypes.c_uint32(n)))
    solver.synchronize()


@pytest.mark.parametrize("steps", [(6, 5)])
def test_static_steps_equal_run_steps(gpu, steps):
    lib = _lib.load()
    # first call: 128x128x96 (1.57M sites, many occupied), second: 48^3
    for spec, n, (rep, seed) in ((crowded((128, 128, 96), 3), steps[0], (1, 7)),
                                 (models.bimolecular(shape=(48, 48, 48), seed=5), steps[1], (2, 9))):
        assert spec.N <= MAX_SITES
        a, b = pair(spec, rep, seed)
        static_steps(a, n)
        _lib.check(lib.lmx_run_steps(b._h, ctypes.c_uint32(n)))
        b.synchronize()
        na, nb = a.getLatticeNative(), b.getLatticeNative()
        assert np.array_equal(na, nb), f"{spec.shape}: {(na != nb).sum()} words differ"
        np.testing.assert_array_equal(a.getSpeciesCounts(), b.getSpeciesCounts())
        assert a.getOverflowStats() == b.getOverflowStats()
        assert a.timestep == b.timestep == n
        a.close()
        b.close()


def test_static_steps_limits(gpu):
    lib = _lib.load()
    # one site over the limit: rejected before any launch
    big = models.bimolecular(shape=(MAX_SITES // 4096 + 1, 64, 64), nA=10, nB=10, seed=1)
    s = lmx.MpdRdmeSolver(0)
    big.load_into(s, replicate=0, seed=1)
    assert lib.lmx_run_steps_static(s._h, ctypes.c_uint32(1)) == _lib.ERR_ARG
    assert s.timestep == 0
    s.close()
    # a rate change not yet uploaded: rejected; lmx_run_steps uploads it
    small = models.bimolecular(shape=(32, 32, 32), seed=2)
    s = lmx.MpdRdmeSolver(0)
    small.load_into(s, replicate=0, seed=2)
    s.setReactionRate(0, 1.0)
    assert lib.lmx_run_steps_static(s._h, ctypes.c_uint32(1)) == _lib.ERR_STATE
    _lib.check(lib.lmx_run_steps(s._h, ctypes.c_uint32(1)))
    static_steps(s, 2)
    assert s.timestep == 3
    s.close()
