"""Protocol state after a failed rebuild.

buildModel and buildDiffusionModel replace the solver's shape, tables and
buffers before all of their checks have passed.  After such a call fails, the
solver must hold no model or lattice that a step or a readback could combine
with the device tables of the earlier build: every call that needs one
returns LMX_ERR_STATE.  No kernel is launched after the failed call; the
device is needed only for the first, successful build.
"""
import ctypes

import numpy as np
import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd import _lib, models

pytestmark = pytest.mark.gpu


def _built_solver():
    spec = models.random_model(shape=(8, 6, 5), P=8, nS=4, nST=3, R=6, density=0.3, seed=3)
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, replicate=1, seed=5)
    solver.runTimesteps(2)
    return spec, solver


def _assert_no_lattice(solver, n_species):
    lib, h = solver._lib, solver._h
    assert lib.lmx_run_steps(h, 1) == _lib.ERR_STATE
    X, Y, Z = solver.shape
    P = solver.particlesPerSite
    buf = np.zeros(X * Y * Z * P, dtype=np.uint8)
    assert lib.lmx_get_lattice(h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), buf.nbytes) == _lib.ERR_STATE
    counts = np.zeros(max(n_species, 1), dtype=np.uint32)
    assert lib.lmx_get_species_counts(h, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                      counts.size) == _lib.ERR_STATE


def test_failed_diffusion_rebuild_leaves_no_lattice(gpu):
    spec, solver = _built_solver()
    # D far above the stable limit: T = D tau / lambda^2 > 0.5 is rejected
    # after the new shape has been taken.
    bad_DF = np.full_like(spec.DF, 1.0)
    X, Y, Z = 16, 12, 10
    with pytest.raises(_lib.InvalidArgException):
        solver.buildDiffusionModel(spec.nST, bad_DF, spec.RL, X, Y, Z, spec.P, 1, spec.spacing,
                                   np.zeros((X, Y, Z, spec.P), np.uint8), np.zeros((X, Y, Z), np.uint8))
    _assert_no_lattice(solver, spec.nS)
    # a successful rebuild restores the protocol
    spec.load_into(solver, replicate=1, seed=5)
    solver.runTimesteps(1)
    solver.close()


def test_failed_model_rebuild_leaves_no_model(gpu):
    spec, solver = _built_solver()
    nS, R = spec.nS + 7, 1
    # a first-order reaction with two reactants fails the order check after
    # nS and R have been taken
    D = np.zeros((nS, R), np.uint32)
    D[0, 0] = D[1, 0] = 1
    with pytest.raises(_lib.InvalidArgException):
        solver.buildModel(nS, R, np.zeros(nS, np.uint32), [1], [1.0], np.zeros((nS, R), np.int32), D)
    _assert_no_lattice(solver, nS)
    # no model either: a rate change and a diffusion build are refused
    assert solver._lib.lmx_set_reaction_rate(solver._h, 0, 1.0) != _lib.OK
    with pytest.raises(_lib.LmxError) as err:
        solver.buildDiffusionModel(spec.nST, spec.DF, spec.RL, *spec.shape, spec.P, 1, spec.spacing,
                                   spec.lattice, spec.sites)
    assert err.value.code == _lib.ERR_STATE
    solver.close()
