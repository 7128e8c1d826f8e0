"""The C ABI from a C++ caller with no Python in its process
(tests/capi/caller.cpp, built by build()): a model written to a file, run
through lmx_create / lmx_initialize / lmx_build_model /
lmx_build_diffusion_model / lmx_run_steps, and its native lattice and species
counts compared bit for bit with the oracle."""
import os
import subprocess

import numpy as np
import pytest

from lattice_microbes_amd import models
import oracle

pytestmark = pytest.mark.gpu

CALLER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "capi", "_build", "capi_caller")


def write_model(path, spec, replicate, seed_top, steps):
    X, Y, Z = spec.shape
    R = spec.R
    with open(path, "wb") as f:
        boundary = 1 if spec.boundary == "reflecting" else 0
        np.array([spec.nS, R, spec.nST, X, Y, Z, spec.P, 1, replicate, seed_top, steps, boundary],
                 np.uint32).tofile(f)
        np.array([spec.tau, spec.spacing], np.float64).tofile(f)
        np.asarray(spec.initial_counts(), np.uint32).tofile(f)
        np.asarray(spec.types, np.uint32).tofile(f)
        np.asarray(spec.K, np.float64).reshape(-1).tofile(f)
        np.asarray(spec.S, np.int32).reshape(-1).tofile(f)
        np.asarray(spec.D, np.uint32).reshape(-1).tofile(f)
        np.asarray(spec.DF, np.float64).reshape(-1).tofile(f)
        np.asarray(spec.RL, np.uint32).reshape(-1).tofile(f)
        np.asarray(spec.lattice, np.uint8).reshape(-1).tofile(f)
        np.asarray(spec.sites, np.uint8).reshape(-1).tofile(f)


@pytest.mark.parametrize("case", [
    dict(shape=(16, 12, 10), P=8, nS=4, nST=3, R=6, density=0.3, boundary="periodic"),
    dict(shape=(32, 8, 6), P=16, nS=5, nST=2, R=5, density=0.2, boundary="reflecting"),
], ids=["P8", "P16-reflecting"])
def test_cpp_caller_matches_oracle(gpu, tmp_path, case):
    if not os.path.exists(CALLER):
        pytest.fail("the C++ caller is not built (build() compiles tests/capi/caller.cpp)")
    spec = models.random_model(seed=19, **case)
    steps, replicate, seed_top = 7, 2, 31
    model, out = str(tmp_path / "model.bin"), str(tmp_path / "out.bin")
    write_model(model, spec, replicate, seed_top, steps)
    r = subprocess.run([CALLER, model, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    X, Y, Z = spec.shape
    words = (spec.P // 4) * X * Y * Z
    data = np.fromfile(out, np.uint32)
    native, counts = data[:words].reshape(spec.P // 4, Z, Y, X), data[words:]
    orc = oracle.Oracle(spec, seed_top=seed_top, replicate=replicate)
    assert orc.run(steps) == 0
    np.testing.assert_array_equal(native, orc.native())
    np.testing.assert_array_equal(counts, orc.species_counts())
