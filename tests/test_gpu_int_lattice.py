"""The uint32 lattice (IntMpdRdmeSolver, one particle per 32-bit word).

Every case is compared word for word with the oracle's uint32 mode (pinned
against its byte mode in test_oracle_words.py).  The reference's word kernels
(word_diffusion_1d_dev.cu, word_reaction_dev.cu) apply the byte kernels' rules
to 32-bit particles with the same RNG keys (slot index w under the same
MPD_PARTICLE_COUNT_BITS), so also:

* for a model that fits a byte, the uint32 lattice's trajectory equals the
  byte lattice's particle for particle, step by step, including overflow
  placement (the byte path is itself bit-exact against the oracle);
* for more than 255 species, renumbering species by a constant offset (the
  extra low-numbered species absent and inert) changes nothing but the ids:
  every table entry, sum order and product order is preserved, so the run
  with ids + offset equals the byte run of the compact model.
"""
import numpy as np
import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd import InvalidArgException, models
import oracle

pytestmark = pytest.mark.gpu


def shifted(spec, off):
    """The same model with species s renumbered s + off (species 1..off
    absent, immobile and in no reaction)."""
    nS = spec.nS + off
    S = np.zeros((nS, spec.R), np.int32)
    S[off:] = spec.S
    D = np.zeros((nS, spec.R), np.uint32)
    D[off:] = spec.D
    DF = np.zeros((spec.nST, spec.nST, nS))
    DF[..., off:] = spec.DF
    lat = spec.lattice.astype(np.uint32)
    lat[lat > 0] += off
    return models.ModelSpec(name=spec.name + f"+{off}", shape=spec.shape, spacing=spec.spacing, tau=spec.tau,
                            P=spec.P, nS=nS, nST=spec.nST, types=spec.types, K=spec.K, S=S, D=D, DF=DF,
                            RL=spec.RL, lattice=lat, sites=spec.sites, boundary=spec.boundary)


def run_both(spec, int_spec, off, steps, chunks, seed):
    """The uint32 lattice against the oracle's uint32 mode word for word, and
    against the byte lattice with ids lowered by off."""
    b = lmx.MpdRdmeSolver(0)
    spec.load_into(b, seed=seed)
    w = lmx.IntMpdRdmeSolver(0)
    int_spec.load_into(w, seed=seed)
    orc = oracle.Oracle(int_spec, seed_top=seed)
    assert b.seed == w.seed == orc.seed32 and orc.words
    for c in range(chunks):
        b.runTimesteps(steps // chunks)
        w.runTimesteps(steps // chunks)
        assert orc.run(steps // chunks) == 0
        np.testing.assert_array_equal(w.getLatticeNative(), orc.native(), err_msg=f"oracle, chunk {c}")
        lb = b.getLattice().astype(np.uint32)
        lb[lb > 0] += off
        lw = w.getLattice()
        assert lw.dtype == np.uint32
        if not np.array_equal(lw, lb):
            bad = np.argwhere(lw != lb)
            raise AssertionError(f"{int_spec.name} chunk {c}: {len(bad)} slots differ, first {bad[:5].tolist()}")
    cb, cw = b.getSpeciesCounts(), w.getSpeciesCounts()
    assert cw.size == int_spec.nS
    np.testing.assert_array_equal(cw[off:], cb)
    assert not cw[:off].any()
    assert w.getOverflowStats() == b.getOverflowStats()
    return b, w


CASES = [
    # shape, P, nS, nST, R, density, boundary
    ((8, 6, 5), 8, 4, 3, 6, 0.3, "periodic"),
    ((7, 5, 3), 16, 5, 2, 5, 0.25, "reflecting"),
    ((7, 5, 3), 4, 3, 2, 4, 0.4, "periodic"),
    ((33, 17, 9), 8, 6, 4, 10, 0.1, "reflecting"),
    ((16, 12, 10), 16, 8, 3, 12, 0.15, "periodic"),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-P{c[1]}-{c[6]}" for c in CASES])
@pytest.mark.parametrize("off", [0, 300])
def test_uint32_lattice_matches_byte_lattice(gpu, case, off):
    shape, P, nS, nST, R, density, boundary = case
    spec = models.random_model(shape=shape, P=P, nS=nS, nST=nST, R=R, density=density, seed=3, boundary=boundary)
    run_both(spec, shifted(spec, off), off, steps=12, chunks=3, seed=11)


@pytest.mark.parametrize("off", [0, 1000])
def test_uint32_lattice_overflows(gpu, off):
    # Nearly half full: many overflow entries per step, placed in canonical
    # order with 32-bit particles.
    spec = models.random_model(shape=(12, 10, 8), P=8, nS=3, nST=1, R=0, density=0.45, seed=11, qmax=0.5)
    b, w = run_both(spec, shifted(spec, off), off, steps=10, chunks=5, seed=13)
    assert w.getOverflowStats()[0] > 0


def test_uint32_lattice_bimolecular(gpu):
    spec = models.bimolecular(shape=(32, 32, 32), nA=1500, nB=1500, kf=1e9, seed=2)
    run_both(spec, shifted(spec, 500), 500, steps=30, chunks=3, seed=9)


def test_uint32_lattice_arguments(gpu):
    spec = models.random_model(shape=(6, 5, 4), P=8, nS=3, nST=2, R=2, density=0.2, seed=1)
    with pytest.raises(InvalidArgException, match="4 bytes per particle"):
        w = lmx.IntMpdRdmeSolver(0)
        w.initialize(0, spec.parameters(1), boundary=spec.boundary)
        w.buildModel(spec.nS, spec.R, spec.initial_counts(), spec.types, spec.K, spec.S, spec.D, 1)
        X, Y, Z = spec.shape
        w.buildDiffusionModel(spec.nST, spec.DF, spec.RL, X, Y, Z, spec.P, 1, spec.spacing, spec.lattice,
                              spec.sites, True)
    big = shifted(spec, 300)
    with pytest.raises(InvalidArgException, match="255 species"):
        big.load_into(lmx.MpdRdmeSolver(0), seed=1)
    w = lmx.IntMpdRdmeSolver(0)
    big.load_into(w, seed=1)
    lat = w.getLattice()
    lat[0, 0, 0, 0] = big.nS + 1
    with pytest.raises(InvalidArgException, match="invalid species"):
        w.setLattice(lat)
    lat[0, 0, 0, 0] = big.nS
    lat[0, 0, 0, 1:] = 0
    w.setLattice(lat)
    np.testing.assert_array_equal(w.getLattice()[0, 0, 0, :1], [big.nS])
    nat = w.getLatticeNative()
    assert nat.shape == (spec.P, 4, 5, 6)
    assert nat[0, 0, 0, 0] == big.nS
    nat[1, 0, 0, 0] = big.nS + 7
    with pytest.raises(InvalidArgException, match="invalid species"):
        w.setLatticeNative(nat)
    f = lmx.MESolverFactory()
    f.setSolver(lmx.INT_SOLVER_NAME)
    assert isinstance(f.instantiate(0), lmx.IntMpdRdmeSolver)


def test_uint32_lattice_whole_cell_full_size(gpu):
    """At the whole-cell size (256x256x512, P=8, 100 species renumbered past
    a byte), the uint32 lattice ends on the byte lattice's state after the
    same steps."""
    spec = models.whole_cell(seed=1)
    off = 300
    b = lmx.MpdRdmeSolver(0)
    spec.load_into(b, seed=3)
    b.runTimesteps(12)
    ref = b.getLatticeNative()
    b.close()
    w = lmx.IntMpdRdmeSolver(0)
    shifted(spec, off).load_into(w, seed=3)
    w.runTimesteps(12)
    got = w.getLatticeNative().astype(np.int64)
    w.close()
    got[got > 0] -= off
    P = spec.P
    as_bytes = got.astype(np.uint8).reshape(P // 4, 4, *got.shape[1:]).transpose(0, 2, 3, 4, 1)
    np.testing.assert_array_equal(np.ascontiguousarray(as_bytes).view(np.uint32)[..., 0], ref)


def test_uint32_lattice_is_not_batched(gpu):
    spec = models.random_model(shape=(8, 6, 5), P=8, nS=4, nST=2, R=3, density=0.2, seed=2)
    w = lmx.IntMpdRdmeSolver(0)
    spec.load_into(w, seed=1)
    ens = lmx.Ensemble(0)
    try:
        with pytest.raises(InvalidArgException, match="byte lattice"):
            ens.add(w)
    finally:
        ens.close()
        w.close()
