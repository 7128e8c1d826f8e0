"""The oracle's uint32 lattice (IntMpdRdmeSolver: one particle per word).

The word kernels of the reference (word_diffusion_1d_dev.cu,
word_reaction_dev.cu) make the byte kernels' decisions with the same RNG keys,
so the oracle's word mode is pinned against its byte mode (itself pinned by the
rule, RNG and table tests): the same trajectory for a model that fits a byte,
and for more than 255 species the trajectory of the compact model with every
species id raised by a constant (the extra low ids absent and inert, so every
table entry, sum order and product order is unchanged).  Layout and counting
follow IntLattice (IntLattice.cpp:172-190, 419-435) and are checked on hand
cases.
"""
import numpy as np
import pytest

import oracle
from lattice_microbes_amd import models


def shifted(spec, off):
    nS = spec.nS + off
    S = np.zeros((nS, spec.R), np.int32)
    S[off:] = spec.S
    D = np.zeros((nS, spec.R), np.uint32)
    D[off:] = spec.D
    DF = np.zeros((spec.nST, spec.nST, nS))
    DF[..., off:] = spec.DF
    lat = spec.lattice.astype(np.uint32)
    lat[lat > 0] += off
    return models.ModelSpec(name=spec.name, shape=spec.shape, spacing=spec.spacing, tau=spec.tau, P=spec.P,
                            nS=nS, nST=spec.nST, types=spec.types, K=spec.K, S=S, D=D, DF=DF, RL=spec.RL,
                            lattice=lat, sites=spec.sites, boundary=spec.boundary)


def words_as_bytes(native, off):
    """(P, Z, Y, X) words with ids raised by off -> (P/4, Z, Y, X) byte-native words."""
    w = native.astype(np.int64)
    w[w > 0] -= off
    assert w.min() >= 0 and w.max() < 256
    P = w.shape[0]
    b = w.astype(np.uint8).reshape(P // 4, 4, *w.shape[1:]).transpose(0, 2, 3, 4, 1)
    return np.ascontiguousarray(b).view(np.uint32)[..., 0]


CASES = [
    ((8, 6, 5), 8, 4, 3, 6, 0.3, "periodic"),
    ((7, 5, 3), 16, 5, 2, 5, 0.25, "reflecting"),
    ((7, 5, 3), 4, 3, 2, 4, 0.4, "periodic"),
    ((12, 10, 8), 8, 3, 1, 0, 0.45, "periodic"),   # overflow-heavy
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-P{c[1]}-{c[6]}" for c in CASES])
@pytest.mark.parametrize("off", [0, 300])
def test_word_mode_matches_byte_mode(case, off):
    shape, P, nS, nST, R, density, boundary = case
    spec = models.random_model(shape=shape, P=P, nS=nS, nST=nST, R=R, density=density, seed=5, boundary=boundary,
                               qmax=0.5)
    b = oracle.Oracle(spec, seed_top=3)
    w = oracle.Oracle(shifted(spec, off), seed_top=3)
    assert w.words and not b.words
    for _ in range(4):
        assert b.run(3) == 0 and w.run(3) == 0
        np.testing.assert_array_equal(words_as_bytes(w.native(), off), b.native())
    assert w.stats.overflow_entries == b.stats.overflow_entries
    assert w.stats.reactions == b.stats.reactions
    cw = w.species_counts()
    np.testing.assert_array_equal(cw[off:], b.species_counts())
    assert not cw[:off].any()
    rw = w.row_major().astype(np.int64)
    rw[rw > 0] -= off
    np.testing.assert_array_equal(rw, b.row_major())
    if R == 0:
        assert b.stats.overflow_entries > 0


def test_word_layout_hand_case():
    """Particles inserted at the first empty word of their site in row-major
    order; counts are every nonzero word; the row-major view reads them back."""
    X, Y, Z, P = 3, 2, 2, 4
    rm = np.zeros((X, Y, Z, P), np.uint32)
    rm[1, 0, 1, 2] = 300   # slot 2 in the file: packed into slot 0
    rm[1, 0, 1, 3] = 5
    rm[2, 1, 0, 0] = 300
    spec = models.ModelSpec(name="hand", shape=(X, Y, Z), spacing=1e-8, tau=1e-6, P=P, nS=300, nST=1,
                            types=np.zeros(0, np.uint32), K=np.zeros(0), S=np.zeros((300, 0), np.int32),
                            D=np.zeros((300, 0), np.uint32), DF=np.zeros((1, 1, 300)),
                            RL=np.zeros((0, 1), np.uint32), lattice=rm, sites=np.zeros((X, Y, Z), np.uint8))
    o = oracle.Oracle(spec)
    nat = o.native()
    assert nat.shape == (P, Z, Y, X)
    assert nat[0, 1, 0, 1] == 300 and nat[1, 1, 0, 1] == 5 and nat[2:, 1, 0, 1].sum() == 0
    assert nat[0, 0, 1, 2] == 300
    counts = o.species_counts()
    assert counts[300 - 1] == 2 and counts[5 - 1] == 1 and counts.sum() == 3
    back = o.row_major()
    assert back[1, 0, 1, :2].tolist() == [300, 5] and back[2, 1, 0, 0] == 300
