"""The oracle's relaxed overflow handler (correct_overflows,
MpdRdmeSolver.cu:1677-1761), pinned three ways:

* a known-answer case worked by hand from the kernel's rules: the list is
  the carried entries then the step's, entries of one site go in list order
  (the kernel's rounds), each into its site's first empty slot, and an entry
  whose site is full stays listed;
* equality with the classic handler on every single step whose overflows all
  found room in their own site: classic tries the site itself first (radius
  0 of Lattice::getNearbySites), so a wrong slot or order would differ;
* conservation: with no reactions, the particles on the lattice plus the
  carried ones are the initial counts after every step, and every carried
  entry's site is full.
"""
import ctypes
from ctypes import c_uint32

import numpy as np
import pytest

from lattice_microbes_amd import models
import oracle


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(c_uint32))


def _tiny(words):
    spec = models.random_model(shape=(4, 1, 1), P=8 if not words else 4, nS=3, nST=1, R=0, density=0.0, seed=1)
    if words:
        spec.lattice = spec.lattice.astype(np.uint32)
    return oracle.Oracle(spec, seed_top=3, overflow="relaxed")


def _slots(orc, site):
    lat = orc.lattice.reshape(orc.W, orc.N)
    if orc.words:
        return [int(v) for v in lat[:, site]]
    return [int((lat[w, site] >> (8 * b)) & 0xFF) for w in range(orc.W) for b in range(4)]


def _set_site(orc, site, particles):
    lat = orc.lattice.reshape(orc.W, orc.N)
    lat[:, site] = 0
    for j, v in enumerate(particles):
        if orc.words:
            lat[j, site] = v
        else:
            lat[j // 4, site] |= np.uint32(v << (8 * (j % 4)))


@pytest.mark.parametrize("words", [False, True], ids=["byte", "uint32"])
def test_relaxed_known_answer(words):
    orc = _tiny(words)
    P = orc.P
    big = 300 if words else 9
    _set_site(orc, 0, list(range(1, P)))        # site 0: one free slot
    _set_site(orc, 1, [2] * P)                  # site 1: full
    carry_site = np.zeros(2048, np.uint32)
    carry_part = np.zeros(2048, np.uint32)
    carry_n = np.zeros(1, np.uint32)
    carry_site[0], carry_part[0], carry_n[0] = 0, 3, 1          # carried from an earlier step
    ovf_site = np.array([0, 1, 0, 2], np.uint32)
    ovf_part = np.array([5, 6, 7, big], np.uint32)
    rc = orc.lib.orc_resolve_relaxed(ctypes.byref(orc.model), _p(orc.lattice), _p(ovf_site), _p(ovf_part), 4,
                                     _p(carry_site), _p(carry_part), _p(carry_n))
    assert rc == 0
    assert _slots(orc, 0) == list(range(1, P)) + [3]            # the carried entry goes first
    assert _slots(orc, 1) == [2] * P
    assert _slots(orc, 2) == [big] + [0] * (P - 1)
    assert _slots(orc, 3) == [0] * P
    assert int(carry_n[0]) == 3
    assert carry_site[:3].tolist() == [0, 1, 0] and carry_part[:3].tolist() == [5, 6, 7]
    # site 0 empties: its two carried entries land in list order; site 1 stays full
    _set_site(orc, 0, [])
    rc = orc.lib.orc_resolve_relaxed(ctypes.byref(orc.model), _p(orc.lattice), _p(ovf_site), _p(ovf_part), 0,
                                     _p(carry_site), _p(carry_part), _p(carry_n))
    assert rc == 0
    assert _slots(orc, 0) == [5, 7] + [0] * (P - 2)
    assert int(carry_n[0]) == 1 and (carry_site[0], carry_part[0]) == (1, 6)
    # carried plus new beyond the list's 2048 entries is the capacity error
    carry_n[0] = 2000
    rc = orc.lib.orc_resolve_relaxed(ctypes.byref(orc.model), _p(orc.lattice), _p(ovf_site), _p(ovf_part), 49,
                                     _p(carry_site), _p(carry_part), _p(carry_n))
    assert rc == oracle.ORC_ERR_OVERFLOW_CAP


def test_relaxed_equals_classic_when_own_sites_have_room():
    spec = models.random_model(shape=(12, 10, 8), P=8, nS=3, nST=2, R=4, density=0.3, seed=3, qmax=0.5)
    checked = 0
    state = oracle.Oracle(spec, seed_top=13)
    for t0 in range(12):
        a = oracle.Oracle(spec, seed_top=13, overflow="relaxed")
        b = oracle.Oracle(spec, seed_top=13)
        a.lattice[:] = state.lattice
        b.lattice[:] = state.lattice
        a.timestep = b.timestep = t0
        assert a.run(1) == 0 and b.run(1) == 0
        if a.stats.overflow_entries > 0 and a.unresolved == 0:
            np.testing.assert_array_equal(a.native(), b.native())
            checked += 1
        assert state.run(1) == 0
    assert checked >= 3


@pytest.mark.parametrize("P,density", [(4, 0.6), (8, 0.45)])
def test_relaxed_conserves_particles(P, density):
    spec = models.random_model(shape=(12, 10, 8), P=P, nS=3, nST=1, R=0, density=density, seed=11, qmax=0.5)
    orc = oracle.Oracle(spec, seed_top=13, overflow="relaxed")
    initial = orc.species_counts().astype(np.int64)
    carried_any = False
    for _ in range(8):
        assert orc.run(1) == 0
        n = orc.unresolved
        carried_any |= n > 0
        carried = np.bincount(orc.carry_particle[:n], minlength=spec.nS + 1)[1:]
        np.testing.assert_array_equal(orc.species_counts() + carried, initial)
        for site in orc.carry_site[:n]:
            assert 0 not in _slots(orc, int(site))
    assert carried_any
