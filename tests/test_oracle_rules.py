"""Pins the oracle's sweep, reaction and overflow rules on hand-built cases
whose outcome is predicted here from the independent Python RNG
transcription (tests/golden/make_rng_kat.py) and the reference's rules:
direction and periodic wrap of moves, arrival order, reflecting edges, hop
frequencies against T, reaction occurrence and selection, stoichiometry, and
overflow placement order."""
import ctypes
import os
import sys
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle
from helpers import q_to_D, site_contents, spec
from lattice_microbes_amd import models

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_rng_kat import hash32, timestep_hash  # noqa: E402

getcontext().prec = 50


def U(v1, shift, v2, seed):
    return np.float32(hash32((v1 << shift) | v2, seed)) * np.float32(2.0 ** -32)


def sweep(o, axis, t, formulation=0):
    """One oracle sweep along axis at timestep t; returns (native, overflows)."""
    lib = o.lib
    dst = np.zeros_like(o.lattice)
    ovs = np.zeros(2048, np.uint32)
    ovp = np.zeros(2048, np.uint32)
    cnt = ctypes.c_uint32(0)
    h = timestep_hash(o.seed32, t, axis)
    rc = lib.orc_sweep(ctypes.byref(o.model), oracle._p(o.lattice, ctypes.c_uint32), oracle._p(o.sites, ctypes.c_uint8),
                       oracle._p(dst, ctypes.c_uint32), axis, h, formulation, oracle._p(ovs, ctypes.c_uint32),
                       oracle._p(ovp, ctypes.c_uint32), ctypes.byref(cnt))
    assert rc == 0
    X, Y, Z = o.shape
    n = cnt.value
    return dst.reshape(o.P // 4, Z, Y, X), list(zip(ovs[:n].tolist(), ovp[:n].tolist()))


def idx(shape, x, y, z):
    X, Y, _ = shape
    return x + X * y + X * Y * z


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("boundary", ["periodic", "reflecting"])
def test_single_particle_moves_follow_hash(axis, boundary):
    """q = 0.5 both ways: the particle moves minus iff u < 0.5 and plus iff
    0.5 <= u < 1, wrapping periodically or blocked at a reflecting edge."""
    shape = (5, 4, 3)
    L = shape[axis]
    D = q_to_D([0.5])
    for seed_top in range(1, 13):
        for c in (0, L - 1, 1):
            pos = [1, 1, 1]
            pos[axis] = c
            s = spec(shape, DF=D.reshape(1, 1, 1), particles=[(*pos, 1)], boundary=boundary)
            o = oracle.Oracle(s, seed_top=seed_top)
            nat, ovf = sweep(o, axis, 0)
            assert not ovf
            i = idx(shape, *pos)
            u = U(i, 3, 0, timestep_hash(o.seed32, 0, axis))
            dest = list(pos)
            if u < 0.5:
                if c > 0:
                    dest[axis] = c - 1
                elif boundary == "periodic":
                    dest[axis] = L - 1
            elif u < 1.0:
                if c < L - 1:
                    dest[axis] = c + 1
                elif boundary == "periodic":
                    dest[axis] = 0
            got = np.argwhere(nat[0] != 0)
            assert got.tolist() == [[dest[2], dest[1], dest[0]]]


def test_arrival_order_stay_then_minus_then_plus():
    """performPropagation order: own stayers, then the minus neighbour's plus
    movers, then the plus neighbour's minus movers (slot order within each)."""
    shape = (6, 1, 1)
    D = q_to_D([0.0, 0.5, 0.5]).reshape(1, 1, 3)
    found = 0
    for seed_top in range(1, 400):
        s32 = oracle.seed32(seed_top, 0)
        h = timestep_hash(s32, 0, 0)
        if not (0.5 <= U(1, 3, 0, h) < 1.0 and U(3, 3, 0, h) < 0.5):
            continue
        s = spec(shape, nS=3, DF=D, particles=[(2, 0, 0, 1), (1, 0, 0, 2), (3, 0, 0, 3)])
        o = oracle.Oracle(s, seed_top=seed_top)
        nat, _ = sweep(o, 0, 0)
        assert site_contents(nat, shape, 8, 2, 0, 0) == [1, 2, 3, 0, 0, 0, 0, 0]
        found += 1
        if found == 3:
            break
    assert found == 3


def test_hop_frequencies_match_transition_table():
    """Alternating site types along x: a particle at x = 1 mod 4 (type 0) has
    a type-0 minus neighbour and a type-1 plus neighbour, so it moves minus
    with probability T[0][0] and plus with probability T[0][1]."""
    X, Y, Z = 256, 16, 16
    types = np.array([0, 0, 1, 1] * (X // 4), dtype=np.uint8)
    sites = np.broadcast_to(types[:, None, None], (X, Y, Z)).copy()
    q = np.zeros((2, 2, 1))
    q[0, 0, 0], q[0, 1, 0] = 0.1, 0.35
    parts = [(x, y, z, 1) for x in range(1, X, 4) for y in range(Y) for z in range(Z)]
    s = spec((X, Y, Z), nST=2, DF=q_to_D(q), particles=parts, sites=sites)
    o = oracle.Oracle(s, seed_top=3)
    nat, _ = sweep(o, 0, 0)
    occ = nat[0] != 0
    n = len(parts)
    minus = occ[:, :, 0::4].sum() / n
    plus = occ[:, :, 2::4].sum() / n
    tol = 5 * np.sqrt(0.35 * 0.65 / n)
    assert abs(minus - 0.1) < tol and abs(plus - 0.35) < tol


@pytest.mark.parametrize("P", [4, 8, 16])
@pytest.mark.parametrize("boundary", ["periodic", "reflecting"])
def test_two_formulations_agree(P, boundary):
    for seed in range(4):
        m = models.random_model(shape=(9, 7, 6), P=P, nS=5, nST=3, R=0, density=0.45, seed=seed, boundary=boundary)
        a = oracle.Oracle(m, seed_top=seed + 1)
        b = oracle.Oracle(m, seed_top=seed + 1)
        for axis in range(3):
            na, oa = sweep(a, axis, 7, formulation=0)
            nb, ob = sweep(b, axis, 7, formulation=1)
            assert np.array_equal(na, nb) and oa == ob


def test_conservation_and_compaction():
    m = models.random_model(shape=(12, 10, 8), P=8, nS=4, nST=2, R=0, density=0.2, seed=5)
    o = oracle.Oracle(m, seed_top=9)
    before = o.species_counts().copy()
    assert o.run(25) == 0
    assert np.array_equal(o.species_counts(), before)
    rm = o.row_major().reshape(-1, 8)
    nz = rm != 0
    # packed: no empty slot before an occupied one
    assert not np.any(~nz[:, :-1] & nz[:, 1:])
    assert np.array_equal(o.sites, m.native_sites().reshape(-1))


def exact_prob(a32):
    return np.float32(1.0 - float(Decimal(float(a32)).exp() ** -1))


def test_reaction_occurrence_and_selection_per_site():
    """Sites holding one A each, two first-order channels A->B (k1) and A->C
    (k2): a site reacts iff U(i,1,0) <= 1-exp(-a) with a = (k1+k2) tau, and
    picks channel 0 iff U(i,1,1)*a - p0 <= 0 (determineReactionIndex)."""
    shape = (16, 16, 16)
    k1, k2 = 2.0e5, 0.7e5
    rx = [(1, k1, {0: -1, 1: 1}, [0]), (1, k2, {0: -1, 2: 1}, [0])]
    parts = [(x, y, z, 1) for x in range(0, 16, 2) for y in range(16) for z in range(16)]
    s = spec(shape, nS=3, reactions=rx, particles=parts)
    o = oracle.Oracle(s, seed_top=21)
    assert o.run(1) == 0
    nat = o.native()
    h3 = timestep_hash(o.seed32, 0, 3)
    # qp1[A] accumulates in double and stores float (computePropensities)
    a = np.float32(float(np.float32(k1 * 1e-6)) + k2 * 1e-6)
    assert o.qp1[0] == a
    prob = exact_prob(a)
    assert prob == np.float32(oracle.reaction_probability(float(a)))
    p0 = np.float32(k1 * 1e-6)
    reacted = 0
    for (x, y, z, _) in parts:
        i = idx(shape, x, y, z)
        got = site_contents(nat, shape, 8, x, y, z)[0]
        if U(i, 1, 0, h3) <= prob:
            rp = np.float32(U(i, 1, 1, h3) * a)
            # channel 0 taken first (rp > 0 -> idx 0), then channel 1 if rp - p0 > 0
            ch = 1 if np.float32(rp - p0) > 0 else 0
            assert got == (2 if ch == 0 else 3)
            reacted += 1
        else:
            assert got == 1
    assert 0 < reacted < len(parts)


def test_stoichiometry_consumes_first_reactants_and_appends_products():
    """evaluateReaction: surviving particles keep their order, the first
    matching reactants are consumed, products follow in species order."""
    rx = [(2, 1e12, {0: -1, 1: -1, 3: 1}, [0, 1])]  # A + B -> D, certain to fire
    s = spec((4, 4, 4), nS=4, reactions=rx,
             particles=[(1, 1, 1, 1), (1, 1, 1, 3), (1, 1, 1, 1), (1, 1, 1, 2)])
    o = oracle.Oracle(s, seed_top=2)
    assert o.run(1) == 0
    assert site_contents(o.native(), (4, 4, 4), 8, 1, 1, 1) == [3, 1, 4, 0, 0, 0, 0, 0]
    # self reaction 2A -> B with three A: two consumed, one kept, B appended
    rx = [(3, 1e12, {0: -2, 1: 1}, [0])]
    s = spec((4, 4, 4), nS=2, reactions=rx, particles=[(2, 2, 2, 1)] * 3)
    o = oracle.Oracle(s, seed_top=2)
    assert o.run(1) == 0
    assert site_contents(o.native(), (4, 4, 4), 8, 2, 2, 2)[:3] == [1, 2, 0]


def test_overflow_placed_at_first_free_candidate():
    """A full site receives a fifth particle: the extra goes to the first
    candidate of the radius-ordered search (x outer, then y, then z) with room
    and the origin's site type."""
    shape = (5, 5, 5)
    q = np.zeros((2, 2, 2))
    q[0, 0, 1] = 0.5  # species 2 moves within type 0
    for seed_top in range(1, 300):
        h = timestep_hash(oracle.seed32(seed_top, 0), 0, 0)
        if 0.5 <= U(idx(shape, 1, 2, 2), 2, 0, h) < 1.0:  # P=4: 2 key bits
            break
    sites = np.zeros(shape, np.uint8)
    full = [(2, 2, 2, 1)] * 4
    s = spec(shape, P=4, nS=2, nST=2, DF=q_to_D(q), particles=full, sites=sites)
    # (1,2,2) holds the mover (species 2, slot 0) and three immobile particles
    s.lattice[1, 2, 2, :] = [2, 1, 1, 1]
    o = oracle.Oracle(s, seed_top=seed_top)
    assert o.run(1) == 0
    nat = o.native()
    # mover left (1,2,2) (3 immobile remain, room for 1), arrived at the full
    # (2,2,2) -> overflow -> radius 1 candidates in order: (1,2,2) has room
    # and type 0 -> placed back there
    assert site_contents(nat, shape, 4, 2, 2, 2) == [1, 1, 1, 1]
    assert site_contents(nat, shape, 4, 1, 2, 2) == [1, 1, 1, 2]
    assert o.stats.overflow_entries == 1


def test_overflow_skips_other_site_types():
    shape = (5, 5, 5)
    rx = [(1, 1e12, {0: 1}, [0])]  # A -> 2A, certain to fire
    sites = np.ones(shape, np.uint8)
    sites[2, 2, 2] = 0
    sites[1, 2, 2] = 1
    sites[2, 1, 2] = 0  # (2,1,2) is the 2nd radius-1 candidate: first of type 0
    s = spec(shape, P=4, nS=1, nST=2, reactions=rx, particles=[(2, 2, 2, 1)] * 4, sites=sites)
    s.RL[:, 1] = 0
    o = oracle.Oracle(s, seed_top=4)
    assert o.run(1) == 0
    nat = o.native()
    assert site_contents(nat, shape, 4, 2, 2, 2) == [1, 1, 1, 1]
    assert site_contents(nat, shape, 4, 2, 1, 2) == [1, 0, 0, 0]


def test_overflow_unplaceable_and_out_of_range_errors():
    rx = [(1, 1e12, {0: 1}, [0])]
    sites = np.ones((5, 5, 5), np.uint8)
    sites[2, 2, 2] = 0
    s = spec((5, 5, 5), P=4, nS=1, nST=2, reactions=rx, particles=[(2, 2, 2, 1)] * 4, sites=sites)
    s.RL[:, 1] = 0
    assert oracle.Oracle(s, seed_top=4).run(1) == oracle.ORC_ERR_OVERFLOW_UNPLACEABLE
    # a 2-wide lattice: at radius 2 the reference's unsigned bound reaches x=2
    sites = np.ones((2, 5, 5), np.uint8)
    sites[0, 2, 2] = 0
    s = spec((2, 5, 5), P=4, nS=1, nST=2, reactions=rx, particles=[(0, 2, 2, 1)] * 4, sites=sites)
    s.RL[:, 1] = 0
    assert oracle.Oracle(s, seed_top=4).run(1) == oracle.ORC_ERR_OVERFLOW_OUT_OF_RANGE


def test_overflow_cap():
    m = models.random_model(shape=(16, 16, 16), P=4, nS=2, nST=1, R=0, density=0.9, seed=3)
    m.DF[:] = 0.4999 * m.spacing ** 2 / m.tau
    assert oracle.Oracle(m, seed_top=3).run(1) == oracle.ORC_ERR_OVERFLOW_CAP


def test_species_counts_stop_at_first_empty_byte_per_word():
    # ByteLattice::getParticleCounts counts each word up to its first 0 byte.
    lib = oracle.load()
    native = np.array([0x00020001, 0x03000001], dtype=np.uint32)  # N=2, P=4, W=1
    out = np.zeros(3, np.uint32)
    lib.orc_species_counts(2, 4, oracle._p(native, ctypes.c_uint32), 3, oracle._p(out, ctypes.c_uint32))
    assert out.tolist() == [2, 0, 0]
