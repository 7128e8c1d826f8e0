"""Pins the next-subvolume counterpart used as the CPU baseline: its event
queue against the reference's own ReactionQueue.h (compiled into
oracle/_ref/ from the reference tree) and the reference test's known
answers (testing/c/lm/reaction/reaction_queue.cpp:103-240), and its
dynamics against the exact stationary distribution of A + B <-> C."""
import ctypes
import math
import os

import numpy as np
import pytest

import oracle
from helpers import q_to_D, spec

REF_HEAP = os.path.join(os.path.dirname(oracle.__file__), "_ref", "libref_heap.so")


def our_heap(n, idx, times):
    lib = oracle.load()
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    times = np.ascontiguousarray(times, dtype=np.float64)
    top = np.zeros(idx.size, np.uint32)
    pos = np.zeros(idx.size * n, np.uint32)
    assert lib.orc_heap_sequence(n, idx.size, oracle._p(idx, ctypes.c_uint32), oracle._p(times, ctypes.c_double),
                                 oracle._p(top, ctypes.c_uint32), oracle._p(pos, ctypes.c_uint32)) == 0
    return top, pos.reshape(idx.size, n)


def test_heap_reference_known_answers():
    top, pos = our_heap(6, [0, 5, 2, 0], [15.765, 1.786, 100.01, 1752.0])
    assert pos.tolist() == [[1, 2, 3, 4, 5, 6], [3, 2, 6, 4, 5, 1], [3, 2, 6, 4, 5, 1], [6, 2, 3, 4, 5, 1]]
    assert top.tolist() == [0, 5, 5, 5]


@pytest.mark.skipif(not os.path.exists(REF_HEAP), reason="reference heap not built (needs the reference tree)")
def test_heap_matches_reference_build():
    ref = ctypes.CDLL(REF_HEAP)
    ref.ref_heap_sequence.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint),
                                      ctypes.POINTER(ctypes.c_uint)]
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 64, 1000):
        nupd = 3 * n + 50
        idx = rng.integers(0, n, nupd).astype(np.uint32)
        times = rng.exponential(1.0, nupd)
        times[rng.random(nupd) < 0.1] = np.inf
        times[rng.random(nupd) < 0.05] = 0.5  # ties
        top, pos = our_heap(n, idx, times)
        rtop = np.zeros(nupd, np.uint32)
        rpos = np.zeros(nupd * n, np.uint32)
        assert ref.ref_heap_sequence(n, nupd, oracle._p(idx, ctypes.c_uint), oracle._p(times, ctypes.c_double),
                                     oracle._p(rtop, ctypes.c_uint), oracle._p(rpos, ctypes.c_uint)) == 0
        assert np.array_equal(top, rtop)
        assert np.array_equal(pos, rpos.reshape(nupd, n))


def nsm_handle(s, seed=1):
    lib = oracle.load()
    o = oracle.Oracle(s, seed_top=seed)
    X, Y, Z = s.shape
    S = np.ascontiguousarray(s.S, dtype=np.int32).reshape(-1)
    DF = np.ascontiguousarray(s.DF, dtype=np.float64).reshape(-1)
    h = lib.orc_nsm_create(X, Y, Z, s.P, s.nS, s.nST, s.R, oracle._p(o.types, ctypes.c_uint32),
                           oracle._p(o.k, ctypes.c_double), oracle._p(o.si, ctypes.c_uint32),
                           oracle._p(o.s2i, ctypes.c_uint32), oracle._p(S, ctypes.c_int32),
                           oracle._p(DF, ctypes.c_double), oracle._p(o.RL32, ctypes.c_uint32), s.spacing,
                           oracle._p(o.lattice, ctypes.c_uint32), oracle._p(o.sites, ctypes.c_uint8), seed, 1)
    assert h
    return lib, h, (o, S, DF)


def test_nsm_single_subvolume_matches_exact_stationary_distribution():
    """One subvolume: NSM is Gillespie's SSA for A + B <-> C.  Detailed
    balance gives pi(c) ~ (kf/kr)^c / ((A0-c)! (B0-c)! c!)."""
    A0 = B0 = 5
    kf, kr = 2.0, 3.0
    rx = [(2, kf, {0: -1, 1: -1, 2: 1}, [0, 1]), (1, kr, {0: 1, 1: 1, 2: -1}, [2])]
    s = spec((1, 1, 1), P=16, nS=3, reactions=rx, particles=[(0, 0, 0, 1)] * A0 + [(0, 0, 0, 2)] * B0)
    lib, h, keep = nsm_handle(s, seed=7)
    nrec = 40000
    dt = 0.5
    rec = np.zeros(nrec * 3, np.uint32)
    nw = ctypes.c_uint32(0)
    assert lib.orc_nsm_run(h, nrec * dt + 1, 1 << 62, dt, nrec, oracle._p(rec, ctypes.c_uint32), ctypes.byref(nw)) == 0
    lib.orc_nsm_destroy(h)
    c = rec.reshape(nrec, 3)[100:, 2].astype(np.int64)
    w = np.array([(kf / kr) ** k / (math.factorial(A0 - k) ** 2 * math.factorial(k)) for k in range(6)])
    pi = w / w.sum()
    mean, var = (pi * np.arange(6)).sum(), (pi * np.arange(6) ** 2).sum() - (pi * np.arange(6)).sum() ** 2
    # records dt apart are correlated (relaxation ~ 1/(kf*5+kr) << dt): treat as independent
    se = math.sqrt(var / c.size)
    assert abs(c.mean() - mean) < 5 * se
    hist = np.bincount(c, minlength=6)[:6] / c.size
    assert np.all(np.abs(hist - pi) < 5 * np.sqrt(pi * (1 - pi) / c.size) + 1e-3)


def test_nsm_hop_rate():
    """A lone particle hops at total rate 6 D / lambda^2 (six neighbours)."""
    s = spec((8, 8, 8), P=8, DF=q_to_D([0.3]).reshape(1, 1, 1), particles=[(3, 3, 3, 1)])
    lib, h, keep = nsm_handle(s, seed=3)
    assert lib.orc_nsm_run(h, 1e30, 200000, 0.0, 0, None, None) == 0
    t = lib.orc_nsm_time(h)
    ev = lib.orc_nsm_events(h)
    lib.orc_nsm_destroy(h)
    rate = 6 * s.DF[0, 0, 0] / s.spacing ** 2
    assert abs(ev / t / rate - 1) < 5 / math.sqrt(ev)
