"""GPU parity: the HIP step path against the serial oracle, bit for bit.

Every case builds the same model arrays for the product (through the C ABI)
and for the oracle, runs the same seeded steps and compares the native
lattices word for word.  Integer state, so the bar is exact equality.
"""
import dataclasses

import numpy as np
import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd import models
import oracle

pytestmark = pytest.mark.gpu


def run_pair(spec, steps, seed=5, replicate=0, chunks=1, path=0):
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, replicate=replicate, seed=seed)
    if path:
        solver.setPath(path)
    orc = oracle.Oracle(spec, seed_top=seed, replicate=replicate)
    assert solver.seed == orc.seed32
    per = steps // chunks
    for _ in range(chunks):
        solver.runTimesteps(per)
        rc = orc.run(per)
        assert rc == 0, f"oracle error {rc}"
        gpu = solver.getLatticeNative()
        ref = orc.native()
        if not np.array_equal(gpu, ref):
            bad = np.argwhere(gpu != ref)
            raise AssertionError(f"{spec.name} {spec.shape} P={spec.P}: {len(bad)} words differ, first {bad[:5].tolist()}")
    return solver, orc


CASES = [
    # shape, P, nS, nST, R, density, boundary
    ((8, 6, 5), 8, 4, 3, 6, 0.3, "periodic"),
    ((8, 6, 5), 8, 4, 3, 6, 0.3, "reflecting"),
    ((7, 5, 3), 4, 3, 2, 4, 0.4, "periodic"),
    ((7, 5, 3), 16, 5, 2, 5, 0.25, "reflecting"),
    ((1, 4, 6), 8, 3, 2, 3, 0.2, "periodic"),
    ((2, 2, 2), 8, 3, 1, 3, 0.3, "periodic"),
    ((33, 17, 9), 8, 6, 4, 10, 0.1, "periodic"),
    ((64, 8, 8), 16, 6, 3, 8, 0.2, "periodic"),
    ((20, 20, 20), 8, 8, 5, 12, 0.05, "reflecting"),
    ((31, 7, 13), 4, 4, 3, 6, 0.5, "periodic"),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-P{c[1]}-{c[6]}" for c in CASES])
def test_random_models_bit_exact(gpu, case):
    shape, P, nS, nST, R, density, boundary = case
    for seed in (1, 2):
        spec = models.random_model(shape=shape, P=P, nS=nS, nST=nST, R=R, density=density, seed=seed,
                                   boundary=boundary)
        run_pair(spec, steps=8, seed=seed * 7 + 1, replicate=seed, chunks=2)


def test_tables_match_oracle(gpu):
    spec = models.random_model(shape=(9, 7, 5), P=8, nS=6, nST=4, R=12, seed=3)
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=3)
    orc = oracle.Oracle(spec, seed_top=3)
    for name, ref in (("T", orc.T), ("qp0", orc.qp0), ("qp1", orc.qp1), ("qp2", orc.qp2), ("rate", orc.rate)):
        got = solver.getTable(name)
        assert got.view(np.uint32).tolist() == ref[: got.size].view(np.uint32).tolist(), name


def test_bimolecular_64(gpu):
    spec = models.bimolecular(shape=(64, 64, 64), nA=3000, nB=3000, kf=1e9, seed=2)
    solver, orc = run_pair(spec, steps=40, seed=9, chunks=4)
    np.testing.assert_array_equal(solver.getSpeciesCounts(), orc.species_counts())


def test_capsule_reduced(gpu):
    spec = models.capsule_minde(shape=(32, 32, 64), counts=(800, 800, 400), seed=4)
    run_pair(spec, steps=30, seed=4, chunks=3)


def test_dense_overflow_heavy(gpu):
    # Nearly half full: many overflows per step, resolved in canonical order.
    spec = models.random_model(shape=(12, 10, 8), P=8, nS=3, nST=1, R=0, density=0.45, seed=11,
                               qmax=0.5)
    solver, orc = run_pair(spec, steps=10, seed=13, chunks=5)
    e_gpu, s_gpu = solver.getOverflowStats()
    assert e_gpu == orc.stats.overflow_entries and e_gpu > 0
    assert s_gpu == orc.stats.overflow_steps


def test_species_counts_and_readback(gpu):
    spec = models.random_model(shape=(10, 9, 8), P=16, nS=7, nST=2, R=5, density=0.2, seed=21)
    solver, orc = run_pair(spec, steps=6, seed=2)
    np.testing.assert_array_equal(solver.getSpeciesCounts(), orc.species_counts())
    np.testing.assert_array_equal(solver.getLattice(), orc.row_major())
    np.testing.assert_array_equal(solver.getSites(), spec.sites)


def test_device_exp_matches_oracle(gpu):
    rng = np.random.default_rng(0)
    x = np.concatenate([-rng.uniform(0, 3, 20000), -10.0 ** rng.uniform(-9, 2.8, 20000),
                        rng.uniform(-750, 710, 5000), [0.0, -0.0, -745.2, 709.8, -1e-300]])
    y = np.zeros_like(x)
    lib = lmx.load_library()
    import ctypes
    assert lib.lmx_device_exp(0, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              y.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), x.size) == 0
    ref = np.array([oracle.exp(float(v)) for v in x])
    assert np.array_equal(y.view(np.uint64), ref.view(np.uint64))
    a = np.concatenate([10.0 ** rng.uniform(-8, 1.5, 20000), [1e-30, 0.5, 20.0]]).astype(np.float32)
    pr = np.zeros_like(a)
    assert lib.lmx_device_reaction_probability(0, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                               pr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.size) == 0
    ref = np.array([oracle.reaction_probability(float(v)) for v in a], dtype=np.float32)
    assert np.array_equal(pr.view(np.uint32), ref.view(np.uint32))


def test_set_reaction_rate_mid_run(gpu):
    spec = models.random_model(shape=(10, 8, 6), P=8, nS=4, nST=2, R=6, density=0.3, seed=5)
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=6)
    orc = oracle.Oracle(spec, seed_top=6)
    solver.runTimesteps(3)
    assert orc.run(3) == 0
    # change reaction 1's rate on both sides
    new_k = float(np.float32(spec.K[1] * 3.0))
    solver.setReactionRate(1, new_k)
    spec2 = models.ModelSpec(**{**spec.__dict__})
    spec2.K = spec.K.copy()
    spec2.K[1] = new_k
    orc2 = oracle.Oracle(spec2, seed_top=6)
    orc2.lattice[:] = orc.lattice
    orc2.timestep = orc.timestep
    solver.runTimesteps(4)
    assert orc2.run(4) == 0
    np.testing.assert_array_equal(solver.getLatticeNative(), orc2.native())


def test_overflow_cap_error_matches(gpu):
    # Every site full and every particle mobile: far more than 2048 overflow
    # entries in the first step; both sides report the too-many-overflows error.
    spec = models.random_model(shape=(16, 16, 16), P=4, nS=2, nST=1, R=0, density=0.9, seed=3, qmax=0.5)
    spec.DF[:] = 0.5 * spec.spacing ** 2 / spec.tau * 0.999
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=3)
    orc = oracle.Oracle(spec, seed_top=3)
    rc = orc.run(1)
    assert rc == oracle.ORC_ERR_OVERFLOW_CAP
    solver.runTimesteps(1, synchronize=False)
    with pytest.raises(lmx.OverflowException) as ei:
        solver.synchronize()
    assert ei.value.code == oracle.TO_LMX[rc]


def test_resume_from_timestep(gpu):
    spec = models.random_model(shape=(9, 9, 9), P=8, nS=3, nST=2, R=4, density=0.2, seed=8)
    a = lmx.MpdRdmeSolver(0)
    spec.load_into(a, seed=4)
    a.runTimesteps(7)
    mid = a.getLatticeNative()
    a.runTimesteps(5)
    b = lmx.MpdRdmeSolver(0)
    spec.load_into(b, seed=4)
    b.setLatticeNative(mid)
    b.setTimestep(7)
    b.runTimesteps(5)
    np.testing.assert_array_equal(a.getLatticeNative(), b.getLatticeNative())


ALIGNED = [
    ((8, 6, 5), 8, "periodic", 0.3),
    ((12, 7, 9), 8, "reflecting", 0.3),
    ((4, 4, 4), 8, "periodic", 0.2),
    ((16, 1, 3), 16, "periodic", 0.25),
    ((32, 12, 20), 4, "reflecting", 0.3),
    ((64, 16, 16), 8, "periodic", 0.05),
    ((8, 8, 16), 8, "reflecting", 0.3),
    ((12, 24, 8), 16, "periodic", 0.2),
    # sparse enough that no wave lists more than a shard holds: the reaction
    # phase evaluates the listed sites only
    ((64, 16, 16), 8, "periodic", 0.01),
    ((32, 24, 20), 16, "reflecting", 0.004),
    ((256, 20, 6), 8, "periodic", 0.01),
    # rows longer than one x chunk of the tile sweeps (8 groups in the
    # default tiles, 64 in the x sweep's flag chunks): several chunks per
    # row, a partial last chunk, whole-chunk flag waves at X = 512
    ((300, 6, 5), 8, "periodic", 0.3),
    ((512, 4, 6), 16, "reflecting", 0.2),
    ((1028, 3, 4), 8, "periodic", 0.05),
]


@pytest.mark.parametrize("case", ALIGNED, ids=[f"{c[0]}-P{c[1]}-{c[2]}" for c in ALIGNED])
def test_vector_kernels_match_generic_and_oracle(gpu, case):
    """X % 4 == 0 selects the 4-sites-per-thread kernels; every y/z kernel
    (dense and compacted tiles, each with the reaction list filled by the z
    sweep (the default) and with the full reaction pass, register columns,
    the march for every segment) must agree with the one-site-per-thread
    kernels and the oracle.  The dense cases overflow the reaction list's
    shards, so its full-pass fallback runs there."""
    shape, P, boundary, density = case
    spec = models.random_model(shape=shape, P=P, nS=5, nST=3, R=7, density=density, seed=17, boundary=boundary)
    orc = oracle.Oracle(spec, seed_top=5, replicate=2)
    assert orc.run(6) == 0
    ref = orc.native()
    for opt in ({}, {"generic": 1}, {"sparse": 1}, {"sparse": 1, "nofuse": 1}, {"nofuse": 1}, {"col": 4}, {"col": 8},
                {"seg": 1}, {"seg": 3}, {"seg": 7}, {"seg": 32}, {"seg": 100000}):
        s = lmx.MpdRdmeSolver(0)
        spec.load_into(s, replicate=2, seed=5)
        if "generic" in opt:
            s.setLaunchOption(lmx._lib.OPT_FORCE_GENERIC, 1)
        if "sparse" in opt:
            s.setLaunchOption(lmx._lib.OPT_YZ_KERNEL, 2)
        if "nofuse" in opt:
            s.setLaunchOption(lmx._lib.OPT_FUSE_REACTION, 0)
        elif "col" in opt:
            s.setLaunchOption(lmx._lib.OPT_YZ_KERNEL, 3)
            s.setLaunchOption(lmx._lib.OPT_MARCH_SEGMENT, opt["col"])
        elif "seg" in opt:
            s.setLaunchOption(lmx._lib.OPT_YZ_KERNEL, 1)
            s.setLaunchOption(lmx._lib.OPT_MARCH_SEGMENT, min(opt["seg"], 65536))
        s.runTimesteps(6)
        got = s.getLatticeNative()
        assert np.array_equal(got, ref), f"{opt}: {(got != ref).sum()} words differ"
        s.close()


FUSED = [
    # shape, P, boundary, density, R
    ((8, 6, 5), 8, "periodic", 0.3, 7),
    ((12, 7, 9), 8, "reflecting", 0.3, 7),
    ((4, 4, 4), 8, "periodic", 0.2, 7),
    ((4, 33, 3), 4, "reflecting", 0.3, 5),     # one row per lane, several y tiles
    ((16, 1, 3), 16, "periodic", 0.25, 7),    # Y = 1: both y neighbours are the row itself
    ((12, 5, 1), 16, "periodic", 0.1, 7),     # Z = 1
    ((12, 5, 1), 16, "reflecting", 0.1, 7),
    ((12, 5, 1), 8, "periodic", 0.3, 7),      # overflow search leaves a one-plane lattice: error parity
    ((32, 40, 20), 4, "reflecting", 0.3, 6),  # rows straddle several tiles
    ((64, 70, 9), 8, "periodic", 0.05, 9),
    ((256, 20, 6), 8, "periodic", 0.05, 9),   # one row per wave
    ((256, 9, 5), 16, "reflecting", 0.1, 9),
    ((12, 10, 8), 8, "periodic", 0.45, 0),    # overflow heavy
]


@pytest.mark.parametrize("case", FUSED, ids=[f"{c[0]}-P{c[1]}-{c[2]}" for c in FUSED])
def test_fused_step_matches_oracle(gpu, case):
    """The fused diffuse+react kernel must give the same lattices, overflow
    statistics and errors as the separate kernels, for every z segmenting."""
    shape, P, boundary, density, R = case
    spec = models.random_model(shape=shape, P=P, nS=5, nST=3, R=R, density=density, seed=23, boundary=boundary,
                               qmax=0.5)
    orc = oracle.Oracle(spec, seed_top=6, replicate=1)
    refs, err = [], 0
    for _ in range(3):
        err = orc.run(3)
        if err:
            break
        refs.append(orc.native())
    for zseg in (0, 1, 2, 3, 1000):
        s = lmx.MpdRdmeSolver(0)
        spec.load_into(s, replicate=1, seed=6)
        s.setPath(lmx._lib.PATH_FUSED)
        s.setLaunchOption(lmx._lib.OPT_FUSED_ZSEG, zseg)
        for c, ref in enumerate(refs):
            s.runTimesteps(3)
            got = s.getLatticeNative()
            assert np.array_equal(got, ref), f"zseg={zseg} chunk {c}: {(got != ref).sum()} words differ"
        if err:
            s.runTimesteps(3, synchronize=False)
            with pytest.raises(lmx.LmxError) as ei:
                s.synchronize()
            assert ei.value.code == oracle.TO_LMX[err], f"zseg={zseg}"
        else:
            e, st = s.getOverflowStats()
            assert (e, st) == (orc.stats.overflow_entries, orc.stats.overflow_steps), f"zseg={zseg}"
        s.close()


def test_fused_whole_cell_reduced(gpu):
    spec = models.whole_cell(shape=(64, 64, 32), n_particles=30000, seed=2)
    run_pair(spec, steps=6, seed=3, chunks=2, path=lmx._lib.PATH_FUSED)


@pytest.mark.parametrize("graph,reaction_list,tile_threads", [(True, True, 0), (False, True, 0), (True, False, 0),
                                                               (True, True, 512), (True, True, 1024)],
                         ids=["graph-list", "nograph-list", "graph-fullpass", "tiles512", "tiles1024"])
def test_ensemble_members_bit_exact(gpu, graph, reaction_list, tile_threads):
    """A batch of uneven lattices, models and seeds: after every run each
    member equals its own oracle run (and so its plain run), with and
    without the captured graph, also after members leave and join.  Members
    without zero-order reactions run their reaction phase over the sites
    their z sweep lists (or, with the list off, the full pass); the others
    always the full pass."""
    shapes = [(8, 6, 5), (12, 7, 9), (32, 12, 20), (16, 1, 3), (4, 33, 3), (64, 16, 16), (300, 4, 6), (512, 3, 4)]
    ens = lmx.Ensemble(0, graph=graph)
    ens.set_reaction_list(reaction_list)
    ens.set_tile_threads(tile_threads)
    pairs = []
    for k, shape in enumerate(shapes):
        spec = models.random_model(shape=shape, P=8, nS=5, nST=3, R=7, density=0.25, seed=40 + k,
                                   boundary="reflecting" if k % 2 else "periodic",
                                   include_zero_order=k % 3 == 2)
        s = lmx.MpdRdmeSolver(0)
        spec.load_into(s, replicate=k, seed=9 + k)
        orc = oracle.Oracle(spec, seed_top=9 + k, replicate=k)
        ens.add(s)
        pairs.append((s, orc))
    for chunk in (1, 4, 3):
        ens.run(chunk)
        for s, orc in pairs:
            assert orc.run(chunk) == 0
            assert np.array_equal(s.getLatticeNative(), orc.native()), f"graph={graph} {s.shape}"
    # a member leaves, steps alone, and rejoins
    s0, o0 = pairs[0]
    ens.remove(s0)
    s0.runTimesteps(2)
    assert o0.run(2) == 0
    ens.add(s0)
    ens.run(2)
    for s, orc in pairs:
        assert orc.run(2) == 0
        assert np.array_equal(s.getLatticeNative(), orc.native())
        assert s.timestep == orc.timestep
    ens.close()


def test_ensemble_grows_past_block_map(gpu):
    """A member joining after the batch has run raises the launches' block
    count past the block -> member map the batch first allocated (4096
    entries; a 256x256x128 lattice alone has 8192 x-sweep blocks): the map
    is reallocated, the captured step recaptured, and every member still
    equals its oracle run."""
    ens = lmx.Ensemble(0)
    pairs = []

    def join(k, shape, density):
        spec = models.random_model(shape=shape, P=8, nS=5, nST=3, R=7, density=density, seed=70 + k)
        s = lmx.MpdRdmeSolver(0)
        spec.load_into(s, replicate=k, seed=5 + k)
        ens.add(s)
        pairs.append((s, oracle.Oracle(spec, seed_top=5 + k, replicate=k)))

    join(0, (32, 12, 20), 0.25)
    join(1, (64, 16, 16), 0.25)
    ens.run(2)
    for s, orc in pairs:
        assert orc.run(2) == 0
    join(2, (256, 256, 128), 0.002)
    join(3, (12, 7, 9), 0.25)
    for chunk in (2, 1):
        ens.run(chunk)
        for s, orc in pairs:
            assert orc.run(chunk) == 0
            assert np.array_equal(s.getLatticeNative(), orc.native()), f"{s.shape}"
            assert s.timestep == orc.timestep
    ens.close()


def test_ensemble_autotune_keeps_state_and_results(gpu):
    """The batch autotuner runs every launch configuration on the members'
    own state and restores it: afterwards each member's lattice, step counter
    and overflow statistics are those before the call, and the steps that
    follow (with the configuration it kept) equal the oracle's."""
    shapes = [(16, 12, 10), (32, 8, 6), (64, 16, 16), (12, 7, 9)]
    ens = lmx.Ensemble(0)
    pairs = []
    for k, shape in enumerate(shapes):
        spec = models.random_model(shape=shape, P=8, nS=5, nST=3, R=7, density=0.4, seed=70 + k,
                                   include_zero_order=k == 3)
        s = lmx.MpdRdmeSolver(0)
        spec.load_into(s, replicate=k, seed=3 + k)
        ens.add(s)
        pairs.append((s, oracle.Oracle(spec, seed_top=3 + k, replicate=k)))
    ens.run(3)
    before = [(s.getLatticeNative(), s.timestep, s.getOverflowStats()) for s, _ in pairs]
    tuned = ens.autotune(steps=4)
    assert tuned["tile_threads"] in (256, 512, 1024) and tuned["ms_per_step"] > 0
    for (s, _), (lat, t, ovf) in zip(pairs, before):
        assert np.array_equal(s.getLatticeNative(), lat)
        assert s.timestep == t and s.getOverflowStats() == ovf
    ens.run(5)
    for s, orc in pairs:
        assert orc.run(8) == 0
        assert np.array_equal(s.getLatticeNative(), orc.native()), f"{s.shape}"
        assert s.getOverflowStats() == (orc.stats.overflow_entries, orc.stats.overflow_steps)
    ens.close()


def test_served_jobs_equal_their_solo_runs(gpu):
    """Continuous batching: jobs that arrive while others run are admitted
    into the running batch (under a tight HBM budget some wait), and each
    job's final lattice equals its solo run, bit for bit."""
    from lattice_microbes_amd import serving
    templates = [models.bimolecular(shape=(32, 32, 32), nA=300, nB=300, kf=1e9, seed=3),
                 models.random_model(shape=(16, 12, 10), P=8, nS=5, nST=3, R=7, density=0.2, seed=5),
                 models.capsule_minde(shape=(32, 32, 48), counts=(300, 300, 100), seed=2)]
    jobs = serving.poisson_jobs(templates, n=12, rate_hz=200.0, steps_range=(5, 60), seed=4)
    budget = 3 * max(serving.footprint(t) for t in templates)
    stats = serving.Server(0, hbm_budget=budget, chunk=7).run(jobs)
    assert stats["jobs"] == 12 and all(j.done_steps == j.steps for j in jobs)
    assert max(j.admitted - j.arrival for j in jobs) >= 0.0
    for j in jobs:
        assert j.digest == serving.run_alone(j), f"job {j.id} differs from its solo run"


def _reverse_slots(native, P, words):
    """Every site's particles moved to its last slots in reverse order: an
    unpacked lattice (empty slots before occupied ones), as a caller may write
    it through the native setter or device pointers."""
    W = native.shape[0]
    flat = native.reshape(W, -1)
    if words:
        slots = flat.T.copy()                                   # (N, P)
    else:
        slots = flat.T.copy().view(np.uint8).reshape(-1, P)     # (N, P): slot 4w+b
    rev = slots[:, ::-1].copy()
    if words:
        out = rev.T
    else:
        out = rev.reshape(-1, W, 4).view(np.uint32).reshape(-1, W).T
    return np.ascontiguousarray(out).reshape(native.shape)


@pytest.mark.parametrize("words", [False, True], ids=["byte", "uint32"])
def test_unpacked_lattice_matches_oracle(gpu, words):
    """A lattice with empty slots before occupied ones (set through the
    native setter) is read slot by slot in the first step, as the oracle reads
    it; the steps after it see packed sites again."""
    spec = models.random_model(shape=(16, 12, 10), P=8, nS=4, nST=2, R=5, density=0.3, seed=8)
    if words:
        spec = dataclasses.replace(spec, lattice=spec.lattice.astype(np.uint32))
        solver = lmx.IntMpdRdmeSolver(0)
    else:
        solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=4)
    orc = oracle.Oracle(spec, seed_top=4)
    unpacked = _reverse_slots(solver.getLatticeNative(), spec.P, words)
    assert not np.array_equal(unpacked, solver.getLatticeNative())
    solver.setLatticeNative(unpacked)
    orc.lattice[:] = unpacked.reshape(-1)
    for steps in (1, 1, 3):
        solver.runTimesteps(steps)
        assert orc.run(steps) == 0
        np.testing.assert_array_equal(solver.getLatticeNative(), orc.native())
    solver.close()


def test_unpacked_lattice_through_device_pointers(gpu):
    """The same through the device-hook path: a particle written on the device
    into the last slot of an empty site is read in the next step."""
    import torch

    if not torch.cuda.is_initialized():
        pytest.skip("torch must start before the library in this process (run with -m gpu)")
    from lattice_microbes_amd.solver import DeviceHookLattice

    spec = models.random_model(shape=(16, 12, 10), P=8, nS=4, nST=2, R=5, density=0.3, seed=9)
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=6)
    orc = oracle.Oracle(spec, seed_top=6)
    solver.runTimesteps(2)
    assert orc.run(2) == 0
    nat = solver.getLatticeNative()
    empty = np.argwhere((nat == 0).all(axis=0))[:20]            # (z, y, x) of empty sites
    dl = DeviceHookLattice(solver)
    dl.refresh()
    view = dl.getParticleLatticeView()
    for z, y, x in empty:
        view[1, z, y, x, 3] = 2                                # slot 7 of 8
    dl.synchronize()
    ov = orc.lattice.reshape(2, *nat.shape[1:])
    for z, y, x in empty:
        ov[1, z, y, x] = np.uint32(2) << 24
    for steps in (1, 2):
        solver.runTimesteps(steps)
        assert orc.run(steps) == 0
        np.testing.assert_array_equal(solver.getLatticeNative(), orc.native())
    solver.close()


@pytest.mark.parametrize("words", [False, True], ids=["byte", "uint32"])
def test_upper_planes_filling(gpu, words):
    """Upper word planes start empty (their reads skipped on packed sites)
    and fill as sites crowd: each step equals the oracle's."""
    spec = models.random_model(shape=(12, 10, 8), P=8, nS=3, nST=1, R=0, density=0.45, seed=12, qmax=0.5)
    lat = spec.lattice.copy()
    if words:
        lat[..., 1:] = 0        # only slot 0: planes 1..7 start empty
        spec = dataclasses.replace(spec, lattice=lat.astype(np.uint32))
        solver = lmx.IntMpdRdmeSolver(0)
    else:
        lat[..., 4:] = 0        # slots 0..3: plane 1 starts empty
        spec = dataclasses.replace(spec, lattice=lat)
        solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=21)
    orc = oracle.Oracle(spec, seed_top=21)
    W = spec.P if words else spec.P // 4
    assert not solver.getLatticeNative()[1:].any()
    filled = False
    for _ in range(12):
        solver.runTimesteps(1)
        assert orc.run(1) == 0
        got = solver.getLatticeNative()
        np.testing.assert_array_equal(got, orc.native())
        filled |= bool(got[1:].any())
    assert filled and W > 1
    solver.close()


@pytest.mark.parametrize("other", ["march", "fused", "generic"])
def test_option_switches_mid_run(gpu, other):
    """Steps by the default kernels (which keep the upper planes' chunk
    flags), then by kernels that do not, then by the default ones again: the
    flags left stale in between are reset, and every phase equals the
    oracle."""
    spec = models.random_model(shape=(64, 8, 8), P=8, nS=3, nST=2, R=4, density=0.4, seed=14, qmax=0.5)
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=17)
    orc = oracle.Oracle(spec, seed_top=17)

    def phase(n):
        solver.runTimesteps(n)
        assert orc.run(n) == 0
        np.testing.assert_array_equal(solver.getLatticeNative(), orc.native())

    phase(3)
    if other == "march":
        solver.setLaunchOption(lmx._lib.OPT_YZ_KERNEL, 1)
    elif other == "fused":
        solver.setPath(lmx._lib.PATH_FUSED)
    else:
        solver.setLaunchOption(lmx._lib.OPT_FORCE_GENERIC, 1)
    phase(3)
    solver.setLaunchOption(lmx._lib.OPT_YZ_KERNEL, 0)
    solver.setPath(lmx._lib.PATH_SWEEPS)
    solver.setLaunchOption(lmx._lib.OPT_FORCE_GENERIC, 0)
    phase(4)
    assert solver.getLatticeNative()[1].any()    # the second plane was in use
    solver.close()


def test_whole_cell_full_size_matches_oracle(gpu):
    """The default kernels at the whole-cell shape (256x256x512, P=8, 100
    species, 300 reactions: whole-chunk waves keeping the plane flags, the
    XCD block order and the reaction-list shards at their real scale) equal
    the oracle word for word after 2 steps (the oracle takes seconds per
    step at this size)."""
    spec = models.whole_cell(seed=1)
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, seed=3)
    solver.runTimesteps(2)
    orc = oracle.Oracle(spec, seed_top=3)
    assert orc.run(2) == 0
    np.testing.assert_array_equal(solver.getLatticeNative(), orc.native())
    np.testing.assert_array_equal(solver.getSpeciesCounts(), orc.species_counts())
    solver.close()


def test_whole_cell_full_size_paths_agree(gpu):
    """At the whole-cell shape the fused step, the marching y/z sweep and the
    one-site-per-thread kernels end 6 steps on the default kernels' state."""
    spec = models.whole_cell(seed=2)
    states = []
    for variant in ("default", "fused", "march", "generic"):
        s = lmx.MpdRdmeSolver(0)
        spec.load_into(s, seed=5)
        if variant == "fused":
            s.setPath(lmx._lib.PATH_FUSED)
        elif variant == "march":
            s.setLaunchOption(lmx._lib.OPT_YZ_KERNEL, 1)
        elif variant == "generic":
            s.setLaunchOption(lmx._lib.OPT_FORCE_GENERIC, 1)
        s.runTimesteps(6)
        states.append(s.getLatticeNative())
        s.close()
    for k in range(1, len(states)):
        np.testing.assert_array_equal(states[k], states[0])
