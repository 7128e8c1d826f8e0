"""GPU parity on the benchmark configurations' own workloads.

* the capsule cell at its full 128x128x256 size against the oracle;
* the batched ensemble as benchmarked: 256 bimolecular replicates on lattices
  drawn from 32^3 ... 128^3, every member equal to its plain single-solver
  run after 50 steps, and one member of each size equal to the oracle;
* the serving mix of the three request templates (bimolecular 64^3, capsule
  128x128x256, whole-cell 256x256x128 with 30 species) with short step
  counts: every job equal to its solo run, the first job of each template
  equal to the oracle;
* a long horizon: 250 steps of a crowded whole-cell-style 64^3 model against
  the oracle every 50 steps, so that what spans steps (the chunk flags of the
  upper word plane, the reaction-list shards and their full-pass fallback,
  the overflow list and its resolution) is checked over many steps.

Integer state: the bar is exact equality.
"""
import hashlib

import numpy as np
import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd import models, serving
import oracle

pytestmark = pytest.mark.gpu


def _digest(native):
    return hashlib.sha256(np.ascontiguousarray(native, dtype=np.uint32).tobytes()).hexdigest()


def test_capsule_full_size_matches_oracle(gpu):
    spec = models.capsule_minde(seed=1)
    assert spec.shape == (128, 128, 256)
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, replicate=2, seed=17)
    orc = oracle.Oracle(spec, seed_top=17, replicate=2)
    for _ in range(3):
        solver.runTimesteps(1)
        assert orc.run(1) == 0
        gpu_lat = solver.getLatticeNative()
        ref = orc.native()
        assert np.array_equal(gpu_lat, ref), f"step {orc.timestep}: {(gpu_lat != ref).sum()} words differ"
    np.testing.assert_array_equal(solver.getSpeciesCounts(), orc.species_counts())
    solver.close()


def test_ensemble_256_replicates_match_plain_runs(gpu):
    steps = 50
    mix = models.ensemble_mix(256, seed=1)
    ens = lmx.Ensemble(0)
    members = []
    for spec, r, seed in mix:
        s = lmx.MpdRdmeSolver(0)
        spec.load_into(s, replicate=r, seed=seed)
        ens.add(s)
        members.append(s)
    ens.run(steps)
    batched = [_digest(s.getLatticeNative()) for s in members]
    ens.close()
    for s in members:
        s.close()
    # every replicate run alone, one after another
    for (spec, r, seed), d in zip(mix, batched):
        s = lmx.MpdRdmeSolver(0)
        spec.load_into(s, replicate=r, seed=seed)
        s.runTimesteps(steps)
        assert _digest(s.getLatticeNative()) == d, f"replicate {r} {spec.shape}: batched != plain"
        s.close()
    # one replicate of each lattice size against the oracle
    first = {}
    for k, (spec, r, seed) in enumerate(mix):
        first.setdefault(spec.shape, k)
    assert len(first) == len(models.ENSEMBLE_SIZES)
    for shape, k in sorted(first.items()):
        spec, r, seed = mix[k]
        orc = oracle.Oracle(spec, seed_top=seed, replicate=r)
        assert orc.run(steps) == 0
        assert _digest(orc.native()) == batched[k], f"replicate {r} {shape}: batched != oracle"


@pytest.mark.parametrize("policy", ["fcfs", "sjf"])
def test_served_template_mix_matches_solo_and_oracle(gpu, policy):
    """The served mix of the three templates: every job equals its solo run
    and one job per template its oracle run, in arrival order and with the
    size-aware admission (smallest work first under a site limit) that the
    benchmark serves with."""
    templates = models.serving_templates(seed=1)
    jobs = serving.poisson_jobs(templates, 9, 20.0, (4, 12), seed=5)
    # every template is requested
    assert {id(j.spec) for j in jobs} == {id(t) for t in templates}
    # a budget that holds two of the largest jobs at once: some requests wait
    budget = 2 * serving.footprint(templates[2]) + serving.footprint(templates[1])
    limit = None if policy == "fcfs" else templates[2].N + templates[1].N
    serving.Server(0, hbm_budget=budget, chunk=3, max_sites=limit, policy=policy).run(jobs)
    for j in jobs:
        assert j.digest == serving.run_alone(j), f"job {j.id} ({j.spec.name}): served != solo"
    seen = set()
    for j in jobs:
        if j.spec.name in seen:
            continue
        seen.add(j.spec.name)
        orc = oracle.Oracle(j.spec, seed_top=j.seed, replicate=j.replicate)
        assert orc.run(j.steps) == 0
        assert _digest(orc.native()) == j.digest, f"job {j.id} ({j.spec.name}): served != oracle"
        np.testing.assert_array_equal(j.counts, orc.species_counts())
    assert len(seen) == 3


def crowded_whole_cell_64(seed=7):
    """The whole-cell model's layout at 64^3 (10 site types, obstacles, 30
    species, 90 reactions), 17% of fluid sites occupied, and its five
    spherical compartments packed with 6 particles per site, so that sites
    overflow and the upper word plane fills where the compartments drain."""
    spec = models.whole_cell(shape=(64, 64, 64), nS=30, R=90, n_particles=30_000, seed=seed)
    rng = np.random.default_rng(seed + 1)
    lat = spec.lattice.copy()
    comp = np.argwhere(spec.sites >= 5)
    for x, y, z in comp:
        site = lat[x, y, z]
        n = int(np.count_nonzero(site))
        add = max(0, 6 - n)
        site[n:n + add] = rng.integers(1, spec.nS + 1, size=add)
    spec.lattice = lat
    return spec


def test_long_horizon_crowded_matches_oracle(gpu):
    spec = crowded_whole_cell_64()
    solver = lmx.MpdRdmeSolver(0)
    spec.load_into(solver, replicate=3, seed=29)
    orc = oracle.Oracle(spec, seed_top=29, replicate=3)
    for chunk in range(5):
        solver.runTimesteps(50)
        assert orc.run(50) == 0
        gpu_lat = solver.getLatticeNative()
        ref = orc.native()
        assert np.array_equal(gpu_lat, ref), f"step {orc.timestep}: {(gpu_lat != ref).sum()} words differ"
        np.testing.assert_array_equal(solver.getSpeciesCounts(), orc.species_counts())
        entries, steps_with = solver.getOverflowStats()
        assert entries == orc.stats.overflow_entries and steps_with == orc.stats.overflow_steps
    assert orc.timestep == 250
    assert orc.stats.overflow_entries > 0, "the crowded compartments should overflow"
    solver.close()
