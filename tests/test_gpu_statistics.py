"""Statistical agreement of the MPD-RDME ensemble with the CPU solvers.

The integer state is checked bit for bit elsewhere; these tests check that
the dynamics are the right ones: species-count means and variances of a GPU
ensemble (one batched Ensemble of independent replicates) against

  * the exact stationary distribution of the well-mixed chemical master
    equation for A + B <-> C (detailed balance; the RDME's total-count
    stationary distribution is the same), and
  * an ensemble of the next-subvolume CPU counterpart (exact RDME SSA) at a
    fixed time on the same lattice and model,

each with a 99% confidence interval built from independent replicates (a
replicate's time average is one sample, so autocorrelation within a replicate
cannot narrow the interval).
"""
import ctypes
import math

import numpy as np
import pytest

import lattice_microbes_amd as lmx
import oracle
from helpers import q_to_D, spec, TAU

pytestmark = pytest.mark.gpu

Z99 = 2.5758
A0 = 30
L = 8


def bimolecular_case(a_pair=0.03, kr_step=1e-3, q=0.2, seed=5):
    """A + B <-> C on an 8^3 periodic lattice: a_pair is the per-step reaction
    propensity of one co-located A-B pair (K * tau * N), kr_step the per-step
    dissociation propensity of one C."""
    N = L ** 3
    K = a_pair / (TAU * N)
    kr = kr_step / TAU
    rx = [(2, K, {0: -1, 1: -1, 2: 1}, [0, 1]), (1, kr, {0: 1, 1: 1, 2: -1}, [2])]
    rng = np.random.default_rng(seed)
    pos = rng.integers(0, L, size=(2 * A0, 3))
    parts = [(int(x), int(y), int(z), 1 + (k >= A0)) for k, (x, y, z) in enumerate(pos)]
    s = spec((L, L, L), P=8, nS=3, DF=q_to_D([q, q, q]).reshape(1, 1, 3), reactions=rx, particles=parts)
    return s, K, kr


def exact_stationary(K, kr):
    w = np.array([(K / kr) ** c / (math.factorial(A0 - c) ** 2 * math.factorial(c)) for c in range(A0 + 1)])
    pi = w / w.sum()
    c = np.arange(A0 + 1)
    mu = float((pi * c).sum())
    return mu, float((pi * c * c).sum() - mu * mu)


def ensemble(s, replicates, seed0):
    ens = lmx.Ensemble(0)
    solvers = []
    for r in range(replicates):
        sv = lmx.MpdRdmeSolver(0)
        s.load_into(sv, replicate=r, seed=seed0 + r)
        ens.add(sv)
        solvers.append(sv)
    return ens, solvers


def test_stationary_distribution_matches_cme():
    s, K, kr = bimolecular_case()
    mu, var = exact_stationary(K, kr)
    R, burn, every, nsamp = 256, 4000, 100, 32
    ens, solvers = ensemble(s, R, 100)
    ens.run(burn)  # ~12 relaxation times from C = 0
    samp = np.zeros((R, nsamp))
    for k in range(nsamp):
        ens.run(every)
        samp[:, k] = [int(sv.getSpeciesCounts()[2]) for sv in solvers]
    # conservation: A + C and B + C stay A0 in every replicate
    for sv in solvers[:8]:
        a, b, c = (int(v) for v in sv.getSpeciesCounts())
        assert a + c == A0 and b + c == A0
    ens.close()
    rmean, rsq = samp.mean(axis=1), (samp ** 2).mean(axis=1)
    gmean = rmean.mean()
    se = rmean.std(ddof=1) / math.sqrt(R)
    assert abs(gmean - mu) < Z99 * se, f"mean {gmean:.4f} vs exact {mu:.4f} (99% half-width {Z99 * se:.4f})"
    # variance: 99% bootstrap interval over replicates
    rng = np.random.default_rng(0)
    boot = []
    for _ in range(2000):
        idx = rng.integers(0, R, R)
        m = rmean[idx].mean()
        boot.append(rsq[idx].mean() - m * m)
    lo, hi = np.percentile(boot, [0.5, 99.5])
    assert lo <= var <= hi, f"variance interval [{lo:.4f}, {hi:.4f}] misses exact {var:.4f}"


def test_transient_mean_matches_next_subvolume_ensemble():
    """Mean and variance of C after 200 steps from the same initial lattice:
    MPD on the GPU against the exact RDME (next-subvolume method) on the CPU."""
    s, K, kr = bimolecular_case()
    steps, R_mpd, R_nsm = 200, 512, 256
    ens, solvers = ensemble(s, R_mpd, 300)
    ens.run(steps)
    c_mpd = np.array([int(sv.getSpeciesCounts()[2]) for sv in solvers], dtype=np.float64)
    ens.close()

    lib = oracle.load()
    S = np.ascontiguousarray(s.S, dtype=np.int32).reshape(-1)
    DF = np.ascontiguousarray(s.DF, dtype=np.float64).reshape(-1)
    c_nsm = np.zeros(R_nsm)
    counts = np.zeros(3, np.uint32)
    for r in range(R_nsm):
        o = oracle.Oracle(s, seed_top=7, replicate=r)
        h = lib.orc_nsm_create(L, L, L, s.P, s.nS, s.nST, s.R, oracle._p(o.types, ctypes.c_uint32),
                               oracle._p(o.k, ctypes.c_double), oracle._p(o.si, ctypes.c_uint32),
                               oracle._p(o.s2i, ctypes.c_uint32), oracle._p(S, ctypes.c_int32),
                               oracle._p(DF, ctypes.c_double), oracle._p(o.RL32, ctypes.c_uint32), s.spacing,
                               oracle._p(o.lattice, ctypes.c_uint32), oracle._p(o.sites, ctypes.c_uint8), 11, r + 1)
        assert h
        assert lib.orc_nsm_run(h, steps * s.tau, 1 << 62, 0.0, 0, None, None) == 0
        lib.orc_nsm_counts(h, oracle._p(counts, ctypes.c_uint32))
        lib.orc_nsm_destroy(h)
        c_nsm[r] = counts[2]
    diff = c_mpd.mean() - c_nsm.mean()
    se = math.sqrt(c_mpd.var(ddof=1) / R_mpd + c_nsm.var(ddof=1) / R_nsm)
    assert abs(diff) < Z99 * se, f"MPD {c_mpd.mean():.3f} vs NSM {c_nsm.mean():.3f} (99% half-width {Z99 * se:.3f})"
    assert c_mpd.mean() > 3.0  # the reaction has progressed: the comparison is not between two zeros
    # variance ratio within the 99% F interval
    from scipy import stats
    f = c_mpd.var(ddof=1) / c_nsm.var(ddof=1)
    lo, hi = stats.f.ppf([0.005, 0.995], R_mpd - 1, R_nsm - 1)
    assert lo <= f <= hi, f"variance ratio {f:.3f} outside [{lo:.3f}, {hi:.3f}]"


def test_tracked_particles_diffuse_as_mpd_random_walk(gpu):
    """Every particle has its own species id (the uint32 lattice), so each
    path is tracked.  Inside a periodic lattice of one site type a sweep moves
    a particle by -1 or +1 along its axis with probability q each
    (makeDiffusionChoices, byte_diffusion_1d_dev.cu:86-121), one sweep per
    axis per step, so after K steps each axis displacement has mean 0 and
    variance 2qK, independently per axis.  Nine checks (three means, three
    variances, three correlations), each at 99.9% (normal for means and
    correlations, chi-square for variances), so 99% jointly; a 10% error in q
    fails one of them with probability ~0.9, a 10% directional bias always."""
    from scipy import stats

    from lattice_microbes_amd.models import ModelSpec
    from helpers import SPACING

    L3, n, q, K = 96, 3000, 0.2, 60
    z999 = stats.norm.ppf(0.9995)
    rng = np.random.default_rng(17)
    grid = np.stack(np.meshgrid(*(np.arange(0, L3, 6),) * 3, indexing="ij"), -1).reshape(-1, 3)
    start = grid[rng.choice(len(grid), n, replace=False)]
    P = 8
    lattice = np.zeros((L3, L3, L3, P), np.uint32)
    lattice[start[:, 0], start[:, 1], start[:, 2], 0] = np.arange(1, n + 1, dtype=np.uint32)
    s = ModelSpec("tracked", (L3, L3, L3), SPACING, TAU, P, n, 1, np.zeros(0, np.uint32), np.zeros(0),
                  np.zeros((n, 0), np.int32), np.zeros((n, 0), np.uint32), q_to_D(np.full((1, 1, n), q)),
                  np.zeros((0, 1), np.uint32), lattice, np.zeros((L3, L3, L3), np.uint8), "periodic")
    w = lmx.IntMpdRdmeSolver(0)
    try:
        s.load_into(w, seed=29)
        w.runTimesteps(K)
        lat = w.getLattice()
    finally:
        w.close()
    x, y, z, sl = np.nonzero(lat)
    ids = lat[x, y, z, sl].astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(1, n + 1))   # every particle once
    end = np.zeros((n, 3), np.int64)
    end[ids - 1] = np.stack([x, y, z], -1)
    d = (end - start + L3 // 2) % L3 - L3 // 2                  # periodic displacement
    var = 2 * q * K
    for a in range(3):
        m = d[:, a].mean()
        assert abs(m) < z999 * math.sqrt(var / n), (a, m)
        s2 = d[:, a].var(ddof=1)
        lo, hi = stats.chi2.ppf([0.0005, 0.9995], n - 1) * var / (n - 1)
        assert lo < s2 < hi, (a, s2, lo, hi)
    c = np.corrcoef(d.T)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert abs(c[a, b]) < z999 / math.sqrt(n), (a, b, c[a, b])


def _nsm_counts(s, steps, replicates, seed_top):
    """Species counts at steps * tau of next-subvolume replicates of model s
    (the exact RDME), run in threads over the host cores: the oracle releases
    the interpreter while it runs."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    lib = oracle.load()
    S = np.ascontiguousarray(s.S, dtype=np.int32).reshape(-1)
    DF = np.ascontiguousarray(s.DF, dtype=np.float64).reshape(-1)
    o = oracle.Oracle(s, seed_top=seed_top, replicate=0)
    X, Y, Z = s.shape

    def one(r):
        h = lib.orc_nsm_create(X, Y, Z, s.P, s.nS, s.nST, s.R, oracle._p(o.types, ctypes.c_uint32),
                               oracle._p(o.k, ctypes.c_double), oracle._p(o.si, ctypes.c_uint32),
                               oracle._p(o.s2i, ctypes.c_uint32), oracle._p(S, ctypes.c_int32),
                               oracle._p(DF, ctypes.c_double), oracle._p(o.RL32, ctypes.c_uint32), s.spacing,
                               oracle._p(o.lattice, ctypes.c_uint32), oracle._p(o.sites, ctypes.c_uint8),
                               seed_top, r + 1)
        assert h
        rc = lib.orc_nsm_run(h, steps * s.tau, 1 << 62, 0.0, 0, None, None)
        counts = np.zeros(s.nS, np.uint32)
        lib.orc_nsm_counts(h, oracle._p(counts, ctypes.c_uint32))
        lib.orc_nsm_destroy(h)
        assert rc == 0
        return counts

    workers = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(workers) as ex:
        return np.array(list(ex.map(one, range(replicates))), dtype=np.float64)


def _cme_bimolecular(K, kr, n0, t):
    """Distribution of C at time t for the well-mixed A + B <-> C chemical
    master equation from A = B = n0, C = 0: states c = 0..n0, association
    propensity K (n0-c)^2, dissociation kr c (a tridiagonal generator),
    p(t) = expm(Q t) p(0)."""
    from scipy.sparse import diags
    from scipy.sparse.linalg import expm_multiply

    c = np.arange(n0 + 1, dtype=np.float64)
    up = K * (n0 - c) ** 2          # c -> c + 1
    down = kr * c                    # c -> c - 1
    Q = diags([-(up + down), up[:-1], down[1:]], [0, -1, 1], format="csc")  # dp/dt = Q p
    p0 = np.zeros(n0 + 1)
    p0[0] = 1.0
    p = expm_multiply(Q * t, p0)
    p = np.clip(p, 0.0, None)
    p /= p.sum()
    mu = float((p * c).sum())
    return mu, float((p * c * c).sum() - mu * mu)


def test_bimolecular_64_matches_nsm_and_cme(gpu):
    """The plain benchmark model itself (64^3 reflecting, lambda 32 nm, tau
    50 us, 1000 A + 1000 B, kf 1e6 /M/s, kr 0.1 /s), 2000 steps = 0.1 s:
    mean and variance of C over 512 MPD replicates against 256
    next-subvolume replicates of the same lattice (exact RDME; 99% Welch and
    F intervals), and against the well-mixed chemical master equation solved
    by matrix exponential (99% normal interval for the mean, 99% bootstrap
    interval for the variance).  The CME is the well-mixed limit: on the
    cubic lattice the diffusion-limited correlation lowers the effective
    association rate by k_sv lambda^2 W / (D_A + D_B) ~ 0.7% (W = 0.2527, the
    lattice constant of the return probability), a quarter of the mean's
    interval here.  A rate-unit slip (a factor N, N_A or 1000 in the
    conversion) or a lost factor 2 moves the mean by far more."""
    from lattice_microbes_amd import models
    from scipy import stats

    s = models.bimolecular(seed=1)
    assert s.shape == (64, 64, 64) and s.boundary == "reflecting"
    steps, R_mpd, R_nsm = 2000, 512, 256
    ens, solvers = ensemble(s, R_mpd, 500)
    ens.run(steps)
    counts = np.array([sv.getSpeciesCounts() for sv in solvers], dtype=np.float64)
    ens.close()
    for sv in solvers:
        sv.close()
    assert np.all(counts[:, 0] + counts[:, 2] == 1000) and np.all(counts[:, 1] + counts[:, 2] == 1000)
    c_mpd = counts[:, 2]
    c_nsm = _nsm_counts(s, steps, R_nsm, 13)[:, 2]
    assert c_mpd.mean() > 10.0  # the association has progressed

    diff = c_mpd.mean() - c_nsm.mean()
    se = math.sqrt(c_mpd.var(ddof=1) / R_mpd + c_nsm.var(ddof=1) / R_nsm)
    assert abs(diff) < Z99 * se, f"MPD {c_mpd.mean():.3f} vs NSM {c_nsm.mean():.3f} (99% half-width {Z99 * se:.3f})"
    f = c_mpd.var(ddof=1) / c_nsm.var(ddof=1)
    lo, hi = stats.f.ppf([0.005, 0.995], R_mpd - 1, R_nsm - 1)
    assert lo <= f <= hi, f"MPD/NSM variance ratio {f:.3f} outside [{lo:.3f}, {hi:.3f}]"

    K = float(s.K[0])      # association, per pair in the whole volume (LM units)
    kr = float(s.K[1])
    mu, var = _cme_bimolecular(K, kr, 1000, steps * s.tau)
    se_m = c_mpd.std(ddof=1) / math.sqrt(R_mpd)
    assert abs(c_mpd.mean() - mu) < Z99 * se_m, \
        f"MPD mean {c_mpd.mean():.3f} vs CME {mu:.3f} (99% half-width {Z99 * se_m:.3f})"
    rng = np.random.default_rng(1)
    boot = [c_mpd[rng.integers(0, R_mpd, R_mpd)].var(ddof=1) for _ in range(2000)]
    vlo, vhi = np.percentile(boot, [0.5, 99.5])
    assert vlo <= var <= vhi, f"MPD variance interval [{vlo:.3f}, {vhi:.3f}] misses the CME's {var:.3f}"
