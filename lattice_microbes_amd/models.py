"""Synthetic, seeded MPD-RDME models in the reference's input format.

Each builder returns a :class:`ModelSpec` holding exactly the arrays the
reference's ``buildModel`` / ``buildDiffusionModel`` receive
(src/rdme/MpdRdmeSolver.h:82-83): reaction types, rate constants K in the
reference's internal units, stoichiometry S and dependency D (species-major),
the diffusion matrix DF[src][dst][species] in m^2/s, the reaction location
matrix RL[reaction][site type], and row-major lattice/site bytes.

Rate constants given in M^-1 s^-1 are converted the way jLM does before it
hands them to the solver (src/jlm/jLM/Types.py:316-339):
k_LM = k * (N_A * V_site)^(1-order) * N_sites^(1-order).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

AVOGADRO = 6.02214076e23

# Reaction type codes of the ReactionModel message (src/cme/CMESolver.h:79-106).
ZEROTH, FIRST, SECOND, SECOND_SELF = 0, 1, 2, 3


@dataclass
class ModelSpec:
    name: str
    shape: Tuple[int, int, int]
    spacing: float
    tau: float
    P: int
    nS: int
    nST: int
    types: np.ndarray            # [R] uint32
    K: np.ndarray                # [R] float64 (kCols = 1)
    S: np.ndarray                # [nS, R] int32
    D: np.ndarray                # [nS, R] uint32
    DF: np.ndarray               # [nST, nST, nS] float64
    RL: np.ndarray               # [R, nST] uint32
    lattice: np.ndarray          # [X, Y, Z, P] uint8 (uint32 for more than 255 species)
    sites: np.ndarray            # [X, Y, Z] uint8
    boundary: str = "periodic"
    species_names: List[str] = field(default_factory=list)
    site_type_names: List[str] = field(default_factory=list)

    @property
    def R(self) -> int:
        return int(self.types.size)

    @property
    def N(self) -> int:
        X, Y, Z = self.shape
        return X * Y * Z

    def initial_counts(self) -> np.ndarray:
        c = np.bincount(self.lattice.reshape(-1), minlength=self.nS + 1)[1:self.nS + 1]
        return c.astype(np.uint32)

    def parameters(self, seed: int, max_steps: Optional[int] = None) -> Dict[str, str]:
        p = {"seed": str(seed), "timestep": repr(self.tau)}
        if max_steps is not None:
            p["maxTime"] = repr(max_steps * self.tau)
        return p

    def load_into(self, solver, replicate: int = 0, seed: int = 1, extra: Optional[Dict[str, str]] = None):
        """initialize + buildModel + buildDiffusionModel, the reference's order
        (runSimulation.cpp:227-242)."""
        prm = self.parameters(seed)
        if extra:
            prm.update(extra)
        solver.initialize(replicate, prm, boundary=self.boundary)
        solver.buildModel(self.nS, self.R, self.initial_counts(), self.types, self.K, self.S, self.D, 1)
        X, Y, Z = self.shape
        solver.buildDiffusionModel(self.nST, self.DF, self.RL, X, Y, Z, self.P,
                                   getattr(solver, "bytesPerParticle", 1), self.spacing, self.lattice, self.sites,
                                   True)
        return solver

    @property
    def bytes_per_particle(self) -> int:
        """4 for a uint32 lattice (IntMpdRdmeSolver), else 1."""
        return 4 if self.lattice.dtype == np.uint32 else 1

    # ---------------------------------------------------------------- files
    def reaction_model(self):
        """The ReactionModel message of this model (one rate constant each)."""
        from .simfile import ReactionModel

        return ReactionModel(number_species=self.nS, number_reactions=self.R,
                             initial_species_count=self.initial_counts(), types=self.types.astype(np.uint32),
                             rate_constants=[[float(k)] for k in self.K], S=self.S.astype(np.int32),
                             D=self.D.astype(np.uint32))

    def diffusion_model(self):
        """The DiffusionModel message of this model."""
        from .simfile import DiffusionModel

        X, Y, Z = self.shape
        return DiffusionModel(number_species=self.nS, number_reactions=self.R, number_site_types=self.nST,
                              lattice_spacing=self.spacing, lattice_x_size=X, lattice_y_size=Y, lattice_z_size=Z,
                              particles_per_site=self.P, bytes_per_particle=self.bytes_per_particle,
                              diffusion_matrix=self.DF.astype(np.float64),
                              reaction_location_matrix=self.RL.astype(np.uint32))

    def to_file(self, path: str, parameters: Optional[Dict[str, str]] = None) -> None:
        """A new Lattice Microbes HDF5 file holding this model, its initial
        lattice and the given /Parameters (``timestep`` is always written):
        what pyLM/jLM's ``save`` produces for the solver to run."""
        from .simfile import SimulationFile

        SimulationFile.create(path)
        prm = {"timestep": repr(self.tau)}
        prm.update({k: str(v) for k, v in (parameters or {}).items()})
        with SimulationFile(path) as f:
            for k, v in prm.items():
                f.setParameter(k, v)
            f.setReactionModel(self.reaction_model())
            f.setDiffusionModel(self.diffusion_model(), self.lattice, self.sites)

    @classmethod
    def from_file(cls, path: str, name: Optional[str] = None, boundary: str = "periodic") -> "ModelSpec":
        """The model, initial lattice and ``timestep`` of a Lattice Microbes
        HDF5 file.  Reactions keep their first rate constant, the only one the
        MPD-RDME solver reads."""
        from .simfile import SimulationFile

        with SimulationFile(path, writable=False) as f:
            tau = float(f.getParameter("timestep", "0") or 0)
            rm = f.getReactionModel()
            dm = f.getDiffusionModel()
            lattice, sites = f.getDiffusionModelLattice(dm)
        R = rm.number_reactions
        K = np.array([ks[0] if ks else 0.0 for ks in rm.rate_constants], dtype=np.float64)
        return cls(name=name or path, shape=dm.shape, spacing=dm.lattice_spacing, tau=tau,
                   P=dm.particles_per_site, nS=rm.number_species, nST=dm.number_site_types,
                   types=rm.types.astype(np.uint32).reshape(R), K=K.reshape(R),
                   S=rm.S.astype(np.int32).reshape(rm.number_species, R),
                   D=rm.D.astype(np.uint32).reshape(rm.number_species, R),
                   DF=dm.diffusion_matrix, RL=dm.reaction_location_matrix.reshape(R, dm.number_site_types),
                   lattice=lattice, sites=sites, boundary=boundary)

    def native_lattice(self) -> np.ndarray:
        """(P/4, Z, Y, X) uint32 planes, slots compacted in row-major order."""
        X, Y, Z = self.shape
        lat = np.transpose(self.lattice, (2, 1, 0, 3))  # Z, Y, X, P
        # compact nonzero bytes to the front of each site (addParticle order)
        nz = lat != 0
        order = np.argsort(~nz, axis=-1, kind="stable")
        comp = np.take_along_axis(lat, order, axis=-1)
        b = comp.reshape(Z, Y, X, self.P // 4, 4).astype(np.uint32)
        words = b[..., 0] | (b[..., 1] << 8) | (b[..., 2] << 16) | (b[..., 3] << 24)
        return np.ascontiguousarray(np.transpose(words, (3, 0, 1, 2)))

    def native_sites(self) -> np.ndarray:
        return np.ascontiguousarray(np.transpose(self.sites, (2, 1, 0)))


def site_nav(spacing: float) -> float:
    """N_A * V_site with V_site in litres."""
    return AVOGADRO * (spacing ** 3) * 1000.0


def to_lm_rate(k: float, order: int, spacing: float, nsites: int) -> float:
    """jLM RateConst._toLM (Types.py:336-339)."""
    cp = 1 - order
    return k * site_nav(spacing) ** cp * float(nsites) ** cp


# --------------------------------------------------------------------------- helpers
def _place_particles(rng: np.random.Generator, allowed: np.ndarray, species: np.ndarray,
                     N: int, P: int) -> np.ndarray:
    """Uniformly random placement of the given species codes on the allowed
    flat site indices (x fastest), at most P per site.  Returns [N, P]."""
    slots = np.zeros((N, P), dtype=np.uint8)
    occ = np.zeros(N, dtype=np.int64)
    remaining = species.astype(np.uint8)
    guard = 0
    while remaining.size:
        guard += 1
        if guard > 1000:
            raise RuntimeError("could not place particles (lattice too full)")
        sites = allowed[rng.integers(0, allowed.size, size=remaining.size)]
        order = np.argsort(sites, kind="stable")
        s_sorted = sites[order]
        sp_sorted = remaining[order]
        first = np.r_[0, np.flatnonzero(np.diff(s_sorted)) + 1]
        grp = np.repeat(first, np.diff(np.r_[first, s_sorted.size]))
        rank = np.arange(s_sorted.size) - grp
        slot = occ[s_sorted] + rank
        ok = slot < P
        slots[s_sorted[ok], slot[ok]] = sp_sorted[ok]
        np.add.at(occ, s_sorted[ok], 1)
        remaining = sp_sorted[~ok]
    return slots


def _to_row_major(slots: np.ndarray, shape: Tuple[int, int, int], P: int) -> np.ndarray:
    X, Y, Z = shape
    return np.ascontiguousarray(slots.reshape(Z, Y, X, P).transpose(2, 1, 0, 3))


def _sites_row_major(native_zyx: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.transpose(native_zyx, (2, 1, 0)).astype(np.uint8))


def _reaction_arrays(nS: int, reactions: Sequence[Tuple[Sequence[int], Sequence[int], float]],
                     spacing: float, nsites: int):
    """reactions: (reactant species list, product species list, rate in
    M/s units by order).  Returns types, K (LM units), S, D like pyLM's
    buildReactionModel (src/pylm/pyLM/RDME.py:573-660)."""
    R = len(reactions)
    types = np.zeros(R, dtype=np.uint32)
    K = np.zeros(R, dtype=np.float64)
    S = np.zeros((nS, R), dtype=np.int32)
    D = np.zeros((nS, R), dtype=np.uint32)
    for r, (reac, prod, k) in enumerate(reactions):
        order = len(reac)
        if order == 0:
            types[r] = ZEROTH
        elif order == 1:
            types[r] = FIRST
            D[reac[0], r] = 1
            S[reac[0], r] -= 1
        elif order == 2 and reac[0] != reac[1]:
            types[r] = SECOND
            for s in reac:
                D[s, r] = 1
                S[s, r] -= 1
        elif order == 2:
            types[r] = SECOND_SELF
            D[reac[0], r] = 1
            S[reac[0], r] -= 2
        else:
            raise ValueError("reactions of order > 2 are not supported by MPD")
        for s in prod:
            S[s, r] += 1
        K[r] = to_lm_rate(k, order, spacing, nsites)
    return types, K, S, D


# --------------------------------------------------------------------------- models
def bimolecular(shape=(64, 64, 64), spacing=32e-9, tau=50e-6, P=8, nA=1000, nB=1000, nC=0,
                kf=1e6, kr=0.1, DA=1e-12, DB=1e-12, DC=5e-13, seed=1,
                boundary="reflecting") -> ModelSpec:
    """A + B <-> C in one site type (the plain MPD benchmark model: 64^3,
    lambda = 32 nm, tau = 50 us, kf = 1e6 /M/s, kr = 0.1 /s)."""
    X, Y, Z = shape
    N = X * Y * Z
    rng = np.random.default_rng(seed)
    reactions = [((0, 1), (2,), kf), ((2,), (0, 1), kr)]
    types, K, S, D = _reaction_arrays(3, reactions, spacing, N)
    DF = np.zeros((1, 1, 3))
    DF[0, 0, :] = (DA, DB, DC)
    RL = np.ones((2, 1), dtype=np.uint32)
    sp = np.concatenate([np.full(nA, 1), np.full(nB, 2), np.full(nC, 3)]).astype(np.uint8)
    slots = _place_particles(rng, np.arange(N), sp, N, P)
    return ModelSpec("bimolecular", tuple(shape), spacing, tau, P, 3, 1, types, K, S, D, DF, RL,
                     _to_row_major(slots, shape, P), np.zeros(shape, dtype=np.uint8), boundary,
                     ["A", "B", "C"], ["cytoplasm"])


def _grid(shape):
    X, Y, Z = shape
    z = np.arange(Z, dtype=np.float32)[:, None, None]
    y = np.arange(Y, dtype=np.float32)[None, :, None]
    x = np.arange(X, dtype=np.float32)[None, None, :]
    return x, y, z


def _capsule_types(shape, radius_sites, half_length_sites, membrane_sites=1.0):
    """Spherocylinder along z centred in the box: 0 outside, 1 membrane shell,
    2 cytoplasm.  Returns native (Z, Y, X) uint8."""
    X, Y, Z = shape
    x, y, z = _grid(shape)
    cx, cy, cz = (X - 1) / 2.0, (Y - 1) / 2.0, (Z - 1) / 2.0
    dzc = np.clip(z - cz, -half_length_sites, half_length_sites)
    d2 = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz - dzc) ** 2
    t = np.zeros((Z, Y, X), dtype=np.uint8)
    t[d2 <= np.float32(radius_sites ** 2)] = 1
    t[d2 <= np.float32(max(radius_sites - membrane_sites, 0.0) ** 2)] = 2
    return t


def capsule_minde(shape=(128, 128, 256), spacing=16e-9, tau=50e-6, P=8, seed=1,
                  counts=(6000, 6000, 3000), boundary="periodic") -> ModelSpec:
    """Capsule cell with a membrane-oscillator (MinD/MinE) scheme: MinD.ATP
    binds membrane and recruits more, MinE binds membrane MinD, the complex
    hydrolyses back to cytoplasm (pattern of pylm-examples/example-minde.py).
    Radius 0.4 um, length 2 um, lambda 16 nm, ~1.5e4 particles."""
    X, Y, Z = shape
    N = X * Y * Z
    r = 0.4e-6 / spacing
    hl = (2.0e-6 / 2 - 0.4e-6) / spacing
    types_zyx = _capsule_types(shape, r, hl)
    names = ["minDatp", "minDadp", "minE", "minDm", "minDEm"]
    DATP, DADP, E, DM, DEM = range(5)
    reactions = [
        ((DADP,), (DATP,), 0.5),             # cytoplasm: nucleotide exchange
        ((DATP,), (DM,), 1.0),               # membrane: MinD binds membrane
        ((DATP, DM), (DM, DM), 9e6),         # membrane: recruitment
        ((DM, E), (DEM,), 1e7),              # membrane: MinE binds MinD
        ((DEM,), (E, DADP), 0.7),            # membrane: hydrolysis back to cytoplasm
    ]
    types, K, S, D = _reaction_arrays(5, reactions, spacing, N)
    RL = np.zeros((5, 3), dtype=np.uint32)
    RL[0, 2] = 1
    RL[1:, 1] = 1
    Dc = 2.5e-12
    DF = np.zeros((3, 3, 5))
    DF[2, 2, :] = Dc
    DF[1, 1, :] = Dc
    DF[1, 1, DM] = 1e-14
    DF[1, 1, DEM] = 1e-14
    for s in (DATP, DADP, E):
        DF[2, 1, s] = Dc
        DF[1, 2, s] = Dc
    rng = np.random.default_rng(seed)
    cyt = np.flatnonzero(types_zyx.reshape(-1) == 2)
    sp = np.concatenate([np.full(c, s + 1) for s, c in enumerate(counts)]).astype(np.uint8)
    slots = _place_particles(rng, cyt, sp, N, P)
    return ModelSpec("capsule_minde", tuple(shape), spacing, tau, P, 5, 3, types, K, S, D, DF, RL,
                     _to_row_major(slots, shape, P), _sites_row_major(types_zyx), boundary, names,
                     ["extracellular", "membrane", "cytoplasm"])


def whole_cell(shape=(256, 256, 512), spacing=10e-9, tau=4e-6, P=8, nS=100, R=300,
               n_particles=1_000_000, obstacle_fraction=0.2, seed=1,
               boundary="periodic") -> ModelSpec:
    """Whole-cell-scale model: 10 site types (extracellular, membrane,
    cytoplasm, obstacle, nucleoid and five spherical compartments), 1-site
    obstacles filling obstacle_fraction of the cytoplasm, nS species with D
    log-uniform in [1e-14, 1e-11] m^2/s, R first/second-order reactions with
    log-uniform rates (isomerisations, splits, and merges on distinct reactant
    pairs: no self reactions, no duplicate pairs), n_particles particles."""
    X, Y, Z = shape
    N = X * Y * Z
    rng = np.random.default_rng(seed)
    radius = 0.47 * min(X, Y)
    half_len = max(Z / 2.0 - radius - 2.0, 0.0)
    t = _capsule_types(shape, radius, half_len)  # 0 ext, 1 membrane, 2 cytoplasm
    x, y, z = _grid(shape)
    cx, cy, cz = (X - 1) / 2.0, (Y - 1) / 2.0, (Z - 1) / 2.0
    # nucleoid: ellipsoid around the centre
    nuc = ((x - cx) / (0.45 * radius)) ** 2 + ((y - cy) / (0.45 * radius)) ** 2 + \
          ((z - cz) / (0.45 * (half_len + radius))) ** 2 <= 1.0
    t[(t == 2) & nuc] = 4
    # five spherical compartments, types 5..9
    for k in range(5):
        rr = 0.12 * radius
        zc = cz + (k - 2) * 0.35 * (half_len + radius) / 2.0
        ang = 2 * np.pi * k / 5
        xc, yc = cx + 0.6 * radius * np.cos(ang), cy + 0.6 * radius * np.sin(ang)
        ball = (x - xc) ** 2 + (y - yc) ** 2 + (z - zc) ** 2 <= rr * rr
        t[(t == 2) & ball] = 5 + k
    del nuc
    cyt = np.flatnonzero(t.reshape(-1) == 2)
    nobs = int(obstacle_fraction * cyt.size)
    obs = rng.choice(cyt, size=nobs, replace=False)
    t.reshape(-1)[obs] = 3
    nST = 10
    fluid = [2, 4, 5, 6, 7, 8, 9]
    Dsp = 10.0 ** rng.uniform(-14, -11, size=nS)
    DF = np.zeros((nST, nST, nS))
    for a in fluid:
        for b in fluid:
            DF[a, b, :] = Dsp
    nmem = min(10, nS)
    DF[1, 1, :nmem] = Dsp[:nmem] / 10.0
    DF[2, 1, :nmem] = Dsp[:nmem]
    DF[1, 2, :nmem] = Dsp[:nmem]
    # reactions: thirds of isomerisations, splits and merges
    reactions = []
    pairs = set()
    n_iso = R // 3
    n_split = R // 3
    n_merge = R - n_iso - n_split
    for _ in range(n_iso):
        a, b = rng.choice(nS, size=2, replace=False)
        reactions.append(((int(a),), (int(b),), 10.0 ** rng.uniform(-1, 2)))
    for _ in range(n_split):
        c, a, b = rng.choice(nS, size=3, replace=False)
        reactions.append(((int(c),), (int(a), int(b)), 10.0 ** rng.uniform(-1, 2)))
    while len(reactions) < R:
        a, b, c = rng.choice(nS, size=3, replace=False)
        key = (min(a, b), max(a, b))
        if key in pairs:
            continue
        pairs.add(key)
        reactions.append(((int(a), int(b)), (int(c),), 10.0 ** rng.uniform(5, 8)))
    types, K, S, D = _reaction_arrays(nS, reactions, spacing, N)
    RL = np.zeros((R, nST), dtype=np.uint32)
    RL[:, 2] = 1
    extra = rng.random((R, len(fluid) - 1)) < 0.5
    RL[:, [4, 5, 6, 7, 8, 9]] = extra.astype(np.uint32)
    allowed = np.flatnonzero(np.isin(t.reshape(-1), fluid))
    sp = rng.integers(1, nS + 1, size=n_particles).astype(np.uint8)
    slots = _place_particles(rng, allowed, sp, N, P)
    names = [f"S{i}" for i in range(nS)]
    tnames = ["extracellular", "membrane", "cytoplasm", "obstacle", "nucleoid",
              "compartment1", "compartment2", "compartment3", "compartment4", "compartment5"]
    return ModelSpec("whole_cell", tuple(shape), spacing, tau, P, nS, nST, types, K, S, D, DF, RL,
                     _to_row_major(slots, shape, P), _sites_row_major(t), boundary, names, tnames)


def random_model(shape=(8, 6, 5), P=8, nS=4, nST=3, R=6, density=0.5, seed=0, boundary="periodic",
                 qmax=0.5, rate_scale=1.0, include_zero_order=True, include_self=True) -> ModelSpec:
    """Small random model for parity tests: random site types, random
    transition probabilities up to qmax (including exactly 0.5 and 0), random
    reactions of all four MPD types, density * P particles per site on average
    (dense lattices exercise the overflow path)."""
    X, Y, Z = shape
    N = X * Y * Z
    rng = np.random.default_rng(seed)
    spacing = 1e-8
    tau = 1e-6
    lam2 = spacing * spacing
    q = rng.uniform(0.0, qmax, size=(nST, nST, nS))
    q[rng.random(q.shape) < 0.15] = 0.0
    if qmax >= 0.5:
        q[rng.random(q.shape) < 0.1] = 0.5
    DF = q * lam2 / tau
    # make sure T computes to <= 0.5 after the float conversion
    for _ in range(4):
        T = (DF * tau / lam2).astype(np.float32)
        bad = T > np.float32(0.5)
        if not bad.any():
            break
        DF[bad] = np.nextafter(DF[bad], 0.0)
    reactions = []
    for r in range(R):
        kind = rng.integers(0, 4)
        if kind == 0 and include_zero_order:
            reactions.append(((), (int(rng.integers(nS)),), rate_scale * 10.0 ** rng.uniform(-1, 0.5)))
        elif kind == 1 or (kind == 0 and not include_zero_order):
            a = int(rng.integers(nS))
            prods = tuple(int(v) for v in rng.integers(0, nS, size=int(rng.integers(0, 3))))
            reactions.append(((a,), prods, rate_scale * 10.0 ** rng.uniform(3, 5.5)))
        elif kind == 2 or not include_self:
            a, b = rng.choice(nS, size=2, replace=False)
            prods = tuple(int(v) for v in rng.integers(0, nS, size=int(rng.integers(0, 3))))
            reactions.append(((int(a), int(b)), prods, rate_scale * 10.0 ** rng.uniform(8, 10)))
        else:
            a = int(rng.integers(nS))
            prods = tuple(int(v) for v in rng.integers(0, nS, size=int(rng.integers(0, 3))))
            reactions.append(((a, a), prods, rate_scale * 10.0 ** rng.uniform(8, 10)))
    # jLM conversion with this model's own site volume
    types, K, S, D = _reaction_arrays(nS, reactions, spacing, N)
    RL = (rng.random((R, nST)) < 0.7).astype(np.uint32)
    st = rng.integers(0, nST, size=N).astype(np.uint8).reshape(Z, Y, X)
    nparticles = int(density * P * N)
    sp = rng.integers(1, nS + 1, size=nparticles).astype(np.uint8)
    slots = _place_particles(rng, np.arange(N), sp, N, P)
    return ModelSpec("random", tuple(shape), spacing, tau, P, nS, nST, types, K, S, D, DF, RL,
                     _to_row_major(slots, shape, P), _sites_row_major(st), boundary)


# --------------------------------------------------------------------------- ensemble and serving workloads
ENSEMBLE_SIZES = (32, 48, 64, 96, 128)


def ensemble_mix(nrep: int = 256, seed: int = 1):
    """The batched-ensemble workload: nrep replicates of the bimolecular
    model, each on a lattice drawn uniformly from 32^3, 48^3, 64^3, 96^3 and
    128^3, with its own seed.  Returns (model, replicate, seed) per replicate:
    replicate r runs its lattice size's model with replicate number r and
    seed 1000 + r."""
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, len(ENSEMBLE_SIZES), nrep)
    templates = {L: bimolecular(shape=(L, L, L), seed=seed + L) for L in ENSEMBLE_SIZES}
    return [(templates[ENSEMBLE_SIZES[int(picks[r])]], r, 1000 + r) for r in range(nrep)]


def whole_cell_reduced(seed: int = 1) -> ModelSpec:
    """The whole-cell model reduced for serving: 256x256x128, 30 species, 90
    reactions, 2.5e5 particles (a quarter of the full lattice's)."""
    return whole_cell(shape=(256, 256, 128), nS=30, R=90, n_particles=250_000, seed=seed)


def serving_templates(seed: int = 1) -> List[ModelSpec]:
    """The request mix of the serving workload: the bimolecular 64^3 model,
    the capsule 128x128x256 and the reduced whole-cell 256x256x128."""
    return [bimolecular(shape=(64, 64, 64), seed=seed), capsule_minde(seed=seed + 1),
            whole_cell_reduced(seed=seed + 2)]
