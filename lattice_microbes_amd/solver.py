"""Python mirror of ``lm::rdme::MpdRdmeSolver`` over the lmx C ABI.

Method names, argument meaning and error behaviour follow the reference class
(src/rdme/MpdRdmeSolver.h:66-156) as exposed to Python by its SWIG module
(src/swig/lm.i:949-967): ``initialize``, ``buildModel``,
``buildDiffusionModel``, ``setReactionRate``, ``generateTrajectory`` and the
overridable hooks.  Every call goes to the HIP library; nothing here computes
the step.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_size_t, c_uint32, c_uint64, c_void_p
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _lib

SOLVER_NAME = "lm::rdme::MpdRdmeSolver"


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def timestep_seed(seed32: int, timestep: int, substep: int) -> int:
    """getTimestepSeed (MpdRdmeSolver.cu:987-994), for reference and tests."""
    m = (1 << 64) - 1
    h = ((((seed32 << 30) + timestep) << 2) + substep) & m
    h = (h * 3202034522624059733 + 4354685564936845319) & m
    h ^= h >> 20
    h = (h ^ (h << 41)) & m
    h ^= h >> 5
    return (h * 7664345821815920749) & m


class MpdRdmeSolver:
    """Drop-in for the reference MPD-RDME solver, running on one MI355X."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        h = c_void_p()
        _lib.check(self._lib.lmx_create(int(device), byref(h)))
        self._h = h
        self.device = device
        self.parameters: Dict[str, str] = {}
        self.replicate = 0
        self.numberSpecies = 0
        self.numberReactions = 0
        self.numberSiteTypes = 0
        self.shape = (0, 0, 0)
        self.particlesPerSite = 0
        self.tau = 0.0
        self.latticeSpacing = 0.0

    # ------------------------------------------------------------------ lifecycle
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lmx_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ MESolver
    def needsReactionModel(self) -> bool:
        return True

    def needsDiffusionModel(self) -> bool:
        return True

    def initialize(self, replicate: int, parameters: Dict[str, str], resources=None,
                   boundary: str = "periodic") -> None:
        """MpdRdmeSolver::initialize (MpdRdmeSolver.cu:107-162).

        ``parameters`` holds the reference's string map: ``seed``, ``timestep``,
        ``rdme.mpd.overflowhandler`` and, for generateTrajectory, ``maxTime``,
        ``writeInterval``, ``latticeWriteInterval``, ``hookInterval`` (steps).
        ``boundary`` selects periodic (every reference build) or reflecting.
        """
        self.parameters = {k: str(v) for k, v in parameters.items()}
        self.replicate = int(replicate)
        seed_top = int(float(self.parameters.get("seed", "0") or 0)) & 0xFFFFFFFF
        tau = float(self.parameters.get("timestep", "0") or 0)
        handler = self.parameters.get("rdme.mpd.overflowhandler", "")
        if handler not in ("", "classic"):
            raise _lib.InvalidArgException(_lib.ERR_ARG,
                                           f"rdme.mpd.overflowhandler '{handler}' is not supported")
        bmode = {"periodic": _lib.BOUNDARY_PERIODIC, "reflecting": _lib.BOUNDARY_REFLECTING}[boundary]
        _lib.check(self._lib.lmx_initialize(self._h, self.replicate, seed_top, tau, 0, bmode))
        self.tau = tau

    @property
    def seed(self) -> int:
        v = c_uint32()
        _lib.check(self._lib.lmx_get_seed(self._h, byref(v)))
        return v.value

    def getTimestepSeed(self, timestep: int, substep: int) -> int:
        return timestep_seed(self.seed, timestep, substep)

    # ------------------------------------------------------------------ models
    def buildModel(self, numberSpecies: int, numberReactions: int, initialSpeciesCounts,
                   reactionTypes, K, S, D, kCols: int = 1) -> None:
        """CMESolver/MpdRdmeSolver::buildModel.  S and D are species-major
        [numberSpecies][numberReactions] as in the ReactionModel message."""
        nS, R = int(numberSpecies), int(numberReactions)
        init = np.ascontiguousarray(np.asarray(initialSpeciesCounts, dtype=np.uint32).reshape(-1))
        types = np.ascontiguousarray(np.asarray(reactionTypes, dtype=np.uint32).reshape(-1))
        Ka = np.ascontiguousarray(np.asarray(K, dtype=np.float64).reshape(-1))
        Sa = np.ascontiguousarray(np.asarray(S, dtype=np.int32).reshape(-1))
        Da = np.ascontiguousarray(np.asarray(D, dtype=np.uint32).reshape(-1))
        if init.size < nS or types.size < R or Ka.size < R * kCols or Sa.size < nS * R or Da.size < nS * R:
            raise _lib.InvalidArgException(_lib.ERR_ARG, "model arrays are smaller than declared")
        _lib.check(self._lib.lmx_build_model(self._h, nS, R, _ptr(init, c_uint32), _ptr(types, c_uint32),
                                             _ptr(Ka, c_double), _ptr(Sa, ctypes.c_int32),
                                             _ptr(Da, c_uint32), int(kCols)))
        self.numberSpecies, self.numberReactions = nS, R

    def buildDiffusionModel(self, numberSiteTypes: int, DF, RL, latticeXSize: int, latticeYSize: int,
                            latticeZSize: int, particlesPerSite: int, bytes_per_particle: int,
                            latticeSpacing: float, latticeData, latticeSitesData,
                            rowMajorData: bool = True) -> None:
        """RDMESolver/MpdRdmeSolver::buildDiffusionModel.  latticeData is the
        row-major (x,y,z,p) byte array, latticeSitesData row-major (x,y,z)."""
        DFa = np.ascontiguousarray(np.asarray(DF, dtype=np.float64).reshape(-1))
        RLa = np.ascontiguousarray(np.asarray(RL, dtype=np.uint32).reshape(-1))
        lat = np.ascontiguousarray(np.asarray(latticeData, dtype=np.uint8).reshape(-1))
        sites = np.ascontiguousarray(np.asarray(latticeSitesData, dtype=np.uint8).reshape(-1))
        X, Y, Z, P = int(latticeXSize), int(latticeYSize), int(latticeZSize), int(particlesPerSite)
        if lat.size != X * Y * Z * P:
            raise _lib.InvalidArgException(_lib.ERR_ARG, "latticeSize: the lattice data size does not agree with the lattice dimensions")
        if sites.size != X * Y * Z:
            raise _lib.InvalidArgException(_lib.ERR_ARG, "latticeSitesSize: the lattice data size does not agree with the lattice dimensions")
        if RLa.size == 0:
            RLa = np.zeros(1, dtype=np.uint32)
        _lib.check(self._lib.lmx_build_diffusion_model(
            self._h, int(numberSiteTypes), _ptr(DFa, c_double), _ptr(RLa, c_uint32), X, Y, Z, P,
            int(bytes_per_particle), float(latticeSpacing), _ptr(lat, ctypes.c_uint8),
            _ptr(sites, ctypes.c_uint8), 1 if rowMajorData else 0))
        self.numberSiteTypes = int(numberSiteTypes)
        self.shape = (X, Y, Z)
        self.particlesPerSite = P
        self.latticeSpacing = float(latticeSpacing)

    def setReactionRate(self, rxid: int, rate: float) -> None:
        _lib.check(self._lib.lmx_set_reaction_rate(self._h, int(rxid), float(rate)))

    def setPath(self, path: int) -> None:
        _lib.check(self._lib.lmx_set_path(self._h, int(path)))

    def setLaunchOption(self, option: int, value: int) -> None:
        """Launch-shape options (lmx_option); results do not depend on them."""
        _lib.check(self._lib.lmx_set_launch_option(self._h, int(option), int(value)))

    # ------------------------------------------------------------------ stepping
    def runTimesteps(self, nsteps: int, synchronize: bool = True) -> None:
        """runTimestep (MpdRdmeSolver.cu:996-1193) nsteps times."""
        _lib.check(self._lib.lmx_run_steps(self._h, int(nsteps)))
        if synchronize:
            self.synchronize()

    def synchronize(self) -> None:
        _lib.check(self._lib.lmx_synchronize(self._h))

    @property
    def timestep(self) -> int:
        v = c_uint32()
        _lib.check(self._lib.lmx_get_timestep(self._h, byref(v)))
        return v.value

    def setTimestep(self, t: int) -> None:
        _lib.check(self._lib.lmx_set_timestep(self._h, int(t)))

    def timeSteps(self, nsteps: int, per_kernel: bool = False):
        """Run nsteps bracketed by HIP events; returns (total_ms, {kernel: (ms, launches)})."""
        total = c_double()
        nk = len(_lib.KERNEL_NAMES)
        kms = (c_double * nk)()
        kl = (c_uint64 * nk)()
        _lib.check(self._lib.lmx_time_steps(self._h, int(nsteps), byref(total),
                                            kms if per_kernel else None,
                                            kl if per_kernel else None, nk))
        per = {}
        if per_kernel:
            per = {_lib.KERNEL_NAMES[k]: (kms[k], int(kl[k])) for k in range(nk) if kl[k]}
        return total.value, per

    # ------------------------------------------------------------------ lattice
    def getLattice(self) -> np.ndarray:
        """Particles as row-major (X, Y, Z, P) uint8."""
        X, Y, Z = self.shape
        out = np.empty((X, Y, Z, self.particlesPerSite), dtype=np.uint8)
        _lib.check(self._lib.lmx_get_lattice(self._h, _ptr(out, ctypes.c_uint8), out.size))
        return out

    def getLatticeNative(self) -> np.ndarray:
        """Particles in the native layout: (P/4, Z, Y, X) uint32 planes."""
        X, Y, Z = self.shape
        out = np.empty((self.particlesPerSite // 4, Z, Y, X), dtype=np.uint32)
        _lib.check(self._lib.lmx_get_lattice_native(self._h, _ptr(out, c_uint32), out.size))
        return out

    def getSites(self) -> np.ndarray:
        X, Y, Z = self.shape
        out = np.empty((X, Y, Z), dtype=np.uint8)
        _lib.check(self._lib.lmx_get_sites(self._h, _ptr(out, ctypes.c_uint8), out.size))
        return out

    def setLattice(self, latticeData) -> None:
        lat = np.ascontiguousarray(np.asarray(latticeData, dtype=np.uint8).reshape(-1))
        _lib.check(self._lib.lmx_set_lattice(self._h, _ptr(lat, ctypes.c_uint8), lat.size))

    def setLatticeNative(self, native) -> None:
        lat = np.ascontiguousarray(np.asarray(native, dtype=np.uint32).reshape(-1))
        _lib.check(self._lib.lmx_set_lattice_native(self._h, _ptr(lat, c_uint32), lat.size))

    def setSites(self, sitesData) -> None:
        s = np.ascontiguousarray(np.asarray(sitesData, dtype=np.uint8).reshape(-1))
        _lib.check(self._lib.lmx_set_sites(self._h, _ptr(s, ctypes.c_uint8), s.size))

    def getSpeciesCounts(self) -> np.ndarray:
        """ByteLattice::getParticleCounts for species 1..numberSpecies."""
        n = max(self.numberSpecies, 1)
        out = np.zeros(n, dtype=np.uint32)
        _lib.check(self._lib.lmx_get_species_counts(self._h, _ptr(out, c_uint32), n))
        return out[: self.numberSpecies]

    def getOverflowStats(self):
        e, s = c_uint64(), c_uint64()
        _lib.check(self._lib.lmx_get_overflow_stats(self._h, byref(e), byref(s)))
        return e.value, s.value

    def getTable(self, which: str) -> np.ndarray:
        idx = {"T": 0, "qp0": 1, "qp1": 2, "qp2": 3, "rate": 4}[which]
        n = c_size_t()
        _lib.check(self._lib.lmx_get_table_f32(self._h, idx, None, 0, byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.float32)
        _lib.check(self._lib.lmx_get_table_f32(self._h, idx, _ptr(out, ctypes.c_float), out.size, byref(n)))
        return out[: n.value]

    # ------------------------------------------------------------------ trajectory
    def hookSimulation(self, time: float, solver: "MpdRdmeSolver") -> int:
        """Overridable; return codes as MpdRdmeSolver::hookCheckSimulation
        (MpdRdmeSolver.cu:958-985): 0 unchanged, 1 lattice changed (already
        written back with setLattice), 2 also sites changed, 3 "stop" (which,
        as in the reference, does not stop)."""
        return 0

    def onBeginTrajectory(self, solver: "MpdRdmeSolver") -> int:
        return 0

    def onEndTrajectory(self, solver: "MpdRdmeSolver") -> int:
        return 0

    def generateTrajectory(self) -> Dict[str, object]:
        """MpdRdmeSolver::generateTrajectory (MpdRdmeSolver.cu:485-668) without
        file output: species counts are recorded every ``writeInterval`` steps,
        lattices every ``latticeWriteInterval`` steps and the hook runs every
        ``hookInterval`` steps.  Returns the recorded data."""
        prm = self.parameters
        maxTime = float(prm.get("maxTime", "0") or 0)
        wi = int(float(prm.get("writeInterval", "0") or 0))
        lwi = int(float(prm.get("latticeWriteInterval", "0") or 0))
        hook_iv = prm.get("hookInterval", "")
        hi = int(float(hook_iv)) if hook_iv not in ("", None) else 0
        hook_enabled = hook_iv not in ("", None)
        wi, lwi, hi = max(wi, 1), max(lwi, 1), max(hi, 1)  # an interval of 0 steps fires every step
        tau = self.tau
        counts_t: List[float] = [0.0]
        counts: List[np.ndarray] = []
        lattice_t: List[float] = [0.0]
        lattices: List[np.ndarray] = []
        self.onBeginTrajectory(self)
        counts.append(self.getSpeciesCounts())
        lattices.append(self.getLattice())
        t = 0
        if hook_enabled:
            self.hookSimulation(0.0, self)
        next_c, next_l, next_h = wi, lwi, hi
        time = 0.0
        while time < maxTime:
            # Run until the next event step in one batch (no per-step host sync).
            targets = [x for x in (next_c, next_l, next_h if hook_enabled else None) if x is not None and x > t]
            stop = int(np.ceil(maxTime / tau)) if tau > 0 else t + 1
            nxt = min(targets + [stop]) if targets else stop
            n = max(nxt - t, 1)
            self.runTimesteps(n, synchronize=True)
            t += n
            time = t * tau
            if t >= next_c or t >= next_l or (hook_enabled and t >= next_h):
                if hook_enabled and t >= next_h:
                    self.hookSimulation(time, self)
                    next_h += hi
                if t >= next_l:
                    lattice_t.append(time)
                    lattices.append(self.getLattice())
                    next_l += lwi
                if t >= next_c:
                    counts_t.append(time)
                    counts.append(self.getSpeciesCounts())
                    next_c += wi
        self.onEndTrajectory(self)
        return {"species_count_times": np.array(counts_t), "species_counts": np.array(counts),
                "lattice_times": np.array(lattice_t), "lattices": lattices}


class MESolverFactory:
    """String-based selection (MESolverFactory.cpp:75-96,115-190)."""

    _solvers: Dict[str, Callable[..., MpdRdmeSolver]] = {SOLVER_NAME: MpdRdmeSolver}

    def __init__(self):
        self.name: Optional[str] = None

    def setSolver(self, name: str) -> None:
        if name not in self._solvers:
            raise _lib.InvalidArgException(_lib.ERR_ARG, f"Invalid solver class specified: {name}")
        self.name = name

    def instantiate(self, device: int = 0) -> MpdRdmeSolver:
        if self.name is None:
            raise _lib.LmxError(_lib.ERR_STATE, "no solver selected")
        return self._solvers[self.name](device)
