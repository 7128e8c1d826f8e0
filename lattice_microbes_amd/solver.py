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
import signal
import sys
import threading
import time as _time
import warnings
from ctypes import byref, c_double, c_size_t, c_uint32, c_uint64, c_void_p
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _lib
from .lattice import ByteLattice, IntLattice

SOLVER_NAME = "lm::rdme::MpdRdmeSolver"
INT_SOLVER_NAME = "lm::rdme::IntMpdRdmeSolver"


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


# globalAbort (src/core/SignalHandler.cpp:109): a stop request that every
# running trajectory sees before its next batch of steps.
_abort = threading.Event()
# Steps per lmx_run_steps batch at most: a stop request and the overflow-use
# window of 1000 steps are seen within this many steps.
_MAX_BATCH = 1000


def request_abort() -> None:
    """Ask running trajectories to stop (the reference's globalAbort): each
    ends after its current batch of steps, and its end of trajectory still
    runs (final lattice to onEndTrajectory, buffered counts and max counts
    written), MpdRdmeSolver.cu:561-565, 662-667."""
    _abort.set()


def clear_abort() -> None:
    _abort.clear()


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

    # The byte lattice: one byte per particle, at most 255 species
    # (MpdRdmeSolver::allocateLattice, MpdRdmeSolver.cu:189-199).
    bytesPerParticle = 1
    # MpdRdmeSolver reads rdme.mpd.overflowhandler ("classic", "relaxed");
    # IntMpdRdmeSolver has only the classic handler and ignores it.
    _parses_overflow_handler = True
    overflowHandler = "classic"

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
        overflow_mode = _lib.OVERFLOW_CLASSIC
        if self._parses_overflow_handler:
            if handler == "relaxed":
                overflow_mode = _lib.OVERFLOW_RELAXED
            elif handler not in ("", "classic"):
                # the reference warns and keeps the classic handler (MpdRdmeSolver.cu:142-146)
                warnings.warn(f"Unknown overflow handler requested: '{handler}'")
        self.overflowHandler = "relaxed" if overflow_mode == _lib.OVERFLOW_RELAXED else "classic"
        bmode = {"periodic": _lib.BOUNDARY_PERIODIC, "reflecting": _lib.BOUNDARY_REFLECTING}[boundary]
        _lib.check(self._lib.lmx_initialize(self._h, self.replicate, seed_top, tau, overflow_mode, bmode))
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
        row-major (x,y,z,p) particle array (uint8, or uint32 for the uint32
        lattice), latticeSitesData row-major (x,y,z)."""
        if int(bytes_per_particle) != self.bytesPerParticle:
            raise _lib.InvalidArgException(
                _lib.ERR_ARG, f"{type(self).__name__} works only with lattices of {self.bytesPerParticle} "
                              f"bytes per particle, not {int(bytes_per_particle)}")
        if self.bytesPerParticle == 1 and self.numberSpecies > 255:
            raise _lib.InvalidArgException(_lib.ERR_ARG, "a byte lattice holds at most 255 species")
        DFa = np.ascontiguousarray(np.asarray(DF, dtype=np.float64).reshape(-1))
        RLa = np.ascontiguousarray(np.asarray(RL, dtype=np.uint32).reshape(-1))
        lat = self._particle_array(latticeData)
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
            int(bytes_per_particle), float(latticeSpacing), _ptr(lat.view(np.uint8), ctypes.c_uint8),
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

    def autotune(self, steps: int = 8) -> Dict[str, object]:
        """Times every y/z sweep kernel and window shape for ``steps`` steps
        from the current state and keeps the fastest (lmx_autotune); the state
        is unchanged afterwards and results do not depend on the choice."""
        best = c_double()
        chosen = (ctypes.c_int * 4)()
        _lib.check(self._lib.lmx_autotune(self._h, int(steps), byref(best), chosen))
        return {"path": int(chosen[0]), "yz_kernel": int(chosen[1]), "march_segment": int(chosen[2]),
                "fuse_reaction": bool(chosen[3]), "ms_per_step": best.value}

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
    @property
    def particleDtype(self):
        return np.uint32 if self.bytesPerParticle == 4 else np.uint8

    @property
    def wordsPerSite(self) -> int:
        return self.particlesPerSite if self.bytesPerParticle == 4 else self.particlesPerSite // 4

    def _particle_array(self, data) -> np.ndarray:
        a = np.asarray(data)
        if a.dtype != self.particleDtype and a.size and (a.min() < 0 or a.max() > np.iinfo(self.particleDtype).max):
            raise _lib.InvalidArgException(_lib.ERR_ARG, "latticeData: an invalid species was found")
        return np.ascontiguousarray(a.astype(self.particleDtype, copy=False).reshape(-1))

    def getLattice(self) -> np.ndarray:
        """Particles as row-major (X, Y, Z, P): uint8, or uint32 for the
        uint32 lattice."""
        X, Y, Z = self.shape
        out = np.empty((X, Y, Z, self.particlesPerSite), dtype=self.particleDtype)
        _lib.check(self._lib.lmx_get_lattice(self._h, _ptr(out.view(np.uint8), ctypes.c_uint8), out.nbytes))
        return out

    def getLatticeNative(self) -> np.ndarray:
        """Particles in the native layout: (W, Z, Y, X) uint32 planes, W = P/4
        (four byte slots per word) or P (one particle per word)."""
        X, Y, Z = self.shape
        out = np.empty((self.wordsPerSite, Z, Y, X), dtype=np.uint32)
        _lib.check(self._lib.lmx_get_lattice_native(self._h, _ptr(out, c_uint32), out.size))
        return out

    def getSites(self) -> np.ndarray:
        X, Y, Z = self.shape
        out = np.empty((X, Y, Z), dtype=np.uint8)
        _lib.check(self._lib.lmx_get_sites(self._h, _ptr(out, ctypes.c_uint8), out.size))
        return out

    def setLattice(self, latticeData) -> None:
        lat = self._particle_array(latticeData)
        _lib.check(self._lib.lmx_set_lattice(self._h, _ptr(lat.view(np.uint8), ctypes.c_uint8), lat.nbytes))

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

    def getUnresolvedOverflows(self) -> int:
        """Relaxed handler: overflow entries whose own site was full after the
        last step; they are retried after the next one (MpdRdmeSolver.cu:1094-1101)."""
        n = c_uint32()
        _lib.check(self._lib.lmx_get_unresolved_overflows(self._h, byref(n)))
        return int(n.value)

    def getTable(self, which: str) -> np.ndarray:
        idx = {"T": 0, "qp0": 1, "qp1": 2, "qp2": 3, "rate": 4}[which]
        n = c_size_t()
        _lib.check(self._lib.lmx_get_table_f32(self._h, idx, None, 0, byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.float32)
        _lib.check(self._lib.lmx_get_table_f32(self._h, idx, _ptr(out, ctypes.c_float), out.size, byref(n)))
        return out[: n.value]

    # ------------------------------------------------------------------ models from a file
    def setReactionModel(self, rm) -> None:
        """CMESolver::setReactionModel (CMESolver.cpp:278-305): buildModel from a
        ReactionModel (``simfile.ReactionModel``), K with as many columns as the
        reaction with the most rate constants."""
        R = int(rm.number_reactions)
        if R != len(rm.types) or R != len(rm.rate_constants):
            raise _lib.InvalidArgException(_lib.ERR_ARG, "rm: number of reactions does not agree with reaction list size")
        kCols = max([len(k) for k in rm.rate_constants] + [0])
        K = np.zeros((R, max(kCols, 1)))
        for r, ks in enumerate(rm.rate_constants):
            K[r, :len(ks)] = ks
        self.buildModel(rm.number_species, R, rm.initial_species_count, rm.types, K, rm.S, rm.D, max(kCols, 1))

    def setDiffusionModel(self, dm, lattice, latticeSites) -> None:
        """RDMESolver::setDiffusionModel (RDMESolver.cpp:97-108): a
        DiffusionModel (``simfile.DiffusionModel``) with its row-major lattice
        and site arrays, checked against the reaction model's sizes."""
        if dm.number_species != self.numberSpecies:
            raise _lib.InvalidArgException(_lib.ERR_ARG, "dm.number_species: number of species in the diffusion model does not agree with the number in the reaction model")
        if dm.number_reactions != self.numberReactions:
            raise _lib.InvalidArgException(_lib.ERR_ARG, "dm.number_reactions: number of reactions in the diffusion model does not agree with the number in the reaction model")
        RL = dm.reaction_location_matrix if dm.number_reactions else np.zeros(0, np.uint32)
        self.buildDiffusionModel(dm.number_site_types, dm.diffusion_matrix, RL, dm.lattice_x_size,
                                 dm.lattice_y_size, dm.lattice_z_size, dm.particles_per_site,
                                 dm.bytes_per_particle, dm.lattice_spacing, lattice, latticeSites, True)

    # ------------------------------------------------------------------ trajectory
    def hookSimulation(self, time: float, lattice: "HookLattice") -> int:
        """Overridable (MpdRdmeSolver.cu:789-799).  ``lattice`` is a host copy
        of the solver's lattice; return 0 if unchanged, 1 if the particles or
        sites were changed (they are copied back to the GPU), 2 if so and the
        sites are also to be recorded, 3 to record the sites without copying
        back (hookCheckSimulation, MpdRdmeSolver.cu:958-985; as there, 3 does
        not end the trajectory).  setReactionRate may be called from here."""
        return 0

    def onBeginTrajectory(self, lattice: "HookLattice") -> int:
        """Overridable (MpdRdmeSolver.cu:813-816); changes to ``lattice`` are
        copied to the GPU before the first step."""
        return 0

    def onEndTrajectory(self, lattice: "HookLattice") -> int:
        """Overridable (MpdRdmeSolver.cu:818-821); sees the final lattice."""
        return 0

    def generateTrajectory(self, output=None) -> Dict[str, object]:
        """MpdRdmeSolver::generateTrajectory (MpdRdmeSolver.cu:485-668).

        Species counts are recorded at time 0 and every ``writeInterval``
        steps, lattices at time 0 and every ``latticeWriteInterval`` steps, the
        hook runs at time 0 and every ``hookInterval`` steps (all intervals in
        steps, as the reference compares them with the step counter), until
        ``maxTime``.  Steps between events run back to back on the GPU; the
        lattice comes to the host only at an event.  With ``output`` (an open
        ``simfile.SimulationFile``) the records go to replicate
        ``self.replicate`` of that file the way the reference's output queue
        writes them: SpeciesCounts/SpeciesCountTimes, Lattice/LatticeTimes,
        Sites/SiteTimes (hook codes 2 and 3) and, at the end,
        MaxParticleCounts and MaxSiteCounts.  Returns the same records."""
        prm = self.parameters
        maxTime = float(prm.get("maxTime", "0") or 0)
        wi = int(float(prm.get("writeInterval", "0") or 0))
        lwi = int(float(prm.get("latticeWriteInterval", "0") or 0))
        hook_iv = prm.get("hookInterval", "")
        hook_enabled = hook_iv not in ("", None)
        hi = int(float(hook_iv)) if hook_enabled else 0
        rec = _Recorder(self, output)
        lattice = IntHookLattice(self) if self.bytesPerParticle == 4 else HookLattice(self)
        device = bool(getattr(self, "deviceHooks", False))
        dlat = DeviceHookLattice(self) if device else None

        if device:
            self._device_call(self.onBeginTrajectory, None, dlat, validate=True)
            lattice.copyFromGPU()
        else:
            lattice.copyFromGPU()
            self.onBeginTrajectory(lattice)
            lattice.copyToGPU()
        time, t = 0.0, 0
        rec.species_counts(time, lattice)
        rec.lattice(time, lattice)
        rec.init_max_counts(lattice)
        self._hook_check(time, lattice, dlat, rec)

        next_c, next_l, next_h = wi, lwi, hi
        end = _last_step(self.tau, maxTime)
        perf_iv = float(prm.get("perfPrintInterval", "") or 60.0)
        window_end, ovf_at_window = 1000, self.getOverflowStats()[1]
        perf_t0, perf_steps = _time.perf_counter(), 0
        # SIGINT and SIGTERM become a stop request for the duration of the
        # trajectory (the reference's signal thread sets globalAbort).
        restore = {}
        if threading.current_thread() is threading.main_thread():
            for sig in (signal.SIGINT, signal.SIGTERM):
                restore[sig] = signal.signal(sig, lambda *_: _abort.set())
        try:
            while time < maxTime:
                if _abort.is_set():
                    print("Global abort: terminating solver", file=sys.stderr, flush=True)
                    break
                # Run to the next step at which anything happens (an interval of
                # 0 fires after every step, as in the reference) or the last step.
                due = min([next_c, next_l] + ([next_h] if hook_enabled else []))
                stop = min(max(due, t + 1), max(end, t + 1), window_end, t + _MAX_BATCH)
                self.runTimesteps(stop - t, synchronize=True)
                perf_steps += stop - t
                t = stop
                time = t * self.tau
                if self.overflowHandler == "relaxed":
                    unresolved = self.getUnresolvedOverflows()
                    if unresolved > 0:
                        warnings.warn(f"{unresolved} unresolved overflows")
                if t >= window_end:
                    # overflowTimesteps / overflowListUses (MpdRdmeSolver.cu:1180-1186)
                    used = self.getOverflowStats()[1]
                    if self.overflowHandler != "relaxed" and used - ovf_at_window > 10:
                        warnings.warn(f"{used - ovf_at_window} uses of the particle overflow list in the last "
                                      "1000 timesteps, performance may be degraded.")
                    ovf_at_window, window_end = used, window_end + 1000
                elapsed = _time.perf_counter() - perf_t0
                if elapsed >= perf_iv:
                    step_s = elapsed / perf_steps
                    print(f"Average walltime per timestep: {1e3 * step_s:.2f} ms. Progress: {time:.4f}s/"
                          f"{maxTime:.4f}s ({100.0 * time / maxTime:.3g}% done / "
                          f"{step_s * (maxTime - time) / self.tau:.2g} seconds walltime remaining)",
                          file=sys.stderr, flush=True)
                    perf_t0, perf_steps = _time.perf_counter(), 0
                self._events(t, time, next_c, next_l, next_h, hook_enabled, device, lattice, dlat, rec)
                if hook_enabled and t >= next_h:
                    next_h += hi
                if t >= next_l:
                    next_l += lwi
                if t >= next_c:
                    next_c += wi
        except BaseException:
            # what was recorded so far still reaches the file
            try:
                rec.finish()
            except Exception:
                pass
            raise
        finally:
            for sig, h in restore.items():
                signal.signal(sig, h)
        if device:
            self._device_call(self.onEndTrajectory, None, dlat, validate=False)
        lattice.copyFromGPU()
        if not device:
            self.onEndTrajectory(lattice)
        return rec.finish()

    def _events(self, t, time, next_c, next_l, next_h, hook_enabled, device, lattice, dlat, rec) -> None:
        """The records of an event step: hook, then lattice frame, then species
        counts (MpdRdmeSolver.cu:612-659)."""
        # The lattice comes to the host for a host hook or a lattice frame;
        # counts alone are taken on the GPU (the same numbers as counting
        # the host copy, which then equals the device lattice).
        hook_due = hook_enabled and t >= next_h
        host = t >= next_l or (hook_due and not device)
        if host and not device:
            lattice.copyFromGPU()
        if hook_due:
            host |= self._hook_check(time, lattice, dlat, rec)
        if t >= next_l:
            if device:
                lattice.copyFromGPU()
            rec.lattice(time, lattice)
        if t >= next_c:
            rec.species_counts(time, lattice if host else None)

    def _device_call(self, fn, time, dlat: "DeviceHookLattice", validate: bool) -> int:
        """Calls a device-side hook with fresh device views, waits for the
        device work it queued, and checks the lattice it may have changed."""
        dlat.refresh()
        code = fn(dlat) if time is None else fn(time, dlat)
        dlat.synchronize()
        if validate or code in (1, 2):
            _lib.check(self._lib.lmx_validate_lattice(self._h))
        return code

    def _hook_check(self, time: float, lattice: "HookLattice", dlat: Optional["DeviceHookLattice"],
                    rec: "_Recorder") -> bool:
        """hookCheckSimulation (MpdRdmeSolver.cu:958-985).  With device hooks
        the changes are already on the device (codes 1 and 2 need no copy) and
        the sites of codes 2 and 3 come to the host to be recorded.  Returns
        whether the host copy was refreshed."""
        if dlat is not None:
            code = self._device_call(self.hookSimulation, time, dlat, validate=False)
            if code not in (0, 1, 2, 3):
                raise _lib.LmxError(_lib.ERR_ARG, f"Unknown hook return value {code}")
            if code in (2, 3):
                lattice.copyFromGPU()
                rec.sites(time, lattice)
                return True
            return False
        code = self.hookSimulation(time, lattice)
        if code == 0:
            return False
        if code == 1:
            lattice.copyToGPU()
        elif code == 2:
            lattice.copyToGPU()
            rec.sites(time, lattice)
        elif code == 3:
            rec.sites(time, lattice)
        else:
            raise _lib.LmxError(_lib.ERR_ARG, f"Unknown hook return value {code}")
        return False


def _last_step(tau: float, maxTime: float) -> int:
    """The step after which ``step * tau < maxTime`` first fails: the
    reference's loop (``while (time < maxTime)`` with ``time =
    current_timestep * tau``) runs exactly this many steps."""
    if not maxTime > 0.0:
        return 0
    k = max(int(np.ceil(maxTime / tau)), 1)
    while k > 1 and (k - 1) * tau >= maxTime:
        k -= 1
    while k * tau < maxTime:
        k += 1
    return k


class _HookCopies:
    """copyFromGPU/copyToGPU of a hook lattice: a host copy of the solver's
    particles and site types in the native planes."""

    def __init__(self, solver: MpdRdmeSolver):
        X, Y, Z = solver.shape
        super().__init__(X, Y, Z, solver.latticeSpacing, solver.particlesPerSite)
        self._solver = solver

    def copyFromGPU(self) -> None:
        self._words[:] = self._solver.getLatticeNative()
        self._sites[:] = np.transpose(self._solver.getSites(), (2, 1, 0))

    def copyToGPU(self) -> None:
        self._solver.setLatticeNative(self._words)
        self._solver.setSites(self.sites_row_major())


class HookLattice(_HookCopies, ByteLattice):
    """The lattice a hook sees, in place of ``lm::rdme::CudaByteLattice``: a
    host copy of the solver's particles and site types (``ByteLattice``
    accessors and views over it).  ``copyFromGPU``/``copyToGPU`` move it
    between the host and the solver; the trajectory loop copies it back when a
    hook returns 1 or 2."""


class IntHookLattice(_HookCopies, IntLattice):
    """The same for the uint32 lattice, in place of ``lm::rdme::CudaIntLattice``."""


class _DeviceArray:
    """__cuda_array_interface__ (version 3) of device memory the solver owns."""

    def __init__(self, ptr: int, shape, typestr: str):
        self.__cuda_array_interface__ = {"shape": tuple(int(v) for v in shape), "typestr": typestr,
                                         "data": (int(ptr), False), "version": 3, "strides": None}


class DeviceHookLattice:
    """What a device-side hook sees (a solver class with ``deviceHooks =
    True``): the solver's current device buffers as torch tensors, with no
    copies.  ``getParticleLatticeView()`` is the (W, Z, Y, X, 4) uint8 view of
    the byte lattice's particle planes, or the (P, Z, Y, X, 1) int32 view of
    the uint32 lattice's (int32, the 32-bit integer type torch computes with;
    the bits are the particle words), ``getSiteLatticeView()`` the (Z, Y, X)
    site types;
    both are valid during the hook call only (the solver alternates between
    two particle buffers).  Changes made with torch operations, on any stream,
    are complete before the next step (the solver waits for the device), and
    are checked on the device (particle types at most numberSpecies, site
    types below numberSiteTypes)."""

    def __init__(self, solver: MpdRdmeSolver):
        import torch

        if not torch.cuda.is_initialized():
            raise _lib.LmxError(_lib.ERR_STATE, "device hooks need lattice_microbes_amd.torch_interop() before the "
                                                "first solver of the process")
        self._torch = torch
        self._solver = solver
        self._dev = torch.device("cuda", solver.device)
        self._particles = self._sites = None

    def refresh(self) -> None:
        s = self._solver
        pp, sp = c_void_p(), c_void_p()
        _lib.check(s._lib.lmx_device_lattice(s._h, byref(pp), byref(sp)))
        X, Y, Z = s.shape
        torch = self._torch
        if s.bytesPerParticle == 4:
            shape, typestr = (s.particlesPerSite, Z, Y, X, 1), "<i4"
        else:
            shape, typestr = (s.particlesPerSite // 4, Z, Y, X, 4), "|u1"
        self._particles = torch.as_tensor(_DeviceArray(pp.value, shape, typestr), device=self._dev)
        self._sites = torch.as_tensor(_DeviceArray(sp.value, (Z, Y, X), "|u1"), device=self._dev)

    def synchronize(self) -> None:
        self._torch.cuda.synchronize(self._dev)
        self._particles = self._sites = None

    def getParticleLatticeView(self):
        return self._particles

    def getSiteLatticeView(self):
        return self._sites

    def getXSize(self) -> int:
        return self._solver.shape[0]

    def getYSize(self) -> int:
        return self._solver.shape[1]

    def getZSize(self) -> int:
        return self._solver.shape[2]

    def getSize(self):
        return tuple(self._solver.shape)

    def getMaxOccupancy(self) -> int:
        return self._solver.particlesPerSite

    def getSpacing(self) -> float:
        return self._solver.latticeSpacing

    def getParticleCounts(self) -> Dict[int, int]:
        c = self._torch.bincount(self._particles.reshape(-1).to(self._torch.int64),
                                 minlength=self._solver.numberSpecies + 1).cpu().numpy()
        return {int(p): int(c[p]) for p in np.flatnonzero(c) if p != 0}


class _Recorder:
    """What the reference's solver pushes to its output queue during a
    trajectory (MpdRdmeSolver.cu:824-956), kept in memory and, with a file,
    written to it."""

    def __init__(self, solver: MpdRdmeSolver, output):
        self.solver = solver
        self.out = output
        self.counts_t: List[float] = []
        self.counts: List[np.ndarray] = []
        self.lattice_t: List[float] = []
        self.lattices: List[np.ndarray] = []
        self.site_t: List[float] = []
        self.site_frames: List[np.ndarray] = []
        self.max_particles = np.zeros(solver.numberSpecies + 1, np.uint32)
        self.max_sites = np.zeros(solver.numberSiteTypes, np.uint32)

    def _particle_counts(self, lattice: HookLattice) -> np.ndarray:
        c = np.bincount(lattice.getParticleLatticeView().reshape(-1),
                        minlength=self.max_particles.size)[: self.max_particles.size]
        c[0] = 0
        return c.astype(np.uint32)

    def _site_counts(self, lattice: HookLattice) -> np.ndarray:
        c = np.bincount(lattice.getSiteLatticeView().reshape(-1), minlength=256)[: self.max_sites.size]
        if c.size:
            c[0] = 0
        return c.astype(np.uint32)

    def init_max_counts(self, lattice: HookLattice) -> None:
        np.maximum(self.max_sites, self._site_counts(lattice), out=self.max_sites)
        np.maximum(self.max_particles, self._particle_counts(lattice), out=self.max_particles)

    def species_counts(self, time: float, lattice: Optional[HookLattice] = None) -> None:
        """recordSpeciesCounts (MpdRdmeSolver.cu:933-942): particles of types
        1..numberSpecies, from the host copy when there is one."""
        self.counts_t.append(time)
        if lattice is None:
            self.counts.append(self.solver.getSpeciesCounts())
        else:
            n = self.solver.numberSpecies
            c = np.bincount(lattice.getParticleLatticeView().reshape(-1), minlength=n + 1)
            self.counts.append(c[1:n + 1].astype(np.uint32))

    def lattice(self, time: float, lattice: HookLattice) -> None:
        np.maximum(self.max_particles, self._particle_counts(lattice), out=self.max_particles)
        frame = lattice.row_major()
        self.lattice_t.append(time)
        self.lattices.append(frame)
        if self.out is not None:
            self.out.appendLattice(self.solver.replicate, time, frame)

    def sites(self, time: float, lattice: HookLattice) -> None:
        np.maximum(self.max_sites, self._site_counts(lattice), out=self.max_sites)
        frame = lattice.sites_row_major()
        self.site_t.append(time)
        self.site_frames.append(frame)
        if self.out is not None:
            self.out.appendSites(self.solver.replicate, time, frame)

    def finish(self) -> Dict[str, object]:
        counts = np.array(self.counts, dtype=np.uint32).reshape(len(self.counts), self.solver.numberSpecies)
        if self.out is not None:
            r = self.solver.replicate
            if len(self.counts):
                self.out.appendSpeciesCounts(r, np.array(self.counts_t), counts)
            self.out.writeReplicateDataset(r, "MaxParticleCounts", self.max_particles)
            if self.max_sites.size:
                self.out.writeReplicateDataset(r, "MaxSiteCounts", self.max_sites)
        return {"species_count_times": np.array(self.counts_t), "species_counts": counts,
                "lattice_times": np.array(self.lattice_t), "lattices": self.lattices,
                "site_times": np.array(self.site_t), "sites": self.site_frames,
                "max_particle_counts": self.max_particles.copy(), "max_site_counts": self.max_sites.copy()}


class IntMpdRdmeSolver(MpdRdmeSolver):
    """Drop-in for lm::rdme::IntMpdRdmeSolver (IntMpdRdmeSolver.cu): the
    same MPD-RDME step on the uint32 lattice, one particle per 32-bit word
    (P planes of N words), so that models may have more than 255 species
    (allocateLattice, :155-170, requires 4 bytes per particle).  The update
    rules and RNG keys are those of the byte lattice
    (word_diffusion_1d_dev.cu:106 keys word w as the byte kernels key slot w),
    so both solvers give the same trajectory for a model that fits a byte.
    Hooks receive an ``IntHookLattice`` (``IntLattice`` accessors); lattice
    frames are written with 4 bytes per particle."""

    bytesPerParticle = 4
    _parses_overflow_handler = False


class MESolverFactory:
    """String-based selection (MESolverFactory.cpp:75-96,115-190)."""

    _solvers: Dict[str, Callable[..., MpdRdmeSolver]] = {SOLVER_NAME: MpdRdmeSolver,
                                                          INT_SOLVER_NAME: IntMpdRdmeSolver}

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
