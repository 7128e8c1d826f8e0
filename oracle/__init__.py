"""ctypes access to the serial C oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module.  The product (lattice_microbes_amd/) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, Structure, c_double, c_float, c_int, c_uint8, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

ORC_OK = 0
ORC_ERR_ARG = 1
ORC_ERR_OVERFLOW_CAP = 2
ORC_ERR_OVERFLOW_UNPLACEABLE = 3
ORC_ERR_OVERFLOW_OUT_OF_RANGE = 4
ORC_ERR_REACTION_UNDERFLOW = 5

# oracle code -> lmx_status code of the product
TO_LMX = {ORC_OK: 0, ORC_ERR_ARG: 1, ORC_ERR_OVERFLOW_CAP: 3, ORC_ERR_OVERFLOW_UNPLACEABLE: 4,
          ORC_ERR_OVERFLOW_OUT_OF_RANGE: 5, ORC_ERR_REACTION_UNDERFLOW: 6}


class OrcModel(Structure):
    _fields_ = [
        ("X", c_uint32), ("Y", c_uint32), ("Z", c_uint32), ("P", c_uint32),
        ("nS", c_uint32), ("nST", c_uint32), ("R", c_uint32), ("boundary", c_uint32),
        ("seed32", c_uint32),
        ("T", POINTER(c_float)), ("qp0", POINTER(c_float)), ("qp1", POINTER(c_float)),
        ("qp2", POINTER(c_float)), ("rate", POINTER(c_float)),
        ("order", POINTER(c_uint32)), ("D1", POINTER(c_uint32)), ("D2", POINTER(c_uint32)),
        ("S", POINTER(ctypes.c_int8)), ("RL", POINTER(c_uint8)),
        ("bpp", c_uint32),
    ]


class OrcStats(Structure):
    _fields_ = [("overflow_entries", c_uint64), ("overflow_steps", c_uint64),
                ("reactions", c_uint64), ("last_error_step", c_uint32), ("pad", c_uint32)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    u32p, f32p, f64p, u8p = POINTER(c_uint32), POINTER(c_float), POINTER(c_double), POINTER(c_uint8)
    lib.orc_hash.restype = c_uint32
    lib.orc_hash.argtypes = [c_uint64, c_uint64]
    lib.orc_timestep_hash.restype = c_uint64
    lib.orc_timestep_hash.argtypes = [c_uint32, c_uint32, c_uint32]
    lib.orc_uniform.restype = c_float
    lib.orc_uniform.argtypes = [c_uint32, c_uint32, c_uint32, c_uint64]
    lib.orc_exp.restype = c_double
    lib.orc_exp.argtypes = [c_double]
    lib.orc_reaction_probability.restype = c_float
    lib.orc_reaction_probability.argtypes = [c_float]
    lib.orc_build_reaction_tables.restype = c_int
    lib.orc_build_reaction_tables.argtypes = [c_uint32, c_uint32, u32p, f64p, c_uint32,
                                              POINTER(ctypes.c_int32), u32p, u32p, u32p, u32p,
                                              POINTER(ctypes.c_int8), u32p, u32p, f64p]
    lib.orc_build_T.restype = c_int
    lib.orc_build_T.argtypes = [c_uint32, c_uint32, f64p, c_double, c_double, f32p]
    lib.orc_compute_propensities.restype = c_int
    lib.orc_compute_propensities.argtypes = [c_uint32, c_uint32, c_uint32, u32p, f64p, u32p, u32p, u32p,
                                             c_double, c_uint32, c_uint32, c_uint32, f32p, f32p, f32p, f32p]
    lib.orc_lattice_from_row_major.restype = c_int
    lib.orc_lattice_from_row_major.argtypes = [c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, u8p, u32p]
    lib.orc_lattice_to_row_major.restype = None
    lib.orc_lattice_to_row_major.argtypes = [c_uint32, c_uint32, c_uint32, c_uint32, u32p, u8p]
    lib.orc_sites_from_row_major.restype = c_int
    lib.orc_sites_from_row_major.argtypes = [c_uint32, c_uint32, c_uint32, c_uint32, u8p, u8p]
    lib.orc_species_counts.restype = None
    lib.orc_species_counts.argtypes = [c_uint32, c_uint32, u32p, c_uint32, u32p]
    lib.orc_sweep.restype = c_int
    lib.orc_lattice_from_row_major_words.restype = c_int
    lib.orc_lattice_from_row_major_words.argtypes = [c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, u32p, u32p]
    lib.orc_lattice_to_row_major_words.restype = None
    lib.orc_lattice_to_row_major_words.argtypes = [c_uint32, c_uint32, c_uint32, c_uint32, u32p, u32p]
    lib.orc_species_counts_words.restype = None
    lib.orc_species_counts_words.argtypes = [c_uint32, c_uint32, u32p, c_uint32, u32p]
    lib.orc_sweep.argtypes = [POINTER(OrcModel), u32p, u8p, u32p, c_uint32, c_uint64, c_int, u32p, u32p, u32p]
    lib.orc_react.restype = c_int
    lib.orc_react.argtypes = [POINTER(OrcModel), u32p, u8p, c_uint64, u32p, u32p, u32p, POINTER(c_uint64)]
    lib.orc_resolve_overflows.restype = c_int
    lib.orc_resolve_overflows.argtypes = [POINTER(OrcModel), u32p, u8p, u32p, u32p, c_uint32]
    lib.orc_resolve_relaxed.restype = c_int
    lib.orc_resolve_relaxed.argtypes = [POINTER(OrcModel), u32p, u32p, u32p, c_uint32, u32p, u32p, u32p]
    lib.orc_run_handler.restype = c_int
    lib.orc_run_handler.argtypes = [POINTER(OrcModel), u32p, u8p, c_uint32, c_uint32, c_int, c_int, u32p, u32p,
                                    u32p, POINTER(OrcStats)]
    lib.orc_run.restype = c_int
    lib.orc_run.argtypes = [POINTER(OrcModel), u32p, u8p, c_uint32, c_uint32, c_int, POINTER(OrcStats)]
    lib.orc_nsm_create.restype = c_void_p
    lib.orc_nsm_create.argtypes = [c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, c_uint32, c_uint32,
                                   u32p, f64p, u32p, u32p, POINTER(ctypes.c_int32), f64p, u32p, c_double,
                                   u32p, u8p, c_uint32, c_uint32]
    lib.orc_nsm_destroy.restype = None
    lib.orc_nsm_destroy.argtypes = [c_void_p]
    lib.orc_nsm_run.restype = c_int
    lib.orc_nsm_run.argtypes = [c_void_p, c_double, c_uint64, c_double, c_uint32, u32p, u32p]
    lib.orc_nsm_time.restype = c_double
    lib.orc_nsm_time.argtypes = [c_void_p]
    lib.orc_nsm_events.restype = c_uint64
    lib.orc_nsm_events.argtypes = [c_void_p]
    lib.orc_nsm_reaction_events.restype = c_uint64
    lib.orc_nsm_reaction_events.argtypes = [c_void_p]
    lib.orc_nsm_counts.restype = None
    lib.orc_nsm_counts.argtypes = [c_void_p, u32p]
    lib.orc_nsm_lattice.restype = None
    lib.orc_nsm_lattice.argtypes = [c_void_p, u32p]
    lib.orc_heap_sequence.restype = c_int
    lib.orc_heap_sequence.argtypes = [c_uint32, c_uint32, u32p, f64p, u32p, u32p]
    _lib = lib
    return lib


def _p(a, ct):
    return a.ctypes.data_as(POINTER(ct))


def seed32(seed_top: int, replicate: int) -> int:
    """MpdRdmeSolver::initialize seed word (MpdRdmeSolver.cu:123)."""
    return ((seed_top << 16) & 0xFFFFFFFF) | (replicate & 0xFFFF)


class Oracle:
    """Tables + state for one model, built from the same arrays the solver
    receives (a lattice_microbes_amd.models.ModelSpec or compatible).  A
    uint32 spec lattice selects the uint32 lattice (one particle per word,
    IntMpdRdmeSolver); otherwise the byte lattice."""

    def __init__(self, spec, seed_top: int = 1, replicate: int = 0, overflow: str = "classic"):
        if overflow not in ("classic", "relaxed"):
            raise ValueError(f"overflow handler {overflow!r}")
        lib = load()
        self.relaxed = overflow == "relaxed"
        self.carry_site = np.zeros(2048, np.uint32)
        self.carry_particle = np.zeros(2048, np.uint32)
        self.carry_count = np.zeros(1, np.uint32)
        self.lib = lib
        self.spec = spec
        X, Y, Z = spec.shape
        self.shape = (X, Y, Z)
        self.N = X * Y * Z
        self.P = spec.P
        self.words = np.asarray(spec.lattice).dtype == np.uint32
        self.W = self.P if self.words else self.P // 4
        nS, nST, R = spec.nS, spec.nST, spec.R
        types = np.ascontiguousarray(spec.types, dtype=np.uint32)
        K = np.ascontiguousarray(spec.K, dtype=np.float64).reshape(-1)
        S = np.ascontiguousarray(spec.S, dtype=np.int32).reshape(-1)
        D = np.ascontiguousarray(spec.D, dtype=np.uint32).reshape(-1)
        self.order = np.zeros(max(R, 1), np.uint32)
        self.D1 = np.zeros(max(R, 1), np.uint32)
        self.D2 = np.zeros(max(R, 1), np.uint32)
        self.S_rm = np.zeros(max(R * nS, 1), np.int8)
        self.si = np.zeros(max(R, 1), np.uint32)
        self.s2i = np.zeros(max(R, 1), np.uint32)
        self.k = np.zeros(max(R, 1), np.float64)
        rc = lib.orc_build_reaction_tables(nS, R, _p(types, c_uint32), _p(K, c_double), 1,
                                           _p(S, ctypes.c_int32), _p(D, c_uint32), _p(self.order, c_uint32),
                                           _p(self.D1, c_uint32), _p(self.D2, c_uint32),
                                           _p(self.S_rm, ctypes.c_int8), _p(self.si, c_uint32),
                                           _p(self.s2i, c_uint32), _p(self.k, c_double))
        if rc:
            raise ValueError(f"oracle reaction tables rejected the model (code {rc})")
        DF = np.ascontiguousarray(spec.DF, dtype=np.float64).reshape(-1)
        self.T = np.zeros(max(DF.size, 1), np.float32)
        rc = lib.orc_build_T(nS, nST, _p(DF, c_double), spec.tau, spec.spacing, _p(self.T, c_float))
        if rc:
            raise ValueError("oracle T table rejected the model (q > 0.5)")
        self.RL32 = np.ascontiguousarray(spec.RL, dtype=np.uint32).reshape(-1)
        if self.RL32.size == 0:
            self.RL32 = np.zeros(1, np.uint32)
        self.RL8 = self.RL32.astype(np.uint8)
        self.rate = np.zeros(max(R, 1), np.float32)
        self.qp0 = np.zeros(max(nST, 1), np.float32)
        self.qp1 = np.zeros(max(nST * nS, 1), np.float32)
        self.qp2 = np.zeros(max(nST * nS * nS, 1), np.float32)
        self.types = types if types.size else np.zeros(1, np.uint32)
        rc = lib.orc_compute_propensities(nS, nST, R, _p(self.types, c_uint32), _p(self.k, c_double),
                                          _p(self.si, c_uint32), _p(self.s2i, c_uint32),
                                          _p(self.RL32, c_uint32), spec.tau, X, Y, Z,
                                          _p(self.rate, c_float), _p(self.qp0, c_float),
                                          _p(self.qp1, c_float), _p(self.qp2, c_float))
        if rc:
            raise ValueError("oracle propensity tables rejected the model")
        self.lattice = np.zeros(self.N * self.W, np.uint32)
        if self.words:
            lat_rm = np.ascontiguousarray(spec.lattice, dtype=np.uint32).reshape(-1)
            rc = lib.orc_lattice_from_row_major_words(X, Y, Z, self.P, nS, _p(lat_rm, c_uint32),
                                                      _p(self.lattice, c_uint32))
        else:
            lat_rm = np.ascontiguousarray(spec.lattice, dtype=np.uint8).reshape(-1)
            rc = lib.orc_lattice_from_row_major(X, Y, Z, self.P, nS, _p(lat_rm, c_uint8),
                                                _p(self.lattice, c_uint32))
        if rc:
            raise ValueError("oracle rejected the lattice")
        sites_rm = np.ascontiguousarray(spec.sites, dtype=np.uint8).reshape(-1)
        self.sites = np.zeros(self.N, np.uint8)
        rc = lib.orc_sites_from_row_major(X, Y, Z, nST, _p(sites_rm, c_uint8), _p(self.sites, c_uint8))
        if rc:
            raise ValueError("oracle rejected the site types")
        self.model = OrcModel()
        m = self.model
        m.X, m.Y, m.Z, m.P = X, Y, Z, self.P
        m.nS, m.nST, m.R = nS, nST, R
        m.boundary = 1 if getattr(spec, "boundary", "periodic") == "reflecting" else 0
        m.seed32 = seed32(seed_top, replicate)
        m.T = _p(self.T, c_float)
        m.qp0, m.qp1, m.qp2 = _p(self.qp0, c_float), _p(self.qp1, c_float), _p(self.qp2, c_float)
        m.rate = _p(self.rate, c_float)
        m.order, m.D1, m.D2 = _p(self.order, c_uint32), _p(self.D1, c_uint32), _p(self.D2, c_uint32)
        m.S = _p(self.S_rm, ctypes.c_int8)
        m.RL = _p(self.RL8, c_uint8)
        m.bpp = 4 if self.words else 1
        self.timestep = 0
        self.stats = OrcStats()

    @property
    def seed32(self) -> int:
        return int(self.model.seed32)

    def run(self, nsteps: int, formulation: int = 0) -> int:
        rc = self.lib.orc_run_handler(ctypes.byref(self.model), _p(self.lattice, c_uint32),
                                      _p(self.sites, c_uint8), self.timestep, nsteps, formulation,
                                      int(self.relaxed), _p(self.carry_site, c_uint32),
                                      _p(self.carry_particle, c_uint32), _p(self.carry_count, c_uint32),
                                      ctypes.byref(self.stats))
        if rc == 0:
            self.timestep += nsteps
        return rc

    @property
    def unresolved(self) -> int:
        """Overflow entries the relaxed handler is carrying to the next step."""
        return int(self.carry_count[0])

    def native(self) -> np.ndarray:
        X, Y, Z = self.shape
        return self.lattice.reshape(self.W, Z, Y, X).copy()

    def row_major(self) -> np.ndarray:
        X, Y, Z = self.shape
        if self.words:
            out = np.zeros(self.N * self.P, np.uint32)
            self.lib.orc_lattice_to_row_major_words(X, Y, Z, self.P, _p(self.lattice, c_uint32), _p(out, c_uint32))
        else:
            out = np.zeros(self.N * self.P, np.uint8)
            self.lib.orc_lattice_to_row_major(X, Y, Z, self.P, _p(self.lattice, c_uint32), _p(out, c_uint8))
        return out.reshape(X, Y, Z, self.P)

    def species_counts(self) -> np.ndarray:
        out = np.zeros(max(self.spec.nS, 1), np.uint32)
        fn = self.lib.orc_species_counts_words if self.words else self.lib.orc_species_counts
        fn(self.N, self.P, _p(self.lattice, c_uint32), self.spec.nS, _p(out, c_uint32))
        return out[: self.spec.nS]


def hash32(v: int, seed: int) -> int:
    return int(load().orc_hash(v, seed))


def timestep_hash(seed32_: int, t: int, s: int) -> int:
    return int(load().orc_timestep_hash(seed32_, t, s))


def uniform(v1: int, shift: int, v2: int, seed: int) -> float:
    return float(load().orc_uniform(v1, shift, v2, seed))


def exp(x: float) -> float:
    return float(load().orc_exp(x))


def reaction_probability(a: float) -> float:
    return float(load().orc_reaction_probability(a))
