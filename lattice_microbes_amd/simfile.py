"""Python mirror of ``lm::io::hdf5::SimulationFile`` over the lmx_io C ABI.

Method names and argument meaning follow the reference class
(src/io/SimulationFile.h, exposed to Python by src/swig/lm.i): ``create``,
``getParameters``/``getParameter``/``setParameter``, ``getReactionModel``/
``setReactionModel``, ``getDiffusionModel``/``getDiffusionModelLattice``/
``setDiffusionModel``/``setDiffusionModelLattice``, ``appendSpeciesCounts``,
``appendLattice``, ``appendSites``, ``replicateExists``, ``getSpeciesCounts``,
``getSpeciesCountTimes``, ``getLatticeTimes``, ``getLattice``.  The reference
passes protobuf ``ReactionModel``/``DiffusionModel`` messages; here they are
the dataclasses below with the same fields.  Every call goes to
liblmx_io.so (HDF5), which must have been built.
"""
from __future__ import annotations

import ctypes
import os
import shutil
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_size_t, c_uint32, c_void_p
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.path.join(_HERE, "_build", "liblmx_io.so")

OK, ERR_ARG, ERR_IO, ERR_FORMAT = 0, 1, 2, 3
MAX_RATE_CONSTANTS = 10
NOISE_COLS = 2


class _ReactionInfo(ctypes.Structure):
    _fields_ = [("present", c_uint32), ("numberSpecies", c_uint32), ("hasTables", c_uint32),
                ("numberReactions", c_uint32), ("hasRateNoise", c_uint32)]


class _DiffusionInfo(ctypes.Structure):
    _fields_ = [("present", c_uint32), ("numberSpecies", c_uint32), ("numberReactions", c_uint32),
                ("numberSiteTypes", c_uint32), ("X", c_uint32), ("Y", c_uint32), ("Z", c_uint32), ("P", c_uint32),
                ("bytesPerParticle", c_uint32), ("latticeSpacing", c_double)]


_u32p, _i32p, _f64p, _u8p = POINTER(c_uint32), POINTER(ctypes.c_int32), POINTER(c_double), POINTER(ctypes.c_uint8)

SIGNATURES = {
    "lmxio_last_error": (c_char_p, []),
    "lmxio_create": (c_int, [c_char_p, c_int, c_uint32]),
    "lmxio_flush": (c_int, [c_void_p]),
    "lmxio_open": (c_int, [c_char_p, c_int, POINTER(c_void_p)]),
    "lmxio_close": (c_int, [c_void_p]),
    "lmxio_format_version": (c_int, [c_void_p, _u32p]),
    "lmxio_get_parameters": (c_int, [c_void_p, c_char_p, c_size_t, POINTER(c_size_t)]),
    "lmxio_set_parameter": (c_int, [c_void_p, c_char_p, c_char_p]),
    "lmxio_get_reaction_info": (c_int, [c_void_p, POINTER(_ReactionInfo)]),
    "lmxio_get_reaction_model": (c_int, [c_void_p, _u32p, _u32p, _f64p, _f64p, _i32p, _u32p]),
    "lmxio_set_reaction_model": (c_int, [c_void_p, c_uint32, c_uint32, _u32p, _u32p, _f64p, _f64p, _i32p, _u32p]),
    "lmxio_get_diffusion_info": (c_int, [c_void_p, POINTER(_DiffusionInfo)]),
    "lmxio_get_diffusion_model": (c_int, [c_void_p, _f64p, _u32p, c_void_p, c_size_t, _u8p, c_size_t]),
    "lmxio_set_diffusion_model": (c_int, [c_void_p, c_uint32, c_uint32, c_uint32, _f64p, _u32p, c_double, c_uint32,
                                          c_uint32, c_uint32, c_uint32, c_uint32, c_void_p, _u8p]),
    "lmxio_set_diffusion_model_lattice": (c_int, [c_void_p, c_uint32, c_uint32, c_uint32, c_uint32, c_uint32,
                                                  c_void_p, _u8p]),
    "lmxio_set_empty_spatial_model": (c_int, [c_void_p]),
    "lmxio_append_species_counts": (c_int, [c_void_p, c_uint32, c_uint32, c_uint32, _f64p, _u32p]),
    "lmxio_append_lattice": (c_int, [c_void_p, c_uint32, c_double, c_uint32, c_uint32, c_uint32, c_uint32, c_uint32,
                                     c_void_p]),
    "lmxio_append_sites": (c_int, [c_void_p, c_uint32, c_double, c_uint32, c_uint32, c_uint32, _u8p]),
    "lmxio_replicate_exists": (c_int, [c_void_p, c_uint32, POINTER(c_int)]),
    "lmxio_get_species_counts_size": (c_int, [c_void_p, c_uint32, _u32p, _u32p]),
    "lmxio_get_species_counts": (c_int, [c_void_p, c_uint32, _f64p, _u32p]),
    "lmxio_get_lattice_count": (c_int, [c_void_p, c_uint32, _u32p]),
    "lmxio_get_lattice": (c_int, [c_void_p, c_uint32, c_uint32, _f64p, c_void_p, c_size_t, _u32p]),
    "lmxio_new_replicate_dataset_u32": (c_int, [c_void_p, c_uint32, c_char_p, c_uint32, _u32p]),
    "lmxio_get_replicate_dataset_u32": (c_int, [c_void_p, c_uint32, c_char_p, _u32p, c_uint32, _u32p]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load liblmx_io.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(IO_LIB_PATH):
        raise RuntimeError(f"lmx_io library not found at {IO_LIB_PATH}; build it with "
                           "`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(IO_LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class IOException(RuntimeError):
    """lm::IOException / lm::Exception of the reference reader and writer."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[lmx_io status {code}] {message}")
        self.code = code


class InvalidArgException(IOException):
    pass


def _check(code: int) -> None:
    if code == OK:
        return
    msg = load().lmxio_last_error().decode(errors="replace")
    if code == ERR_ARG:
        raise InvalidArgException(code, msg)
    raise IOException(code, msg)


def _p(a: np.ndarray, ctype):
    return a.ctypes.data_as(POINTER(ctype)) if a is not None else None


@dataclass
class ReactionModel:
    """Fields of the ReactionModel message (src/protobuf/ReactionModel.proto)."""
    number_species: int = 0
    number_reactions: int = 0
    initial_species_count: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    types: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    rate_constants: List[List[float]] = field(default_factory=list)
    # (variance, tau) per reaction; a reaction has noise iff both are > 0
    rate_noise: Optional[np.ndarray] = None
    S: np.ndarray = field(default_factory=lambda: np.zeros((0, 0), np.int32))   # [nS, R]
    D: np.ndarray = field(default_factory=lambda: np.zeros((0, 0), np.uint32))  # [nS, R]

    def rate_has_noise(self, r: int) -> bool:
        if self.rate_noise is None:
            return False
        v, t = self.rate_noise[r]
        return bool(v > 0.0 and t > 0.0 and not np.isnan(v) and not np.isnan(t))

    def rate_matrix(self) -> np.ndarray:
        """[R, 10] with NaN past each reaction's constants, as stored."""
        K = np.full((self.number_reactions, MAX_RATE_CONSTANTS), np.nan)
        for r, ks in enumerate(self.rate_constants):
            K[r, :len(ks)] = ks
        return K


@dataclass
class DiffusionModel:
    """Fields of the DiffusionModel message (src/protobuf/DiffusionModel.proto)."""
    number_species: int = 0
    number_reactions: int = 0
    number_site_types: int = 0
    lattice_spacing: float = 0.0
    lattice_x_size: int = 0
    lattice_y_size: int = 0
    lattice_z_size: int = 0
    particles_per_site: int = 0
    bytes_per_particle: int = 1
    diffusion_matrix: Optional[np.ndarray] = None          # [nST, nST, nS] float64
    reaction_location_matrix: Optional[np.ndarray] = None  # [R, nST] uint32

    @property
    def shape(self) -> Tuple[int, int, int]:
        return (self.lattice_x_size, self.lattice_y_size, self.lattice_z_size)


class SimulationFile:
    """An open Lattice Microbes HDF5 file (format versions 2-4)."""

    CURRENT_VERSION = 4
    MIN_VERSION = 2

    @staticmethod
    def create(filename: str, numberSpecies: Optional[int] = None) -> None:
        """A new version-4 file, with a reaction model of numberSpecies
        species when given; fails if the file exists
        (SimulationFile_create.cpp:56-122)."""
        init = numberSpecies is not None
        _check(load().lmxio_create(os.fsencode(filename), int(init), int(numberSpecies or 0)))

    @staticmethod
    def isValidFile(filename: str) -> bool:
        try:
            SimulationFile(filename, writable=False).close()
            return True
        except IOException:
            return False

    def __init__(self, filename: str, writable: bool = True):
        self._lib = load()
        self.filename = filename
        self.writable = bool(writable)
        self._open()

    def _open(self) -> None:
        h = c_void_p()
        _check(self._lib.lmxio_open(os.fsencode(self.filename), int(self.writable), byref(h)))
        self._h = h
        v = c_uint32()
        _check(self._lib.lmxio_format_version(self._h, byref(v)))
        self.version = v.value

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _check(self._lib.lmxio_close(self._h))
            self._h = None

    def flush(self) -> None:
        _check(self._lib.lmxio_flush(self._h))

    def checkpoint(self) -> str:
        """Copies the file to ``filename + ".chk"`` with everything written so
        far and reopens it (SimulationFile.cpp:158-188)."""
        self.close()
        chk = self.filename + ".chk"
        try:
            shutil.copyfile(self.filename, chk)
        finally:
            self._open()
        return chk

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---------------------------------------------------------------- parameters
    def getParameters(self) -> Dict[str, str]:
        n = c_size_t()
        _check(self._lib.lmxio_get_parameters(self._h, None, 0, byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(self._lib.lmxio_get_parameters(self._h, buf, n.value, byref(n)))
        out: Dict[str, str] = {}
        for line in buf.value.decode().splitlines():
            k, _, v = line.partition("=")
            out[k] = v
        return out

    def getParameter(self, key: str, defaultValue: str = "") -> str:
        return self.getParameters().get(key, defaultValue)

    def setParameter(self, key: str, value: str) -> None:
        _check(self._lib.lmxio_set_parameter(self._h, key.encode(), str(value).encode()))

    # ---------------------------------------------------------------- models
    def getReactionModel(self) -> ReactionModel:
        info = _ReactionInfo()
        _check(self._lib.lmxio_get_reaction_info(self._h, byref(info)))
        m = ReactionModel()
        if not info.present:
            return m
        m.number_species = info.numberSpecies
        if not info.hasTables:
            return m
        nS, R = info.numberSpecies, info.numberReactions
        m.number_reactions = R
        ic = np.zeros(nS, np.uint32)
        types = np.zeros(R, np.uint32)
        K = np.zeros((R, MAX_RATE_CONSTANTS))
        noise = np.zeros((R, NOISE_COLS))
        S = np.zeros((nS, R), np.int32)
        D = np.zeros((nS, R), np.uint32)
        _check(self._lib.lmxio_get_reaction_model(self._h, _p(ic, c_uint32), _p(types, c_uint32), _p(K, c_double),
                                                  _p(noise, c_double), _p(S, ctypes.c_int32), _p(D, c_uint32)))
        m.initial_species_count, m.types, m.S, m.D = ic, types, S, D
        # constants up to the first NaN (SimulationFile.cpp:434-441)
        m.rate_constants = []
        for r in range(R):
            row = []
            for k in K[r]:
                if np.isnan(k):
                    break
                row.append(float(k))
            m.rate_constants.append(row)
        m.rate_noise = noise if info.hasRateNoise else None
        return m

    def setReactionModel(self, m: ReactionModel) -> None:
        nS, R = int(m.number_species), int(m.number_reactions)
        ic = np.ascontiguousarray(m.initial_species_count, dtype=np.uint32).reshape(-1)
        if nS == 0 or ic.size != nS or len(m.types) != R:
            raise InvalidArgException(ERR_ARG, "reactionModel: inconsistent sizes")
        types = np.ascontiguousarray(m.types, dtype=np.uint32).reshape(-1)
        K = np.ascontiguousarray(m.rate_matrix())
        S = np.ascontiguousarray(m.S, dtype=np.int32).reshape(-1)
        D = np.ascontiguousarray(m.D, dtype=np.uint32).reshape(-1)
        if S.size != nS * R or D.size != nS * R:
            raise InvalidArgException(ERR_ARG, "reactionModel: inconsistent matrix sizes")
        noise = None
        if m.rate_noise is not None:
            noise = np.ascontiguousarray(m.rate_noise, dtype=np.float64).reshape(R, NOISE_COLS)
        _check(self._lib.lmxio_set_reaction_model(self._h, nS, R, _p(ic, c_uint32), _p(types, c_uint32),
                                                  _p(K, c_double), _p(noise, c_double), _p(S, ctypes.c_int32),
                                                  _p(D, c_uint32)))

    def _diffusion_info(self) -> _DiffusionInfo:
        info = _DiffusionInfo()
        _check(self._lib.lmxio_get_diffusion_info(self._h, byref(info)))
        return info

    def getDiffusionModel(self) -> DiffusionModel:
        info = self._diffusion_info()
        d = DiffusionModel()
        if not info.present:
            return d
        d.number_species, d.number_reactions, d.number_site_types = (info.numberSpecies, info.numberReactions,
                                                                      info.numberSiteTypes)
        d.lattice_spacing = info.latticeSpacing
        d.lattice_x_size, d.lattice_y_size, d.lattice_z_size = info.X, info.Y, info.Z
        d.particles_per_site, d.bytes_per_particle = info.P, info.bytesPerParticle
        if info.numberSpecies and info.numberSiteTypes:
            DF = np.zeros((info.numberSiteTypes, info.numberSiteTypes, info.numberSpecies))
            RL = np.zeros((info.numberReactions, info.numberSiteTypes), np.uint32)
            _check(self._lib.lmxio_get_diffusion_model(self._h, _p(DF, c_double), _p(RL, c_uint32), None, 0, None, 0))
            d.diffusion_matrix, d.reaction_location_matrix = DF, RL
        return d

    def getDiffusionModelLattice(self, m: Optional[DiffusionModel] = None) -> Tuple[np.ndarray, np.ndarray]:
        """(lattice [X, Y, Z, P] uint8 or uint32, sites [X, Y, Z] uint8), row-major as stored."""
        info = self._diffusion_info()
        if not info.present:
            raise IOException(ERR_FORMAT, "the file has no diffusion model")
        dt = np.uint8 if info.bytesPerParticle == 1 else np.uint32
        lat = np.zeros((info.X, info.Y, info.Z, info.P), dt)
        sites = np.zeros((info.X, info.Y, info.Z), np.uint8)
        _check(self._lib.lmxio_get_diffusion_model(self._h, None, None, lat.ctypes.data_as(c_void_p), lat.nbytes,
                                                   _p(sites, ctypes.c_uint8), sites.nbytes))
        return lat, sites

    def setDiffusionModel(self, m: DiffusionModel, lattice: Optional[np.ndarray] = None,
                          sites: Optional[np.ndarray] = None) -> None:
        """setDiffusionModel (replaces /Model/Diffusion), and with lattice
        and sites also setDiffusionModelLattice."""
        nST, nS, R = int(m.number_site_types), int(m.number_species), int(m.number_reactions)
        DF = np.ascontiguousarray(m.diffusion_matrix, dtype=np.float64).reshape(-1)
        if DF.size != nST * nST * nS:
            raise InvalidArgException(ERR_ARG, "diffusion_matrix: inconsistent size")
        RL = np.ascontiguousarray(m.reaction_location_matrix if R else np.zeros(0), dtype=np.uint32).reshape(-1)
        if RL.size != R * nST:
            raise InvalidArgException(ERR_ARG, "reaction_location_matrix: inconsistent size")
        X, Y, Z, P = m.lattice_x_size, m.lattice_y_size, m.lattice_z_size, m.particles_per_site
        bpp = int(m.bytes_per_particle)
        lat_p = site_p = None
        if lattice is not None:
            lattice = np.ascontiguousarray(lattice, dtype=np.uint8 if bpp == 1 else np.uint32)
            if lattice.shape != (X, Y, Z, P):
                raise InvalidArgException(ERR_ARG, f"lattice shape {lattice.shape} is not {(X, Y, Z, P)}")
            lat_p = lattice.ctypes.data_as(c_void_p)
        if sites is not None:
            sites = np.ascontiguousarray(sites, dtype=np.uint8)
            if sites.shape != (X, Y, Z):
                raise InvalidArgException(ERR_ARG, f"sites shape {sites.shape} is not {(X, Y, Z)}")
            site_p = _p(sites, ctypes.c_uint8)
        _check(self._lib.lmxio_set_diffusion_model(self._h, nS, R, nST, _p(DF, c_double), _p(RL, c_uint32),
                                                   float(m.lattice_spacing), X, Y, Z, P, bpp, lat_p, site_p))

    def setDiffusionModelLattice(self, m: DiffusionModel, lattice: np.ndarray, sites: np.ndarray) -> None:
        """The initial lattice [X, Y, Z, P] and sites [X, Y, Z] (row-major) of
        the diffusion model already in the file (SimulationFile.cpp:833-870)."""
        X, Y, Z, P = m.lattice_x_size, m.lattice_y_size, m.lattice_z_size, m.particles_per_site
        bpp = int(m.bytes_per_particle)
        lattice = np.ascontiguousarray(lattice, dtype=np.uint8 if bpp == 1 else np.uint32)
        sites = np.ascontiguousarray(sites, dtype=np.uint8)
        if lattice.shape != (X, Y, Z, P) or sites.shape != (X, Y, Z):
            raise InvalidArgException(ERR_ARG, "lattice or sites shape does not agree with the diffusion model")
        _check(self._lib.lmxio_set_diffusion_model_lattice(self._h, X, Y, Z, P, bpp, lattice.ctypes.data_as(c_void_p),
                                                           _p(sites, ctypes.c_uint8)))

    def setEmptySpatialModel(self) -> None:
        """setSpatialModel for a model with no regions or obstacles."""
        _check(self._lib.lmxio_set_empty_spatial_model(self._h))

    # ---------------------------------------------------------------- trajectories
    def replicateExists(self, replicate: int) -> bool:
        e = c_int()
        _check(self._lib.lmxio_replicate_exists(self._h, int(replicate), byref(e)))
        return bool(e.value)

    def appendSpeciesCounts(self, replicate: int, times, counts) -> None:
        t = np.ascontiguousarray(times, dtype=np.float64).reshape(-1)
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        if c.ndim != 2 or c.shape[0] != t.size:
            raise InvalidArgException(ERR_ARG, "counts must be [entries, species] with one time per entry")
        _check(self._lib.lmxio_append_species_counts(self._h, int(replicate), t.size, c.shape[1], _p(t, c_double),
                                                     _p(c, c_uint32)))

    def appendLattice(self, replicate: int, time: float, lattice: np.ndarray) -> None:
        """lattice: [X, Y, Z, P] uint8 (or uint32), row-major."""
        lat = np.ascontiguousarray(lattice)
        if lat.ndim != 4 or lat.dtype not in (np.uint8, np.uint32):
            raise InvalidArgException(ERR_ARG, "lattice must be a [X, Y, Z, P] uint8 or uint32 array")
        X, Y, Z, P = lat.shape
        _check(self._lib.lmxio_append_lattice(self._h, int(replicate), float(time), X, Y, Z, P, lat.itemsize,
                                              lat.ctypes.data_as(c_void_p)))

    def appendSites(self, replicate: int, time: float, sites: np.ndarray) -> None:
        s = np.ascontiguousarray(sites, dtype=np.uint8)
        if s.ndim != 3:
            raise InvalidArgException(ERR_ARG, "sites must be a [X, Y, Z] array")
        X, Y, Z = s.shape
        _check(self._lib.lmxio_append_sites(self._h, int(replicate), float(time), X, Y, Z, _p(s, ctypes.c_uint8)))

    def _counts(self, replicate: int) -> Tuple[np.ndarray, np.ndarray]:
        n, ns = c_uint32(), c_uint32()
        _check(self._lib.lmxio_get_species_counts_size(self._h, int(replicate), byref(n), byref(ns)))
        t = np.zeros(n.value)
        c = np.zeros((n.value, ns.value), np.uint32)
        _check(self._lib.lmxio_get_species_counts(self._h, int(replicate), _p(t, c_double), _p(c, c_uint32)))
        return t, c

    def getSpeciesCountTimes(self, replicate: int) -> np.ndarray:
        return self._counts(replicate)[0]

    def getSpeciesCounts(self, replicate: int) -> np.ndarray:
        """[entries, species] counts (the reference returns a map time -> 1-based vector)."""
        return self._counts(replicate)[1]

    def getLatticeTimes(self, replicate: int) -> np.ndarray:
        n = c_uint32()
        _check(self._lib.lmxio_get_lattice_count(self._h, int(replicate), byref(n)))
        out = np.zeros(n.value)
        for k in range(n.value):
            t = c_double()
            _check(self._lib.lmxio_get_lattice(self._h, int(replicate), k, byref(t), None, 0, None))
            out[k] = t.value
        return out

    def getLattice(self, replicate: int, latticeIndex: int,
                   shape: Optional[Tuple[int, int, int, int]] = None) -> np.ndarray:
        """Frame latticeIndex as a [X, Y, Z, P] array.  shape (default: the
        diffusion model's lattice) is checked against the stored frame, as the
        reference checks its Lattice argument (SimulationFile.cpp:2010-2056)."""
        if shape is None:
            info = self._diffusion_info()
            shape = (info.X, info.Y, info.Z, info.P)
        bpp = c_uint32()
        _check(self._lib.lmxio_get_lattice(self._h, int(replicate), int(latticeIndex), None, None, 0, byref(bpp)))
        lat = np.zeros(tuple(int(v) for v in shape), np.uint8 if bpp.value == 1 else np.uint32)
        _check(self._lib.lmxio_get_lattice(self._h, int(replicate), int(latticeIndex), None,
                                           lat.ctypes.data_as(c_void_p), lat.nbytes, byref(bpp)))
        return lat

    def writeReplicateDataset(self, replicate: int, name: str, data) -> None:
        d = np.ascontiguousarray(data, dtype=np.uint32).reshape(-1)
        _check(self._lib.lmxio_new_replicate_dataset_u32(self._h, int(replicate), name.encode(), d.size,
                                                         _p(d, c_uint32)))

    def getReplicateDataset(self, replicate: int, name: str) -> np.ndarray:
        n = c_uint32()
        _check(self._lib.lmxio_get_replicate_dataset_u32(self._h, int(replicate), name.encode(), None, 0, byref(n)))
        out = np.zeros(n.value, np.uint32)
        _check(self._lib.lmxio_get_replicate_dataset_u32(self._h, int(replicate), name.encode(), _p(out, c_uint32),
                                                         n.value, byref(n)))
        return out
