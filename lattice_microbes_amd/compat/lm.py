"""A minimal ``lm`` module: the names pyLM and jLM scripts use to build a
model file and run the MPD-RDME solver, backed by this library.

Put this directory on ``sys.path`` (or call
``lattice_microbes_amd.compat.install()``) and ``import lm`` resolves here, so
that scripts written against the reference's Python module (src/swig/lm.i)
select the MI355X solver unchanged:

* ``ReactionModel``, ``DiffusionModel``, ``SpatialModel``: the protobuf
  messages with their generated accessors (src/protobuf/*.proto);
* ``SimulationFile``: create, parameters, models, initial lattice, output;
* ``ByteLattice``, ``IntLattice``, ``getCompiledLatticeMaxOccupancy``;
* ``MpdRdmeSolver`` (an overridable ``hookSimulation(time, lattice)``),
  ``RDMESolver`` (the same class), ``IntMpdRdmeSolver`` (the uint32
  lattice);
* ``runSolver``, ``runSimulation``.

Spatial geometry (regions, obstacles) and the other solvers of the reference
module are not part of this library.
"""
from __future__ import annotations

import os
from typing import Dict, List

import numpy as np

from lattice_microbes_amd import runner as _runner
from lattice_microbes_amd import simfile as _sf
from lattice_microbes_amd import solver as _solver
from lattice_microbes_amd.lattice import ByteLattice, IntLattice, Lattice
from lattice_microbes_amd.solver import HookLattice as CudaByteLattice  # what a hook receives
from lattice_microbes_amd.solver import IntHookLattice as CudaIntLattice

BUILD_CONFIGURATION = "lattice_microbes_amd (gfx950)"


def getCompiledLatticeMaxOccupancy() -> int:
    """Particles per site of new lattices (the reference's build-time
    MPD_LATTICE_MAX_OCCUPANCY, 8 by default; the solver here takes 8 or 16 per
    model, and LMX_LATTICE_MAX_OCCUPANCY selects the value returned)."""
    return int(os.environ.get("LMX_LATTICE_MAX_OCCUPANCY", "8"))


# ------------------------------------------------------------------ messages
class _Message:
    """protobuf-style accessors for the fields listed in _scalars (name ->
    default), _repeated (scalar lists) and _messages (name -> class)."""

    _scalars: Dict[str, object] = {}
    _repeated: tuple = ()
    _messages: Dict[str, type] = {}

    def __init__(self):
        self.Clear()

    def Clear(self) -> None:
        self._v = dict(self._scalars)
        self._has = set()
        self._r = {k: [] for k in self._repeated}
        self._m = {k: [] for k in self._messages}

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        for prefix in ("set_", "add_", "has_", "clear_", "mutable_"):
            if name.startswith(prefix):
                return self._method(prefix, name[len(prefix):])
        if name.endswith("_size"):
            f = name[:-5]
            if f in self._repeated:
                return lambda: len(self._r[f])
            if f in self._messages:
                return lambda: len(self._m[f])
        if name in self._scalars:
            return lambda: self._v[name]
        if name in self._repeated:
            return lambda i=None: list(self._r[name]) if i is None else self._r[name][i]
        if name in self._messages:
            return lambda i: self._m[name][i]
        raise AttributeError(name)

    def _method(self, prefix, f):
        if prefix == "set_" and f in self._scalars:
            def set_(v):
                self._v[f] = v
                self._has.add(f)
            return set_
        if prefix == "set_" and f in self._repeated:
            def set_i(i, v):
                self._r[f][i] = v
            return set_i
        if prefix == "add_" and f in self._repeated:
            return lambda v: self._r[f].append(v)
        if prefix == "add_" and f in self._messages:
            def add():
                m = self._messages[f]()
                self._m[f].append(m)
                return m
            return add
        if prefix == "mutable_" and f in self._messages:
            return lambda i: self._m[f][i]
        if prefix == "has_" and f in self._scalars:
            return lambda: f in self._has
        if prefix == "clear_":
            def clear():
                if f in self._scalars:
                    self._v[f] = self._scalars[f]
                    self._has.discard(f)
                elif f in self._repeated:
                    self._r[f] = []
                elif f in self._messages:
                    self._m[f] = []
            return clear
        raise AttributeError(prefix + f)


class _Reaction(_Message):
    _scalars = {"type": 0, "rate_has_noise": False, "rate_noise_variance": 0.0, "rate_noise_tau": 0.0}
    _repeated = ("rate_constant",)


class ReactionModel(_Message):
    """lm.io.ReactionModel (src/protobuf/ReactionModel.proto)."""
    _scalars = {"number_species": 0, "number_reactions": 0}
    _repeated = ("initial_species_count", "stoichiometric_matrix", "dependency_matrix")
    _messages = {"reaction": _Reaction}

    def _to_data(self) -> _sf.ReactionModel:
        nS, R = int(self.number_species()), int(self.number_reactions())
        noise = None
        if any(self.reaction(r).rate_has_noise() for r in range(self.reaction_size())):
            noise = np.array([[self.reaction(r).rate_noise_variance(), self.reaction(r).rate_noise_tau()]
                              if self.reaction(r).rate_has_noise() else [0.0, 0.0] for r in range(R)])
        return _sf.ReactionModel(
            number_species=nS, number_reactions=R,
            initial_species_count=np.array(self.initial_species_count(), np.uint32),
            types=np.array([self.reaction(r).type() for r in range(self.reaction_size())], np.uint32),
            rate_constants=[list(self.reaction(r).rate_constant()) for r in range(self.reaction_size())],
            rate_noise=noise,
            S=np.array(self.stoichiometric_matrix(), np.int32).reshape(nS, -1) if R else np.zeros((nS, 0), np.int32),
            D=np.array(self.dependency_matrix(), np.uint32).reshape(nS, -1) if R else np.zeros((nS, 0), np.uint32))

    def _from_data(self, d: _sf.ReactionModel) -> None:
        self.Clear()
        self.set_number_species(int(d.number_species))
        self.set_number_reactions(int(d.number_reactions))
        for c in d.initial_species_count:
            self.add_initial_species_count(int(c))
        for r in range(int(d.number_reactions)):
            rx = self.add_reaction()
            rx.set_type(int(d.types[r]))
            for k in d.rate_constants[r]:
                rx.add_rate_constant(float(k))
            if d.rate_has_noise(r):
                rx.set_rate_has_noise(True)
                rx.set_rate_noise_variance(float(d.rate_noise[r][0]))
                rx.set_rate_noise_tau(float(d.rate_noise[r][1]))
        for v in np.asarray(d.S).reshape(-1):
            self.add_stoichiometric_matrix(int(v))
        for v in np.asarray(d.D).reshape(-1):
            self.add_dependency_matrix(int(v))


class DiffusionModel(_Message):
    """lm.io.DiffusionModel (src/protobuf/DiffusionModel.proto)."""
    _scalars = {"number_species": 0, "number_reactions": 0, "number_site_types": 0, "lattice_spacing": 0.0,
                "lattice_x_size": 0, "lattice_y_size": 0, "lattice_z_size": 0, "particles_per_site": 0,
                "bytes_per_particle": 1}
    _repeated = ("diffusion_matrix", "reaction_location_matrix")

    def _to_data(self) -> _sf.DiffusionModel:
        nS, nST, R = int(self.number_species()), int(self.number_site_types()), int(self.number_reactions())
        return _sf.DiffusionModel(
            number_species=nS, number_reactions=R, number_site_types=nST,
            lattice_spacing=float(self.lattice_spacing()), lattice_x_size=int(self.lattice_x_size()),
            lattice_y_size=int(self.lattice_y_size()), lattice_z_size=int(self.lattice_z_size()),
            particles_per_site=int(self.particles_per_site()), bytes_per_particle=int(self.bytes_per_particle()),
            diffusion_matrix=np.array(self.diffusion_matrix(), np.float64).reshape(nST, nST, nS),
            reaction_location_matrix=np.array(self.reaction_location_matrix(), np.uint32).reshape(R, nST))

    def _from_data(self, d: _sf.DiffusionModel) -> None:
        self.Clear()
        for f in ("number_species", "number_reactions", "number_site_types", "lattice_spacing", "lattice_x_size",
                  "lattice_y_size", "lattice_z_size", "particles_per_site", "bytes_per_particle"):
            getattr(self, "set_" + f)(getattr(d, f))
        if d.diffusion_matrix is not None:
            for v in np.asarray(d.diffusion_matrix).reshape(-1):
                self.add_diffusion_matrix(float(v))
        if d.reaction_location_matrix is not None:
            for v in np.asarray(d.reaction_location_matrix).reshape(-1):
                self.add_reaction_location_matrix(int(v))


class _Shape(_Message):
    _scalars = {"shape": 0, "site_type": 0}
    _repeated = ("shape_parameter",)


class SpatialModel(_Message):
    """lm.io.SpatialModel: only the empty model (no regions or obstacles) can
    be written here."""
    _messages = {"region": _Shape, "obstacle": _Shape}


# ------------------------------------------------------------------ file
class SimulationFile(_sf.SimulationFile):
    """lm.SimulationFile with the reference module's argument forms: models
    are message objects, getters fill the message passed in."""

    def setReactionModel(self, rm) -> None:
        super().setReactionModel(rm._to_data() if isinstance(rm, ReactionModel) else rm)

    def getReactionModel(self, rm=None):
        d = super().getReactionModel()
        if rm is None:
            return d
        rm._from_data(d)
        return rm

    def setDiffusionModel(self, dm, lattice=None, sites=None) -> None:
        super().setDiffusionModel(dm._to_data() if isinstance(dm, DiffusionModel) else dm, lattice, sites)

    def getDiffusionModel(self, dm=None):
        d = super().getDiffusionModel()
        if dm is None:
            return d
        dm._from_data(d)
        return dm

    def setDiffusionModelLattice(self, dm, lattice, sites=None) -> None:
        """With an ``lm.ByteLattice`` or ``lm.IntLattice``, as jLM and pyLM
        call it (SimulationFile.cpp:800-900); or with row-major lattice and site
        arrays."""
        data = dm._to_data() if isinstance(dm, DiffusionModel) else dm
        if isinstance(lattice, Lattice):
            lattice, sites = lattice.row_major(), lattice.sites_row_major()
        super().setDiffusionModelLattice(data, lattice, sites)

    def getDiffusionModelLattice(self, dm=None, lattice=None):
        lat, sites = super().getDiffusionModelLattice()
        if isinstance(lattice, Lattice):
            lattice.setFromRowMajorData(lat, sites)
            return lattice
        return lat, sites

    def setSpatialModel(self, sm) -> None:
        if sm.region_size() or sm.obstacle_size():
            raise NotImplementedError("spatial geometry (regions, obstacles) is not part of this library")
        self.setEmptySpatialModel()


def simulationFile(filename: str) -> SimulationFile:
    return SimulationFile(filename)


# ------------------------------------------------------------------ solver and runners
class MpdRdmeSolver(_solver.MpdRdmeSolver):
    """lm.MpdRdmeSolver: constructible without arguments (device 0), for
    subclassing with ``hookSimulation(self, time, lattice)`` as jLM's
    ``makeSolver`` composes it."""

    def __init__(self, device: int = 0):
        super().__init__(device)


RDMESolver = MpdRdmeSolver


class IntMpdRdmeSolver(_solver.IntMpdRdmeSolver):
    """lm.IntMpdRdmeSolver: the uint32 lattice (4 bytes per particle)."""

    def __init__(self, device: int = 0):
        super().__init__(device)


def runSolver(simulationFilename: str, replicate: int, solver, cudaDevices: List[int] = (0,),
              checkpointInterval: int = 0) -> None:
    """lm.runSolver (src/swig/lm.i:1042)."""
    _runner.runSolver(simulationFilename, replicate, solver, list(cudaDevices) or [0], int(checkpointInterval))


def runSimulation(simulationFilename: str, replicate: int, solverClass: str = _solver.SOLVER_NAME,
                  cudaDevices: List[int] = (0,), checkpointInterval: int = 0) -> None:
    """lm.runSimulation (src/swig/lm.i:1039)."""
    _runner.runSimulation(simulationFilename, replicate, solverClass, list(cudaDevices) or [0],
                          int(checkpointInterval))
