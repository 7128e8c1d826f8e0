"""lattice_microbes_amd: MI355X-native MPD-RDME timestep.

The hot path of Lattice Microbes' lm::rdme::MpdRdmeSolver (diffusion sweeps,
reaction kernel, counter-based hash RNG, overflow resolution) as hand-written
HIP kernels for gfx950 behind the C ABI in include/lmx_mpd.h.  This package is
the Python side of that boundary: a ctypes binding, a mirror of the reference
solver class, and generators for the benchmark models.
"""
from ._lib import (LmxError, InvalidArgException, OverflowException, device_count, library_path,
                   load as load_library, torch_interop)
from .solver import (INT_SOLVER_NAME, IntMpdRdmeSolver, MESolverFactory, MpdRdmeSolver, SOLVER_NAME, clear_abort,
                     request_abort, timestep_seed)
from .ensemble import Ensemble
from . import models

__all__ = ["LmxError", "InvalidArgException", "OverflowException", "device_count", "library_path",
           "load_library", "torch_interop", "MESolverFactory", "MpdRdmeSolver", "SOLVER_NAME", "IntMpdRdmeSolver", "INT_SOLVER_NAME", "timestep_seed", "request_abort", "clear_abort", "Ensemble", "models"]
