"""Running a replicate from a Lattice Microbes HDF5 file.

Mirror of ``runSolver`` / ``runSimulation`` (src/core/runSimulation.cpp:115-255,
exposed to Python as ``lm.runSolver`` / ``lm.runSimulation``, src/swig/lm.i:1039-1043):
open the file, read its parameters, reaction model, diffusion model and
initial lattice, initialize the solver for the replicate, hand it the models
and generate the trajectory into the same file.  The reference writes through
a background output thread; here the solver writes to the open file directly
(same datasets, same contents).
"""
from __future__ import annotations

import sys
from typing import Sequence

from . import _lib
from .simfile import SimulationFile
from .solver import MESolverFactory, MpdRdmeSolver, SOLVER_NAME


def runSolver(simulationFilename: str, replicate: int, solver: MpdRdmeSolver,
              cudaDevices: Sequence[int] = (0,), checkpointInterval: int = 0,
              boundary: str = "periodic") -> bool:
    """runSolver (runSimulation.cpp:130-255) with an already constructed
    solver (for example a subclass overriding hookSimulation).  As in the
    reference, an error while the replicate runs is printed, not raised;
    returns whether the replicate completed."""
    if len(cudaDevices) > 1:
        raise _lib.LmxError(_lib.ERR_ARG, "Multiple CUDA devices specified for single GPU solver")
    if checkpointInterval:
        raise _lib.InvalidArgException(_lib.ERR_ARG, "checkpointing while a replicate runs is not supported")
    with SimulationFile(simulationFilename) as f:
        parameters = f.getParameters()
        rm = f.getReactionModel() if solver.needsReactionModel() else None
        dm = lattice = sites = None
        if solver.needsDiffusionModel():
            dm = f.getDiffusionModel()
            lattice, sites = f.getDiffusionModelLattice(dm)
        try:
            solver.initialize(int(replicate), parameters, boundary=boundary)
            if rm is not None:
                solver.setReactionModel(rm)
            if dm is not None:
                solver.setDiffusionModel(dm, lattice, sites)
            solver.generateTrajectory(output=f)
        except (_lib.LmxError, ValueError, RuntimeError) as e:
            print(f"Exception during execution of replicate: {e}.", file=sys.stderr, flush=True)
            return False
    return True


def runSimulation(simulationFilename: str, replicate: int, solverClass: str = SOLVER_NAME,
                  cudaDevices: Sequence[int] = (0,), checkpointInterval: int = 0,
                  boundary: str = "periodic") -> bool:
    """runSimulation (runSimulation.cpp:115-128): the solver chosen by class
    name through the factory, then runSolver."""
    factory = MESolverFactory()
    factory.setSolver(solverClass)
    solver = factory.instantiate(int(cudaDevices[0]) if len(cudaDevices) else 0)
    try:
        return runSolver(simulationFilename, replicate, solver, cudaDevices, checkpointInterval, boundary)
    finally:
        solver.close()
