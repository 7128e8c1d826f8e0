"""Ensembles of MPD-RDME solvers stepped together on one GPU.

No reference counterpart: the reference runs each replicate in its own
thread or process (src/core/ReplicateRunner.cpp).  Here many independently
configured solvers (any lattice size, model, seed) advance with one kernel
launch per kernel class per step, replayed from a captured HIP graph; each
member's lattice stays bit-identical to running that solver alone.
"""
from __future__ import annotations

from ctypes import byref, c_double, c_int, c_uint32, c_void_p
from typing import List

from . import _lib
from .solver import MpdRdmeSolver


class Ensemble:
    def __init__(self, device: int = 0, graph: bool = True):
        self._lib = _lib.load()
        h = c_void_p()
        _lib.check(self._lib.lmx_batch_create(int(device), byref(h)))
        self._h = h
        self.members: List[MpdRdmeSolver] = []
        if not graph:
            self.set_graph(False)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lmx_batch_destroy(self._h)
            self._h = c_void_p()
        self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, on: bool) -> None:
        _lib.check(self._lib.lmx_batch_set_option(self._h, _lib.BATCH_GRAPH, 1 if on else 0))

    def set_reaction_list(self, on: bool) -> None:
        """Reaction phases over the occupied sites each member's z sweep lists
        (default), or the full in-place pass; the results are identical."""
        _lib.check(self._lib.lmx_batch_set_option(self._h, _lib.BATCH_REACTION_LIST, 1 if on else 0))

    def set_tile_threads(self, threads: int) -> None:
        """Threads per block of the tiled y/z sweeps (0: the default); the
        results are identical."""
        _lib.check(self._lib.lmx_batch_set_option(self._h, _lib.BATCH_TILE_THREADS, int(threads)))

    def autotune(self, steps: int = 8) -> dict:
        """Times the batched launch configurations (reaction list on/off, tile
        blocks of the default and half the default threads) on the members'
        own state, restores it, and keeps the fastest."""
        ms = c_double()
        chosen = (c_int * 2)()
        _lib.check(self._lib.lmx_batch_autotune(self._h, int(steps), byref(ms), chosen))
        return {"reaction_list": bool(chosen[0]), "tile_threads": int(chosen[1]), "ms_per_step": ms.value}

    def add(self, solver: MpdRdmeSolver) -> None:
        _lib.check(self._lib.lmx_batch_add(self._h, solver._h))
        self.members.append(solver)

    def remove(self, solver: MpdRdmeSolver) -> None:
        _lib.check(self._lib.lmx_batch_remove(self._h, solver._h))
        self.members.remove(solver)

    def __len__(self) -> int:
        n = c_uint32()
        _lib.check(self._lib.lmx_batch_size(self._h, byref(n)))
        return n.value

    def run(self, nsteps: int, synchronize: bool = True) -> None:
        """Advance every member by nsteps."""
        _lib.check(self._lib.lmx_batch_run_steps(self._h, int(nsteps)))
        if synchronize:
            self.synchronize()

    def synchronize(self) -> None:
        _lib.check(self._lib.lmx_batch_synchronize(self._h))

    def time_steps(self, nsteps: int) -> float:
        """Run nsteps for every member; returns the elapsed milliseconds."""
        ms = c_double()
        _lib.check(self._lib.lmx_batch_time_steps(self._h, int(nsteps), byref(ms)))
        return ms.value
