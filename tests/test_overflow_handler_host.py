"""Host logic of the rdme.mpd.overflowhandler parameter
(MpdRdmeSolver.cu:126-146), with a stand-in for the C library so that it
runs without a GPU: "relaxed" selects the relaxed handler, "" and "classic"
the classic one, an unknown name warns and keeps classic, and
IntMpdRdmeSolver (whose reference has no relaxed handler) always passes
classic."""
import warnings

import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd import _lib


class StandInLib:
    def __init__(self):
        self.modes = []

    def lmx_initialize(self, h, replicate, seed_top, tau, overflow_mode, boundary):
        self.modes.append((seed_top, overflow_mode, boundary))
        return _lib.OK


def initialized(cls, handler, boundary="periodic"):
    s = cls.__new__(cls)
    s._lib = StandInLib()
    s._h = None
    prm = {"seed": "7", "timestep": "1e-6"}
    if handler is not None:
        prm["rdme.mpd.overflowhandler"] = handler
    s.initialize(0, prm, boundary=boundary)
    return s


@pytest.mark.parametrize("handler,mode", [(None, _lib.OVERFLOW_CLASSIC), ("", _lib.OVERFLOW_CLASSIC),
                                          ("classic", _lib.OVERFLOW_CLASSIC), ("relaxed", _lib.OVERFLOW_RELAXED)])
def test_handler_names(handler, mode):
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        s = initialized(lmx.MpdRdmeSolver, handler)
    assert s._lib.modes == [(7, mode, _lib.BOUNDARY_PERIODIC)]
    assert s.overflowHandler == ("relaxed" if mode == _lib.OVERFLOW_RELAXED else "classic")


def test_unknown_handler_warns_and_keeps_classic():
    with pytest.warns(UserWarning, match="Unknown overflow handler requested: 'other'"):
        s = initialized(lmx.MpdRdmeSolver, "other", boundary="reflecting")
    assert s._lib.modes == [(7, _lib.OVERFLOW_CLASSIC, _lib.BOUNDARY_REFLECTING)]


def test_int_solver_ignores_the_handler():
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        s = initialized(lmx.IntMpdRdmeSolver, "relaxed")
        t = initialized(lmx.IntMpdRdmeSolver, "other")
    assert s._lib.modes[0][1] == _lib.OVERFLOW_CLASSIC and t._lib.modes[0][1] == _lib.OVERFLOW_CLASSIC
    assert s.overflowHandler == "classic"
