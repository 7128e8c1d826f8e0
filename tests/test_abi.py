"""The C-ABI libraries load and export every entry point include/lmx_mpd.h and
include/lmx_io.h declare (no compute calls: this runs without a GPU)."""
import os
import re

import pytest

import lattice_microbes_amd as lmx
from lattice_microbes_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lmx_mpd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lmx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "lmx_create" in names and "lmx_run_steps" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = lmx.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_abi_version_and_device_count():
    lib = lmx.load_library()
    assert lib.lmx_abi_version() == 1
    assert lmx.device_count() >= 0


def test_create_without_device_fails_loudly():
    if lmx.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(lmx.LmxError) as ei:
        lmx.MpdRdmeSolver(0)
    assert "device" in str(ei.value)


def declared_io_functions():
    text = open(os.path.join(ROOT, "include", "lmx_io.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lmxio_[a-z0-9_]+)\s*\(", text)))


def test_io_library_exports_every_declared_symbol():
    from lattice_microbes_amd import simfile

    names = declared_io_functions()
    assert "lmxio_open" in names and "lmxio_append_lattice" in names
    lib = simfile.load()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(simfile.SIGNATURES)
