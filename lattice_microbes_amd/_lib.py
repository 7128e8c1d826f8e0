"""ctypes binding of the C ABI declared in include/lmx_mpd.h.

The shared library is built in-tree by ``__graft_entry__.build()``.  There is
no fallback implementation: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liblmx_mpd.so")

# Status codes (enum lmx_status).
OK = 0
ERR_ARG = 1
ERR_HIP = 2
ERR_OVERFLOW_CAP = 3
ERR_OVERFLOW_UNPLACEABLE = 4
ERR_OVERFLOW_OUT_OF_RANGE = 5
ERR_REACTION_UNDERFLOW = 6
ERR_STATE = 7
ERR_NOMEM = 8
ERR_INVALID_LATTICE = 9

BOUNDARY_PERIODIC = 0
BOUNDARY_REFLECTING = 1
OVERFLOW_CLASSIC = 0
OVERFLOW_RELAXED = 1

OPT_MARCH_SEGMENT = 1
OPT_FORCE_GENERIC = 2
OPT_YZ_KERNEL = 3
OPT_FUSED_ZSEG = 4
OPT_FUSE_REACTION = 5
BATCH_GRAPH = 1
BATCH_REACTION_LIST = 2
BATCH_TILE_THREADS = 3

PATH_SWEEPS = 0
PATH_FUSED = 1

# enum lmx_kernel
KERNEL_NAMES = ("sweep_x", "sweep_y", "sweep_z", "reaction", "resolve", "fused")

_u32p = POINTER(c_uint32)
_u64p = POINTER(c_uint64)
_dp = POINTER(c_double)
_fp = POINTER(c_float)
_u8p = POINTER(ctypes.c_uint8)
_i32p = POINTER(ctypes.c_int32)

# name -> (restype, argtypes)
SIGNATURES = {
    "lmx_last_error": (c_char_p, []),
    "lmx_abi_version": (c_int, []),
    "lmx_device_count": (c_int, [POINTER(c_int)]),
    "lmx_create": (c_int, [c_int, POINTER(c_void_p)]),
    "lmx_destroy": (c_int, [c_void_p]),
    "lmx_initialize": (c_int, [c_void_p, c_uint32, c_uint32, c_double, c_int, c_int]),
    "lmx_get_seed": (c_int, [c_void_p, _u32p]),
    "lmx_build_model": (c_int, [c_void_p, c_uint32, c_uint32, _u32p, _u32p, _dp, _i32p, _u32p, c_uint32]),
    "lmx_build_diffusion_model": (c_int, [c_void_p, c_uint32, _dp, _u32p, c_uint32, c_uint32, c_uint32,
                                          c_uint32, c_uint32, c_double, _u8p, _u8p, c_int]),
    "lmx_set_reaction_rate": (c_int, [c_void_p, c_uint32, c_float]),
    "lmx_set_path": (c_int, [c_void_p, c_int]),
    "lmx_set_launch_option": (c_int, [c_void_p, c_int, c_int]),
    "lmx_run_steps": (c_int, [c_void_p, c_uint32]),
    "lmx_run_steps_static": (c_int, [c_void_p, c_uint32]),
    "lmx_synchronize": (c_int, [c_void_p]),
    "lmx_get_timestep": (c_int, [c_void_p, _u32p]),
    "lmx_set_timestep": (c_int, [c_void_p, c_uint32]),
    "lmx_get_lattice": (c_int, [c_void_p, _u8p, c_size_t]),
    "lmx_get_lattice_native": (c_int, [c_void_p, _u32p, c_size_t]),
    "lmx_get_sites": (c_int, [c_void_p, _u8p, c_size_t]),
    "lmx_set_lattice": (c_int, [c_void_p, _u8p, c_size_t]),
    "lmx_set_lattice_native": (c_int, [c_void_p, _u32p, c_size_t]),
    "lmx_set_sites": (c_int, [c_void_p, _u8p, c_size_t]),
    "lmx_device_lattice": (c_int, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p)]),
    "lmx_validate_lattice": (c_int, [c_void_p]),
    "lmx_copy_bandwidth": (c_int, [c_int, c_size_t, c_int, ctypes.POINTER(c_double)]),
    "lmx_get_species_counts": (c_int, [c_void_p, _u32p, c_uint32]),
    "lmx_get_overflow_stats": (c_int, [c_void_p, _u64p, _u64p]),
    "lmx_get_unresolved_overflows": (c_int, [c_void_p, _u32p]),
    "lmx_autotune": (c_int, [c_void_p, c_uint32, _dp, POINTER(c_int)]),
    "lmx_time_steps": (c_int, [c_void_p, c_uint32, _dp, _dp, _u64p, c_uint32]),
    "lmx_batch_create": (c_int, [c_int, POINTER(c_void_p)]),
    "lmx_batch_destroy": (c_int, [c_void_p]),
    "lmx_batch_add": (c_int, [c_void_p, c_void_p]),
    "lmx_batch_remove": (c_int, [c_void_p, c_void_p]),
    "lmx_batch_size": (c_int, [c_void_p, POINTER(c_uint32)]),
    "lmx_batch_set_option": (c_int, [c_void_p, c_int, c_int]),
    "lmx_batch_run_steps": (c_int, [c_void_p, c_uint32]),
    "lmx_batch_synchronize": (c_int, [c_void_p]),
    "lmx_batch_time_steps": (c_int, [c_void_p, c_uint32, _dp]),
    "lmx_batch_autotune": (c_int, [c_void_p, c_uint32, _dp, POINTER(c_int)]),
    "lmx_get_table_f32": (c_int, [c_void_p, c_int, _fp, c_size_t, POINTER(c_size_t)]),
    "lmx_device_exp": (c_int, [c_int, _dp, _dp, c_size_t]),
    "lmx_device_reaction_probability": (c_int, [c_int, _fp, _fp, c_size_t]),
}

_lib = None


def library_path() -> str:
    return LIB_PATH


def load() -> ctypes.CDLL:
    """Load the HIP extension; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"lmx_mpd HIP library not found at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class LmxError(RuntimeError):
    """Raised for a nonzero lmx_status; ``code`` holds the status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[lmx status {code}] {message}")
        self.code = code


class InvalidArgException(LmxError):
    pass


class OverflowException(LmxError):
    pass


def check(code: int) -> None:
    if code == OK:
        return
    msg = load().lmx_last_error().decode(errors="replace")
    if code in (ERR_ARG, ERR_INVALID_LATTICE):
        raise InvalidArgException(code, msg)
    if code in (ERR_OVERFLOW_CAP, ERR_OVERFLOW_UNPLACEABLE, ERR_OVERFLOW_OUT_OF_RANGE,
                ERR_REACTION_UNDERFLOW):
        raise OverflowException(code, msg)
    raise LmxError(code, msg)


def torch_interop() -> None:
    """Starts PyTorch's HIP runtime before this library's.  PyTorch ships its
    own HIP runtime, which starts only if it is initialised first in the
    process; call this before the first solver or device query when torch
    tensors are to share the GPU with the solver (device-side hooks)."""
    if _lib is not None:
        import torch

        if torch.cuda.is_initialized():
            return
        raise LmxError(ERR_STATE, "torch_interop() must be called before the first solver or device query")
    import torch

    torch.cuda.init()


def device_count() -> int:
    n = c_int(0)
    check(load().lmx_device_count(ctypes.byref(n)))
    return n.value
