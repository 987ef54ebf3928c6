"""ctypes binding of the SPH entry points of libmpm355.so (include/sph.h).

The library must have been built (``__graft_entry__.build()`` or
``make -C webgpu-ocean_amd``); a missing library raises instead of falling
back to anything else.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libmpm355.so")

STATUS_NAMES = {
    0: "SPH_OK",
    1: "SPH_ERR_INVALID_ARGUMENT",
    2: "SPH_ERR_GRID_TOO_LARGE",
    3: "SPH_ERR_OUT_OF_MEMORY",
    4: "SPH_ERR_DEVICE",
    5: "SPH_ERR_NO_DEVICE",
    6: "SPH_ERR_NOT_READY",
}
STAGES = ("count", "scan_reduce", "scan", "scatter", "sort_gather", "density", "force")

# Every symbol include/sph.h declares (checked by tests/test_capi_symbols.py).
EXPORTED = (
    "sph_last_error", "sph_default_params", "sph_create", "sph_destroy", "sph_reset",
    "sph_set_box", "sph_step", "sph_run_stage", "sph_synchronize", "sph_read_posvel",
    "sph_posvel_device_ptr", "sph_read_particles", "sph_write_particles",
    "sph_read_cell_start", "sph_read_sorted_order", "sph_get_counts", "sph_time_steps",
    "sph_time_stages",
)


class SPHError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Params(ctypes.Structure):
    """sph_params: the host constants of sph/sph.ts:41,72-77."""
    _fields_ = [
        ("kernel_radius", ctypes.c_double),
        ("mass", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("stiffness", ctypes.c_double),
        ("near_stiffness", ctypes.c_double),
        ("rest_density", ctypes.c_double),
        ("viscosity", ctypes.c_double),
    ]


_lib = None


def load(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build() "
                          "(make -C webgpu-ocean_amd)")
    L = ctypes.CDLL(path)
    P, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    F3 = ctypes.POINTER(ctypes.c_float)
    FP = ctypes.POINTER(ctypes.c_float)
    sigs = {
        "sph_last_error": ([], ctypes.c_char_p),
        "sph_default_params": ([ctypes.POINTER(Params)], i32),
        "sph_create": ([ctypes.POINTER(Params), P, u32, F3, ctypes.POINTER(P)], i32),
        "sph_destroy": ([P], i32),
        "sph_reset": ([P, P, u32, F3], i32),
        "sph_set_box": ([P, F3], i32),
        "sph_step": ([P, u32], i32),
        "sph_run_stage": ([P, i32], i32),
        "sph_synchronize": ([P], i32),
        "sph_read_posvel": ([P, P, u32], i32),
        "sph_posvel_device_ptr": ([P, ctypes.POINTER(P)], i32),
        "sph_read_particles": ([P, P, u32], i32),
        "sph_write_particles": ([P, P, u32], i32),
        "sph_read_cell_start": ([P, P, u64], i32),
        "sph_read_sorted_order": ([P, P, u32], i32),
        "sph_get_counts": ([P, ctypes.POINTER(u32), ctypes.POINTER(u64)], i32),
        "sph_time_steps": ([P, u32, FP], i32),
        "sph_time_stages": ([P, u32, FP], i32),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(status: int):
    if status != 0:
        msg = load().sph_last_error()
        raise SPHError(status, msg.decode() if msg else "")


def f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])
