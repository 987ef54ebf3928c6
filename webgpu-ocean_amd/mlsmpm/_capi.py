"""ctypes binding of libmpm355.so (the C ABI declared in include/mpm.h).

The library must have been built (``__graft_entry__.build()`` or
``make -C webgpu-ocean_amd``); a missing library raises instead of falling
back to anything else.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libmpm355.so")

MPM_OK = 0
STATUS_NAMES = {
    0: "MPM_OK",
    1: "MPM_ERR_INVALID_ARGUMENT",
    2: "MPM_ERR_GRID_TOO_LARGE",
    3: "MPM_ERR_OUT_OF_MEMORY",
    4: "MPM_ERR_DEVICE",
    5: "MPM_ERR_NO_DEVICE",
    6: "MPM_ERR_NOT_READY",
}
STAGES = ("clear_grid", "p2g_1", "p2g_2", "update_grid", "g2p", "copy_position")

# Every symbol include/mpm.h declares (checked by tests/test_capi_symbols.py).
EXPORTED = (
    "mpm_abi_version", "mpm_last_error", "mpm_default_params", "mpm_create",
    "mpm_destroy", "mpm_reset", "mpm_set_box", "mpm_step", "mpm_run_stage",
    "mpm_synchronize", "mpm_read_posvel", "mpm_posvel_device_ptr",
    "mpm_read_particles", "mpm_write_particles", "mpm_read_grid", "mpm_write_grid",
    "mpm_read_grid_velocity",
    "mpm_get_counts", "mpm_time_steps", "mpm_time_stages", "mpm_get_tuning",
    "mpm_set_tuning", "mpm_autotune", "mpm_backward", "mpm_step_noalloc",
    "mpm_batch_create", "mpm_batch_destroy", "mpm_batch_step", "mpm_batch_set_box",
    "mpm_batch_synchronize", "mpm_batch_get_counts", "mpm_batch_read_posvel",
    "mpm_batch_read_particles", "mpm_batch_get_tuning", "mpm_batch_set_tuning",
    "mpm_batch_autotune", "mpm_batch_time_steps", "mpm_batch_time_stages",
)


class MPMError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Tuning(ctypes.Structure):
    _fields_ = [
        ("path", ctypes.c_int),
        ("particles_per_thread", ctypes.c_int),
        ("grid_blocks", ctypes.c_int),
        ("xcd_remap", ctypes.c_int),
        ("tile_edge", ctypes.c_int),
    ]


class Params(ctypes.Structure):
    _fields_ = [
        ("stiffness", ctypes.c_float),
        ("rest_density", ctypes.c_float),
        ("dynamic_viscosity", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("fixed_point_multiplier", ctypes.c_float),
    ]


class AdjointConfig(ctypes.Structure):
    _fields_ = [
        ("substeps", ctypes.c_uint32),
        ("checkpoint_every", ctypes.c_uint32),
        ("target", ctypes.c_float * 3),
        ("ke_weight", ctypes.c_float),
    ]


class Scene(ctypes.Structure):
    _fields_ = [
        ("particles", ctypes.c_void_p),
        ("n", ctypes.c_uint32),
        ("init_box", ctypes.c_float * 3),
        ("params", Params),
    ]


_lib = None


def load(path: str = LIB_PATH):
    """Loads libmpm355.so and declares the argument types of every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build() "
                          "(make -C webgpu-ocean_amd)")
    L = ctypes.CDLL(path)
    P, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    F3 = ctypes.POINTER(ctypes.c_float)
    sigs = {
        "mpm_abi_version": ([], i32),
        "mpm_last_error": ([], ctypes.c_char_p),
        "mpm_default_params": ([ctypes.POINTER(Params)], i32),
        "mpm_create": ([ctypes.POINTER(Params), P, u32, F3, ctypes.POINTER(P)], i32),
        "mpm_destroy": ([P], i32),
        "mpm_reset": ([P, P, u32, F3], i32),
        "mpm_set_box": ([P, F3], i32),
        "mpm_step": ([P, u32], i32),
        "mpm_run_stage": ([P, i32], i32),
        "mpm_synchronize": ([P], i32),
        "mpm_read_posvel": ([P, P, u32], i32),
        "mpm_posvel_device_ptr": ([P, ctypes.POINTER(P)], i32),
        "mpm_read_particles": ([P, P, u32], i32),
        "mpm_write_particles": ([P, P, u32], i32),
        "mpm_read_grid": ([P, P, u64], i32),
        "mpm_write_grid": ([P, P, u64], i32),
        "mpm_read_grid_velocity": ([P, P, u64], i32),
        "mpm_get_counts": ([P, ctypes.POINTER(u32), ctypes.POINTER(u64)], i32),
        "mpm_time_steps": ([P, u32, u32, ctypes.POINTER(ctypes.c_float)], i32),
        "mpm_time_stages": ([P, u32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(i32),
                             ctypes.POINTER(ctypes.c_char_p), i32, ctypes.POINTER(i32)], i32),
        "mpm_get_tuning": ([P, ctypes.POINTER(Tuning)], i32),
        "mpm_set_tuning": ([P, ctypes.POINTER(Tuning)], i32),
        "mpm_autotune": ([P, u32, ctypes.POINTER(Tuning)], i32),
        "mpm_backward": ([P, ctypes.POINTER(AdjointConfig), P, P, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)], i32),
        "mpm_step_noalloc": ([ctypes.POINTER(Params), P, P, u32, F3, F3, u32, P], i32),
        "mpm_batch_create": ([ctypes.POINTER(Scene), u32, ctypes.POINTER(P)], i32),
        "mpm_batch_destroy": ([P], i32),
        "mpm_batch_step": ([P, u32], i32),
        "mpm_batch_set_box": ([P, u32, F3], i32),
        "mpm_batch_synchronize": ([P], i32),
        "mpm_batch_get_counts": ([P, u32, ctypes.POINTER(u32), ctypes.POINTER(u64)], i32),
        "mpm_batch_read_posvel": ([P, u32, P, u32], i32),
        "mpm_batch_read_particles": ([P, u32, P, u32], i32),
        "mpm_batch_get_tuning": ([P, ctypes.POINTER(Tuning)], i32),
        "mpm_batch_set_tuning": ([P, ctypes.POINTER(Tuning)], i32),
        "mpm_batch_autotune": ([P, u32, ctypes.POINTER(Tuning)], i32),
        "mpm_batch_time_steps": ([P, u32, ctypes.POINTER(ctypes.c_float)], i32),
        "mpm_batch_time_stages": ([P, u32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(i32),
                                   ctypes.POINTER(ctypes.c_char_p), i32, ctypes.POINTER(i32)], i32),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(status: int):
    if status != MPM_OK:
        msg = load().mpm_last_error()
        raise MPMError(status, msg.decode() if msg else "")
    return status


def f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])
