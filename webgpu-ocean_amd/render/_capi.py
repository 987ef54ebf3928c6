"""ctypes binding of the renderer entry points of libmpm355.so (include/render.h).

The library must have been built (``__graft_entry__.build()``); a missing
library raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libmpm355.so")

RENDER_OK = 0
STATUS_NAMES = {0: "RENDER_OK", 1: "RENDER_ERR_INVALID_ARGUMENT", 3: "RENDER_ERR_OUT_OF_MEMORY",
                4: "RENDER_ERR_DEVICE", 5: "RENDER_ERR_NO_DEVICE"}
IMAGES = {"depth": 0, "filtered_depth": 1, "thickness": 2, "filtered_thickness": 3, "color": 4}

# Every symbol include/render.h declares (checked by tests/test_capi_symbols.py).
EXPORTED = ("render_last_error", "render_create", "render_destroy", "render_execute",
            "render_execute_host", "render_read_image", "render_color_device_ptr",
            "render_time_frames")


class RenderError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Params(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("radius", ctypes.c_double), ("fov", ctypes.c_double)]


_lib = None


def load(path: str | None = None):
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    P, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    L.render_last_error.restype = ctypes.c_char_p
    L.render_last_error.argtypes = []
    L.render_create.argtypes = [ctypes.POINTER(Params), P, u32, ctypes.POINTER(P)]
    L.render_destroy.argtypes = [P]
    L.render_execute.argtypes = [P, P, u32, P, i32]
    L.render_execute_host.argtypes = [P, P, u32, P, i32]
    L.render_read_image.argtypes = [P, i32, P]
    L.render_color_device_ptr.argtypes = [P, ctypes.POINTER(P)]
    L.render_time_frames.argtypes = [P, P, u32, P, i32, u32, ctypes.POINTER(ctypes.c_float)]
    for name in EXPORTED:
        if name != "render_last_error":
            getattr(L, name).restype = i32
    _lib = L
    return L


def check(status: int):
    if status != RENDER_OK:
        raise RenderError(status, load().render_last_error().decode(errors="replace"))
