"""CPU oracle for the screen-space fluid renderer -- TEST INFRASTRUCTURE ONLY.

ctypes binding of the ``oracle_render_*`` functions of ``oracle/liboracle.so``
(built from ``render_oracle.c`` by ``oracle/Makefile``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py`` may import it; it is the checker
of the HIP renderer, never part of the product path.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import lib as _mpm_lib


class Uniforms(ctypes.Structure):
    _fields_ = [("f", ctypes.c_float * 68)]


class Consts(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("depth_threshold", ctypes.c_float), ("max_filter_size", ctypes.c_float),
                ("projected_particle_constant", ctypes.c_float)]


class Images(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in (
        "depth", "filtered_depth", "thickness", "filtered_thickness", "color")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = _mpm_lib()
        P, f32, i32 = ctypes.c_void_p, ctypes.c_float, ctypes.c_int
        L.oracle_f16_round.restype = f32
        L.oracle_f16_round.argtypes = [f32]
        L.oracle_render_frame.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(Uniforms),
                                          ctypes.POINTER(Consts), P, i32, i32,
                                          ctypes.POINTER(Images), i32]
        L.oracle_cube_sample.argtypes = [P, i32, ctypes.POINTER(f32), ctypes.POINTER(f32)]
        L.oracle_bilateral_pass.argtypes = [P, P, i32, i32, ctypes.POINTER(f32),
                                            ctypes.POINTER(Consts), i32]
        L.oracle_gaussian_pass.argtypes = [P, P, i32, i32, ctypes.POINTER(f32), i32]
        L.oracle_fluid_pass.argtypes = [ctypes.POINTER(Uniforms), P, P, P, i32, i32, i32, P, i32]
        for name in ("oracle_render_exp", "oracle_render_log", "oracle_render_pow5",
                     "oracle_render_pow250"):
            getattr(L, name).restype = f32
            getattr(L, name).argtypes = [f32]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def consts(width: int, height: int, radius: float, fov: float) -> Consts:
    """The filter constants of fluidRender.ts:38-52 (JavaScript numbers, f32
    as pipeline constants)."""
    diameter = 2 * radius
    return Consts(width, height, radius * 10, 100.0,
                  (12 * diameter * 0.05 * (height / 2)) / math.tan(fov / 2))


def uniforms(u: np.ndarray) -> Uniforms:
    out = Uniforms()
    out.f[:] = [float(x) for x in np.asarray(u, dtype=np.float32).reshape(68)]
    return out


def f16_round(x: float) -> float:
    return float(lib().oracle_f16_round(float(x)))


def cube_sample(cube: np.ndarray, d) -> np.ndarray:
    c = np.ascontiguousarray(cube, dtype=np.float32)
    dv = (ctypes.c_float * 3)(*[float(v) for v in d])
    out = (ctypes.c_float * 3)()
    lib().oracle_cube_sample(_p(c), c.shape[1], dv, out)
    return np.array(out[:], dtype=np.float32)


def bilateral(src: np.ndarray, direction, c: Consts, threads: int = 0) -> np.ndarray:
    s = np.ascontiguousarray(src, dtype=np.float32)
    out = np.empty_like(s)
    dv = (ctypes.c_float * 2)(*[float(v) for v in direction])
    lib().oracle_bilateral_pass(_p(s), _p(out), s.shape[1], s.shape[0], dv, ctypes.byref(c),
                                threads)
    return out


def gaussian(src: np.ndarray, direction, threads: int = 0) -> np.ndarray:
    s = np.ascontiguousarray(src, dtype=np.float32)
    out = np.empty_like(s)
    dv = (ctypes.c_float * 2)(*[float(v) for v in direction])
    lib().oracle_gaussian_pass(_p(s), _p(out), s.shape[1], s.shape[0], dv, threads)
    return out


def render_frame(posvel: np.ndarray, n: int, u: np.ndarray, c: Consts, cube: np.ndarray,
                 sphere_mode: bool = False, threads: int = 0) -> dict:
    """One FluidRenderer.execute frame: dict of depth, filtered_depth,
    thickness, filtered_thickness (h, w) and color (h, w, 4)."""
    w, h = c.width, c.height
    imgs = {k: np.zeros((h, w), dtype=np.float32) for k in
            ("depth", "filtered_depth", "thickness", "filtered_thickness")}
    imgs["color"] = np.zeros((h, w, 4), dtype=np.float32)
    pv = np.ascontiguousarray(posvel, dtype=np.float32).reshape(-1, 8)
    cb = np.ascontiguousarray(cube, dtype=np.float32)
    im = Images(*[_p(imgs[k]) for k in ("depth", "filtered_depth", "thickness",
                                         "filtered_thickness", "color")])
    uu = uniforms(u)
    lib().oracle_render_frame(_p(pv), n, ctypes.byref(uu), ctypes.byref(c), _p(cb), cb.shape[1],
                              int(sphere_mode), ctypes.byref(im), threads)
    return imgs


def fluid(u: np.ndarray, depth: np.ndarray, thickness: np.ndarray, cube: np.ndarray,
          threads: int = 0) -> np.ndarray:
    """The fluid pass alone over given filtered depth and thickness maps:
    (h, w, 4) colour."""
    d = np.ascontiguousarray(depth, dtype=np.float32)
    t = np.ascontiguousarray(thickness, dtype=np.float32)
    cb = np.ascontiguousarray(cube, dtype=np.float32)
    out = np.zeros(d.shape + (4,), dtype=np.float32)
    uu = uniforms(u)
    lib().oracle_fluid_pass(ctypes.byref(uu), _p(d), _p(t), _p(cb), cb.shape[1], d.shape[1],
                            d.shape[0], _p(out), threads)
    return out


def exp(x: float) -> float:
    """The renderer's deterministic exp (the HIP kernels use the same formula)."""
    return float(lib().oracle_render_exp(float(x)))


def log(x: float) -> float:
    return float(lib().oracle_render_log(float(x)))


def pow5(x: float) -> float:
    return float(lib().oracle_render_pow5(float(x)))


def pow250(x: float) -> float:
    return float(lib().oracle_render_pow250(float(x)))
