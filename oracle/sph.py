"""CPU oracle for the SPH step -- TEST INFRASTRUCTURE ONLY.

ctypes binding of the ``sph_oracle_*`` functions of ``oracle/liboracle.so``
(built from ``sph_oracle.c`` by ``oracle/Makefile``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import it; it is the checker of the HIP library, never part of the product
path.  Particles are ``(n, 16) float32`` arrays in the reference's 64-B layout
(sph/sph.ts:348-354); posvel ``(n, 8) float32``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as _mpm_lib


class Params(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in (
        "kernel_radius", "mass", "dt", "stiffness", "near_stiffness", "rest_density",
        "viscosity")]


class Env(ctypes.Structure):
    _fields_ = [("xGrids", ctypes.c_int32), ("yGrids", ctypes.c_int32),
                ("zGrids", ctypes.c_int32)] + [
        (k, ctypes.c_float) for k in ("cellSize", "xHalf", "yHalf", "zHalf", "offset")]


class Uniforms(ctypes.Structure):
    _fields_ = [(k, ctypes.c_float) for k in (
        "mass", "kernelRadius", "kernelRadiusPow2", "kernelRadiusPow5", "kernelRadiusPow6",
        "kernelRadiusPow9", "dt", "stiffness", "nearStiffness", "restDensity", "viscosity")]


def default_params(**over) -> Params:
    """Constants of sph/sph.ts:41,72-77."""
    p = Params(0.07, 1.0, 0.006, 20.0, 1.0, 15000.0, 100.0)
    for k, v in over.items():
        setattr(p, k, v)
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = _mpm_lib()
        P, u32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float
        E, U = ctypes.POINTER(Env), ctypes.POINTER(Uniforms)
        F3 = ctypes.POINTER(ctypes.c_float)
        L.sph_oracle_set_threads.argtypes = [ctypes.c_int]
        L.sph_oracle_setup.argtypes = [ctypes.POINTER(Params), E, U]
        L.sph_oracle_grid_count.restype = ctypes.c_uint64
        L.sph_oracle_grid_count.argtypes = [E]
        L.sph_oracle_cell_position.argtypes = [E, F3, ctypes.POINTER(ctypes.c_int32)]
        L.sph_oracle_cell_id.restype = ctypes.c_int32
        L.sph_oracle_cell_id.argtypes = [E, F3]
        for k in ("density_kernel", "near_density_kernel", "density_gradient",
                  "near_density_gradient", "viscosity_laplacian"):
            fn = getattr(L, "sph_oracle_" + k)
            fn.restype = f32
            fn.argtypes = [U, f32]
        L.sph_oracle_bin.restype = u32
        L.sph_oracle_bin.argtypes = [P, u32, E, P, P, P]
        L.sph_oracle_density.argtypes = [P, u32, P, P, E, U]
        L.sph_oracle_force.argtypes = [P, u32, P, P, E, U]
        L.sph_oracle_integrate.argtypes = [P, u32, F3, U]
        L.sph_oracle_substep.argtypes = [P, u32, F3, E, U, P]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def setup(params: Params | None = None):
    env, uni = Env(), Uniforms()
    lib().sph_oracle_setup(ctypes.byref(params or default_params()), ctypes.byref(env),
                           ctypes.byref(uni))
    return env, uni


def kernel(name: str, uni: Uniforms, r: float) -> float:
    return float(getattr(lib(), "sph_oracle_" + name)(ctypes.byref(uni), r))


def cell_id(env: Env, pos) -> int:
    return int(lib().sph_oracle_cell_id(ctypes.byref(env), _f3(pos)))


def cell_position(env: Env, pos):
    c = (ctypes.c_int32 * 3)()
    lib().sph_oracle_cell_position(ctypes.byref(env), _f3(pos), c)
    return tuple(c)


class Scene:
    """Oracle SPH scene: particles (n,16) f32, posvel (n,8) f32."""

    def __init__(self, particles: np.ndarray, init_half_box, params: Params | None = None):
        self.particles = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 16).copy()
        self.n = self.particles.shape[0]
        self.params = params or default_params()
        self.env, self.uni = setup(self.params)
        self.G = int(lib().sph_oracle_grid_count(ctypes.byref(self.env)))
        self.real_half = [float(np.float32(b)) for b in init_half_box]
        self.posvel = np.zeros((self.n, 8), dtype=np.float32)

    def change_box_size(self, real_half):
        self.real_half = [float(np.float32(b)) for b in real_half]

    def bin(self):
        """(start[G+1], order[m], sorted particles[m]) of the current state."""
        start = np.zeros(self.G + 1, dtype=np.uint32)
        order = np.zeros(max(self.n, 1), dtype=np.uint32)
        srt = np.zeros((max(self.n, 1), 16), dtype=np.float32)
        m = lib().sph_oracle_bin(_ptr(self.particles), self.n, ctypes.byref(self.env),
                                 _ptr(start), _ptr(order), _ptr(srt))
        return start, order[:m].copy(), srt[:m].copy()

    def substep(self, threads: int = 1):
        L = lib()
        L.sph_oracle_set_threads(int(threads))
        L.sph_oracle_substep(_ptr(self.particles), self.n, _f3(self.real_half),
                             ctypes.byref(self.env), ctypes.byref(self.uni), _ptr(self.posvel))

    def execute(self, substeps: int = 2, threads: int = 1):
        for _ in range(substeps):
            self.substep(threads)
