"""CPU oracle for the MLS-MPM substep -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/liboracle.so`` (built from ``mpm_oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this package; it is the checker
of the HIP library, never part of the product path.

Arrays use the reference layouts (numpy structured as float32 / int32):
particles ``(n, 20) float32`` (80-B records), cells ``(G, 4) int32``,
posvel ``(n, 8) float32``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class Params(ctypes.Structure):
    _fields_ = [
        ("stiffness", ctypes.c_float),
        ("rest_density", ctypes.c_float),
        ("dynamic_viscosity", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("fixed_point_multiplier", ctypes.c_float),
    ]


def default_params(**over) -> Params:
    """Constants of mls-mpm/mls-mpm.ts:53-59."""
    p = Params(3.0, 4.0, 0.1, 0.2, 1e7)
    for k, v in over.items():
        setattr(p, k, v)
    return p


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        F3 = ctypes.POINTER(ctypes.c_float)
        PP = ctypes.POINTER(Params)
        u32, u64 = ctypes.c_uint32, ctypes.c_uint64
        L.oracle_grid_count.restype = u64
        L.oracle_grid_count.argtypes = [F3]
        L.oracle_clear_grid.argtypes = [P, u64]
        L.oracle_p2g_1.argtypes = [P, u32, P, u64, F3, PP]
        L.oracle_p2g_2.argtypes = [P, u32, P, u64, F3, PP]
        L.oracle_update_grid.argtypes = [P, u64, F3, F3, PP]
        L.oracle_g2p.argtypes = [P, u32, P, u64, F3, F3, PP]
        L.oracle_copy_position.argtypes = [P, u32, P]
        L.oracle_substep.argtypes = [P, u32, P, u64, P, F3, F3, PP]
        L.oracle_substep_omp.argtypes = [P, u32, P, u64, P, F3, F3, PP, ctypes.c_int]
        L.oracle_encode.restype = ctypes.c_int32
        L.oracle_encode.argtypes = [ctypes.c_float, ctypes.c_float]
        L.oracle_decode.restype = ctypes.c_float
        L.oracle_decode.argtypes = [ctypes.c_int32, ctypes.c_float]
        L.oracle_f32_to_i32.restype = ctypes.c_int32
        L.oracle_f32_to_i32.argtypes = [ctypes.c_float]
        L.oracle_pow5.restype = ctypes.c_float
        L.oracle_pow5.argtypes = [ctypes.c_float]
        L.oracle_weights.argtypes = [ctypes.c_float, F3]
        D3 = ctypes.POINTER(ctypes.c_double)
        L.oracle_rollout_loss_f64.restype = ctypes.c_double
        L.oracle_rollout_loss_f64.argtypes = [P, P, P, u32, F3, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, u32, D3,
                                              ctypes.c_double, P]
        _lib = L
    return _lib


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def grid_count(init_box) -> int:
    return int(lib().oracle_grid_count(_f3(init_box)))


class Scene:
    """Oracle scene: particles (n,20) f32, cells (G,4) i32, posvel (n,8) f32."""

    def __init__(self, particles: np.ndarray, init_box, params: Params | None = None):
        self.particles = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 20).copy()
        self.n = self.particles.shape[0]
        self.init_box = [float(np.float32(b)) for b in init_box]
        self.real_box = list(self.init_box)
        self.params = params or default_params()
        self.G = grid_count(self.init_box)
        self.cells = np.zeros((self.G, 4), dtype=np.int32)
        self.posvel = np.zeros((self.n, 8), dtype=np.float32)

    def change_box_size(self, real_box):
        self.real_box = [float(np.float32(b)) for b in real_box]

    def run_stage(self, stage: int):
        L, p = lib(), ctypes.byref(self.params)
        ib, rb = _f3(self.init_box), _f3(self.real_box)
        if stage == 0:
            L.oracle_clear_grid(_ptr(self.cells), self.G)
        elif stage == 1:
            L.oracle_p2g_1(_ptr(self.particles), self.n, _ptr(self.cells), self.G, ib, p)
        elif stage == 2:
            L.oracle_p2g_2(_ptr(self.particles), self.n, _ptr(self.cells), self.G, ib, p)
        elif stage == 3:
            L.oracle_update_grid(_ptr(self.cells), self.G, rb, ib, p)
        elif stage == 4:
            L.oracle_g2p(_ptr(self.particles), self.n, _ptr(self.cells), self.G, rb, ib, p)
        elif stage == 5:
            L.oracle_copy_position(_ptr(self.particles), self.n, _ptr(self.posvel))
        else:
            raise ValueError(stage)

    def substep(self, threads: int = 0):
        L, p = lib(), ctypes.byref(self.params)
        ib, rb = _f3(self.init_box), _f3(self.real_box)
        if threads == 0:
            L.oracle_substep(_ptr(self.particles), self.n, _ptr(self.cells), self.G,
                             _ptr(self.posvel), rb, ib, p)
        else:
            L.oracle_substep_omp(_ptr(self.particles), self.n, _ptr(self.cells), self.G,
                                 _ptr(self.posvel), rb, ib, p, int(threads))

    def execute(self, substeps: int = 2, threads: int = 0):
        for _ in range(substeps):
            self.substep(threads)


def encode(f: float, mult: float = 1e7) -> int:
    return int(lib().oracle_encode(f, mult))


def decode(i: int, mult: float = 1e7) -> float:
    return float(lib().oracle_decode(i, mult))


def f32_to_i32(f: float) -> int:
    return int(lib().oracle_f32_to_i32(f))


def pow5(r: float) -> float:
    return float(lib().oracle_pow5(r))


def weights(x: float) -> np.ndarray:
    w = (ctypes.c_float * 3)()
    lib().oracle_weights(x, w)
    return np.array(list(w), dtype=np.float32)


def rollout_loss_f64(particles, box, substeps, target, ke_weight=1e-3, v0=None,
                     stiffness=3.0, viscosity=0.1, rest_density=4.0, dt=0.2,
                     return_state=False, x0=None):
    """FP64 unquantized rollout (oracle_rollout_loss_f64): the loss whose
    central differences check the HIP adjoint.  x0 / v0 (n x 3 float64)
    replace the records' initial positions / velocities."""
    rec = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 20)
    n = rec.shape[0]
    v = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64).reshape(n, 3)
    x = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).reshape(n, 3)
    tgt = (ctypes.c_double * 3)(*[float(t) for t in target])
    out = np.zeros((n, 20), dtype=np.float32) if return_state else None
    loss = lib().oracle_rollout_loss_f64(
        _ptr(rec), None if x is None else _ptr(x), None if v is None else _ptr(v), n, _f3(box),
        stiffness, viscosity, rest_density, dt, substeps, tgt, ke_weight,
        None if out is None else _ptr(out))
    return (loss, out) if return_state else loss
