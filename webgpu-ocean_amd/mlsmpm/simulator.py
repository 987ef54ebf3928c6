"""Python mirror of the reference's ``MLSMPMSimulator`` (mls-mpm/mls-mpm.ts:12-276)
over the libmpm355 C ABI.  Same method names and argument meaning:

================================  ==========================================
reference (mls-mpm.ts)            here
================================  ==========================================
constructor(particleBuffer, ...)  MLSMPMSimulator(params=None)
initDambreak(initBoxSize, n)      init_dambreak(init_box_size, n)
reset(numParticles, initBoxSize)  reset(num_particles, init_box_size[, particles])
execute(commandEncoder)           execute()  -- 2 substeps, asynchronous
changeBoxSize(realBoxSize)        change_box_size(real_box_size)
numParticles / gridCount          num_particles / grid_count
(renderer reads posvelBuffer)     read_posvel() / posvel_device_ptr()
================================  ==========================================

Errors raise ``MPMError`` (the reference throws on grid overflow, :230-232).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi, scenes

PARTICLE_STRUCT_SIZE = 80  # mlsmpmParticleStructSize, mls-mpm/mls-mpm.ts:10


def default_params() -> _capi.Params:
    p = _capi.Params()
    _capi.check(_capi.load().mpm_default_params(ctypes.byref(p)))
    return p


class MLSMPMSimulator:
    def __init__(self, params: _capi.Params | None = None, particles: np.ndarray | None = None,
                 init_box_size=(64.0, 64.0, 64.0)):
        self._L = _capi.load()
        self.params = params if params is not None else default_params()
        self._h = ctypes.c_void_p()
        recs = self._records(particles)
        _capi.check(self._L.mpm_create(ctypes.byref(self.params), self._data(recs),
                                       recs.shape[0], _capi.f3(init_box_size),
                                       ctypes.byref(self._h)))
        self.init_box_size = [float(x) for x in init_box_size]
        self.real_box_size = list(self.init_box_size)

    # -- lifetime --------------------------------------------------------
    def close(self):
        if self._h:
            self._L.mpm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers ---------------------------------------------------------
    @staticmethod
    def _records(particles):
        if particles is None:
            return np.zeros((0, 20), dtype=np.float32)
        return np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 20)

    @staticmethod
    def _data(a: np.ndarray):
        return a.ctypes.data_as(ctypes.c_void_p) if a.size else None

    # -- reference surface ---------------------------------------------------
    def init_dambreak(self, init_box_size, num_particles: int, rand=None) -> np.ndarray:
        return scenes.init_dambreak(init_box_size, num_particles, rand)

    def reset(self, num_particles: int, init_box_size, particles: np.ndarray | None = None):
        if particles is None:
            particles = self.init_dambreak(init_box_size, num_particles)
        recs = self._records(particles)
        _capi.check(self._L.mpm_reset(self._h, self._data(recs), recs.shape[0],
                                      _capi.f3(init_box_size)))
        self.init_box_size = [float(x) for x in init_box_size]
        self.real_box_size = list(self.init_box_size)

    def execute(self, substeps: int = 2):
        _capi.check(self._L.mpm_step(self._h, substeps))

    def change_box_size(self, real_box_size):
        _capi.check(self._L.mpm_set_box(self._h, _capi.f3(real_box_size)))
        self.real_box_size = [float(x) for x in real_box_size]

    @property
    def num_particles(self) -> int:
        n, g = ctypes.c_uint32(), ctypes.c_uint64()
        _capi.check(self._L.mpm_get_counts(self._h, ctypes.byref(n), ctypes.byref(g)))
        return n.value

    @property
    def grid_count(self) -> int:
        n, g = ctypes.c_uint32(), ctypes.c_uint64()
        _capi.check(self._L.mpm_get_counts(self._h, ctypes.byref(n), ctypes.byref(g)))
        return g.value

    # -- readout and test hooks -------------------------------------------
    def synchronize(self):
        _capi.check(self._L.mpm_synchronize(self._h))

    def read_posvel(self) -> np.ndarray:
        out = np.zeros((self.num_particles, 8), dtype=np.float32)
        _capi.check(self._L.mpm_read_posvel(self._h, self._data(out), out.shape[0]))
        return out

    def posvel_device_ptr(self) -> int:
        p = ctypes.c_void_p()
        _capi.check(self._L.mpm_posvel_device_ptr(self._h, ctypes.byref(p)))
        return p.value or 0

    def read_particles(self) -> np.ndarray:
        out = np.zeros((self.num_particles, 20), dtype=np.float32)
        _capi.check(self._L.mpm_read_particles(self._h, self._data(out), out.shape[0]))
        return out

    def write_particles(self, particles: np.ndarray):
        recs = self._records(particles)
        _capi.check(self._L.mpm_write_particles(self._h, self._data(recs), recs.shape[0]))

    def read_grid(self) -> np.ndarray:
        out = np.zeros((self.grid_count, 4), dtype=np.int32)
        _capi.check(self._L.mpm_read_grid(self._h, self._data(out), out.shape[0]))
        return out

    def read_grid_velocity(self) -> np.ndarray:
        """(G, 4) float32: decoded grid velocities after updateGrid of the last
        production substep (x, y, z, 0)."""
        out = np.zeros((self.grid_count, 4), dtype=np.float32)
        _capi.check(self._L.mpm_read_grid_velocity(self._h, self._data(out), out.shape[0]))
        return out

    def write_grid(self, cells: np.ndarray):
        cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
        _capi.check(self._L.mpm_write_grid(self._h, self._data(cells), cells.shape[0]))

    def run_stage(self, stage: int):
        _capi.check(self._L.mpm_run_stage(self._h, stage))

    def time_steps(self, frames: int, substeps_per_frame: int = 2) -> float:
        ms = ctypes.c_float()
        _capi.check(self._L.mpm_time_steps(self._h, frames, substeps_per_frame,
                                           ctypes.byref(ms)))
        return ms.value

    def time_stages(self, substeps: int) -> dict:
        """{kernel slot: (total ms, launches)} over `substeps` eager substeps."""
        cap = 16
        ms = (ctypes.c_float * cap)()
        calls = (ctypes.c_int * cap)()
        names = (ctypes.c_char_p * cap)()
        count = ctypes.c_int()
        _capi.check(self._L.mpm_time_stages(self._h, substeps, ms, calls, names, cap,
                                            ctypes.byref(count)))
        return {names[i].decode(): (ms[i], calls[i]) for i in range(count.value)}

    @property
    def tuning(self) -> dict:
        t = _capi.Tuning()
        _capi.check(self._L.mpm_get_tuning(self._h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in _capi.Tuning._fields_}

    def set_tuning(self, **kw):
        cur = self.tuning
        cur.update(kw)
        t = _capi.Tuning(**cur)
        _capi.check(self._L.mpm_set_tuning(self._h, ctypes.byref(t)))

    def backward(self, substeps: int, target, ke_weight: float = 1e-3,
                 checkpoint_every: int = 8) -> dict:
        """Differentiable rollout from the current state (mpm_backward):
        loss = |com - target|^2 + ke_weight * mean|v|^2 after `substeps`.
        Returns loss and d loss / d (initial v, initial x, stiffness,
        viscosity); the simulator ends at the rollout's final state."""
        n = self.num_particles
        cfg = _capi.AdjointConfig(substeps, checkpoint_every,
                                  (ctypes.c_float * 3)(*[float(t) for t in target]), ke_weight)
        gv = np.zeros((n, 3), dtype=np.float32)
        gx = np.zeros((n, 3), dtype=np.float32)
        loss = ctypes.c_double()
        gk, gmu = ctypes.c_float(), ctypes.c_float()
        _capi.check(self._L.mpm_backward(self._h, ctypes.byref(cfg), self._data(gv), self._data(gx),
                                         ctypes.byref(loss), ctypes.byref(gk), ctypes.byref(gmu)))
        return {"loss": loss.value, "grad_v0": gv, "grad_x0": gx,
                "grad_stiffness": gk.value, "grad_viscosity": gmu.value}

    def autotune(self, substeps_per_trial: int = 4) -> dict:
        t = _capi.Tuning()
        _capi.check(self._L.mpm_autotune(self._h, substeps_per_trial, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in _capi.Tuning._fields_}
