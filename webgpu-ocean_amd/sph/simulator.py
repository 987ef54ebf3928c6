"""Python mirror of the reference's ``SPHSimulator`` (sph/sph.ts:15-371) over
the SPH entry points of libmpm355 (include/sph.h).  Same method names and
argument meaning:

==================================  =========================================
reference (sph/sph.ts)              here
==================================  =========================================
constructor(particleBuffer, ...)    SPHSimulator(params=None)
initDambreak(initHalfBoxSize, n)    init_dambreak(init_half_box_size, n)
reset(numParticles, initHalfBox)    reset(num_particles, init_half_box_size[, particles])
execute(commandEncoder)             execute()  -- 2 substeps, asynchronous
changeBoxSize(realBoxSize)          change_box_size(real_box_size)
numParticles / gridCount            num_particles / grid_count
(renderer reads posvelBuffer)       read_posvel() / posvel_device_ptr()
==================================  =========================================

Errors raise ``SPHError``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi, scenes

PARTICLE_STRUCT_SIZE = 64  # sphParticleStructSize, sph/sph.ts:13


def default_params() -> _capi.Params:
    p = _capi.Params()
    _capi.check(_capi.load().sph_default_params(ctypes.byref(p)))
    return p


def _records(particles) -> np.ndarray:
    if particles is None:
        return np.zeros((0, scenes.PARTICLE_FLOATS), dtype=np.float32)
    return np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, scenes.PARTICLE_FLOATS)


def _data(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p) if a.size else None


class SPHSimulator:
    kernel_radius = 0.07  # sph/sph.ts:41

    def __init__(self, params: _capi.Params | None = None, particles=None,
                 init_half_box_size=(1.0, 2.0, 1.0)):
        self._L = _capi.load()
        self.params = params if params is not None else default_params()
        self.kernel_radius = float(self.params.kernel_radius)
        self._h = ctypes.c_void_p()
        recs = _records(particles)
        _capi.check(self._L.sph_create(ctypes.byref(self.params), _data(recs), recs.shape[0],
                                       _capi.f3(init_half_box_size), ctypes.byref(self._h)))
        self.init_half_box_size = [float(x) for x in init_half_box_size]
        self.real_box_size = list(self.init_half_box_size)

    # -- lifetime --------------------------------------------------------
    def close(self):
        if self._h:
            self._L.sph_destroy(self._h)
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

    # -- reference surface -------------------------------------------------
    def init_dambreak(self, init_half_box_size, num_particles: int, rand=None) -> np.ndarray:
        return scenes.init_dambreak(init_half_box_size, num_particles, self.kernel_radius, rand)

    def reset(self, num_particles: int, init_half_box_size, particles=None):
        """reset (sph/sph.ts:281-299): builds the dam break unless `particles`
        (n x 16 f32 records) are given; the real box becomes the initial one."""
        if particles is None:
            particles = self.init_dambreak(init_half_box_size, num_particles)
        recs = _records(particles)
        _capi.check(self._L.sph_reset(self._h, _data(recs), recs.shape[0],
                                      _capi.f3(init_half_box_size)))
        self.init_half_box_size = [float(x) for x in init_half_box_size]
        self.real_box_size = list(self.init_half_box_size)

    def execute(self, substeps: int = 2):
        """execute (sph/sph.ts:301-336): two substeps per frame; asynchronous."""
        _capi.check(self._L.sph_step(self._h, substeps))

    def change_box_size(self, real_box_size):
        """changeBoxSize (sph/sph.ts:365-370): half extents of the walls."""
        _capi.check(self._L.sph_set_box(self._h, _capi.f3(real_box_size)))
        self.real_box_size = [float(x) for x in real_box_size]

    @property
    def num_particles(self) -> int:
        return self._counts()[0]

    @property
    def grid_count(self) -> int:
        return self._counts()[1]

    # -- readout and hooks -------------------------------------------------
    def _counts(self):
        n, g = ctypes.c_uint32(), ctypes.c_uint64()
        _capi.check(self._L.sph_get_counts(self._h, ctypes.byref(n), ctypes.byref(g)))
        return n.value, g.value

    def synchronize(self):
        _capi.check(self._L.sph_synchronize(self._h))

    def run_stage(self, stage):
        if isinstance(stage, str):
            stage = _capi.STAGES.index(stage)
        _capi.check(self._L.sph_run_stage(self._h, int(stage)))

    def read_posvel(self) -> np.ndarray:
        n = self.num_particles
        out = np.zeros((n, 8), dtype=np.float32)
        _capi.check(self._L.sph_read_posvel(self._h, _data(out), n))
        return out

    def posvel_device_ptr(self) -> int:
        p = ctypes.c_void_p()
        _capi.check(self._L.sph_posvel_device_ptr(self._h, ctypes.byref(p)))
        return p.value or 0

    def read_particles(self) -> np.ndarray:
        n = self.num_particles
        out = np.zeros((n, scenes.PARTICLE_FLOATS), dtype=np.float32)
        _capi.check(self._L.sph_read_particles(self._h, _data(out), n))
        return out

    def write_particles(self, particles):
        recs = _records(particles)
        _capi.check(self._L.sph_write_particles(self._h, _data(recs), recs.shape[0]))

    def read_cell_start(self) -> np.ndarray:
        g = self.grid_count + 1
        out = np.zeros(g, dtype=np.uint32)
        _capi.check(self._L.sph_read_cell_start(self._h, _data(out), g))
        return out

    def read_sorted_order(self, count: int | None = None) -> np.ndarray:
        count = self.num_particles if count is None else count
        out = np.zeros(count, dtype=np.uint32)
        _capi.check(self._L.sph_read_sorted_order(self._h, _data(out), count))
        return out

    def time_steps(self, substeps: int) -> float:
        ms = ctypes.c_float()
        _capi.check(self._L.sph_time_steps(self._h, substeps, ctypes.byref(ms)))
        return ms.value

    def time_stages(self, substeps: int) -> dict:
        ms = (ctypes.c_float * len(_capi.STAGES))()
        _capi.check(self._L.sph_time_stages(self._h, substeps, ms))
        return {name: float(ms[i]) for i, name in enumerate(_capi.STAGES)}
