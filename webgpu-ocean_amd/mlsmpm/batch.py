"""Batched scenes over libmpm355's mpm_batch_* ABI: many independent MLS-MPM
scenes (own particle count, box and parameters) advanced by one launch per
stage.  Each scene's results equal a single MLSMPMSimulator's."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi


class MLSMPMBatch:
    def __init__(self, scenes):
        """scenes: iterable of (particles (n,20) f32, init_box[3], params | None)."""
        self._L = _capi.load()
        self._keep = []
        arr = (_capi.Scene * len(scenes))()
        for i, (rec, box, params) in enumerate(scenes):
            rec = np.ascontiguousarray(rec, dtype=np.float32).reshape(-1, 20)
            self._keep.append(rec)
            if params is None:
                params = _capi.Params()
                _capi.check(self._L.mpm_default_params(ctypes.byref(params)))
            arr[i].particles = rec.ctypes.data_as(ctypes.c_void_p) if rec.size else None
            arr[i].n = rec.shape[0]
            arr[i].init_box = (ctypes.c_float * 3)(*[float(b) for b in box])
            arr[i].params = params
        self._h = ctypes.c_void_p()
        _capi.check(self._L.mpm_batch_create(arr, len(scenes), ctypes.byref(self._h)))
        self._keep = []
        self.num_scenes = len(scenes)

    def close(self):
        if self._h:
            self._L.mpm_batch_destroy(self._h)
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

    def step(self, substeps: int = 2):
        _capi.check(self._L.mpm_batch_step(self._h, substeps))

    def change_box_size(self, scene: int, real_box):
        _capi.check(self._L.mpm_batch_set_box(self._h, scene, _capi.f3(real_box)))

    def synchronize(self):
        _capi.check(self._L.mpm_batch_synchronize(self._h))

    def counts(self, scene: int):
        n, g = ctypes.c_uint32(), ctypes.c_uint64()
        _capi.check(self._L.mpm_batch_get_counts(self._h, scene, ctypes.byref(n),
                                                 ctypes.byref(g)))
        return n.value, g.value

    def read_posvel(self, scene: int) -> np.ndarray:
        n, _ = self.counts(scene)
        out = np.zeros((n, 8), dtype=np.float32)
        _capi.check(self._L.mpm_batch_read_posvel(self._h, scene,
                                                  out.ctypes.data_as(ctypes.c_void_p) if n else None, n))
        return out

    def read_particles(self, scene: int) -> np.ndarray:
        n, _ = self.counts(scene)
        out = np.zeros((n, 20), dtype=np.float32)
        _capi.check(self._L.mpm_batch_read_particles(self._h, scene,
                                                     out.ctypes.data_as(ctypes.c_void_p) if n else None, n))
        return out

    @property
    def tuning(self) -> dict:
        t = _capi.Tuning()
        _capi.check(self._L.mpm_batch_get_tuning(self._h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in _capi.Tuning._fields_}

    def set_tuning(self, **kw):
        cur = self.tuning
        cur.update(kw)
        t = _capi.Tuning(**cur)
        _capi.check(self._L.mpm_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, substeps_per_trial: int = 4) -> dict:
        t = _capi.Tuning()
        _capi.check(self._L.mpm_batch_autotune(self._h, substeps_per_trial, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in _capi.Tuning._fields_}

    def time_steps(self, substeps: int) -> float:
        ms = ctypes.c_float()
        _capi.check(self._L.mpm_batch_time_steps(self._h, substeps, ctypes.byref(ms)))
        return ms.value

    def time_stages(self, substeps: int) -> dict:
        cap = 16
        ms = (ctypes.c_float * cap)()
        calls = (ctypes.c_int * cap)()
        names = (ctypes.c_char_p * cap)()
        count = ctypes.c_int()
        _capi.check(self._L.mpm_batch_time_stages(self._h, substeps, ms, calls, names, cap,
                                                  ctypes.byref(count)))
        return {names[i].decode(): (ms[i], calls[i]) for i in range(count.value)}
