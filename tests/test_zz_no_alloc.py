"""The allocation-free entry point (mpm_step_noalloc) against mpm_step.

mpm_step_noalloc runs the reference-shaped substep on caller-owned device
buffers with its grid in a static device array; its particles and PosVel
must equal mpm_step's bit for bit on the same inputs, call after call, also
when a later call is smaller (the array keeps the previous call's cells)."""
import ctypes

import numpy as np
import pytest

from mlsmpm import MLSMPMSimulator, _capi, scenes

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceBuffer:
    """A hipMalloc'd buffer owned by the test (the entry point allocates nothing)."""

    def __init__(self, hip, nbytes):
        self.hip, self.ptr = hip, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(nbytes, 4))) == 0

    def put(self, a: np.ndarray):
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a: np.ndarray):
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.ptr)


@pytest.fixture(scope="module")
def hip():
    L = _capi.load()  # loads the HIP runtime the library uses
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    assert L.mpm_abi_version() == 1
    return h


def step_with_simulator(rec, box, real_box, substeps):
    with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
        if list(real_box) != list(box):
            sim.change_box_size(real_box)
        sim.execute(substeps)
        return sim.read_particles(), sim.read_posvel()


def step_noalloc(hip, dparts, dposvel, rec, box, real_box, substeps):
    L = _capi.load()
    dparts.put(rec)
    rc = L.mpm_step_noalloc(None, dparts.ptr, dposvel.ptr, rec.shape[0], _capi.f3(box),
                            _capi.f3(real_box), substeps, None)
    assert rc == 0, L.mpm_last_error()
    out = dparts.get(np.zeros_like(rec))
    pv = dposvel.get(np.zeros((rec.shape[0], 8), dtype=np.float32))
    return out, pv


def same(a, b):
    return np.array_equal(a.view(np.int32), b.view(np.int32))


def test_noalloc_step_equals_mpm_step_twice_second_smaller(hip):
    big_box, small_box = [64.0, 64.0, 64.0], [24.0, 20.0, 28.0]
    big, _ = scenes.config0()                                    # 40k particles, 64^3
    small = scenes.small_scene(3000, small_box, seed=5, velocity=1.0, affine=0.3)
    dparts = DeviceBuffer(hip, big.nbytes)
    dposvel = DeviceBuffer(hip, big.shape[0] * 32)
    try:
        for rec, box, real, sub in ((big, big_box, [64.0, 60.0, 50.5], 4),
                                    (small, small_box, small_box, 2)):
            want_p, want_pv = step_with_simulator(rec, box, real, sub)
            got_p, got_pv = step_noalloc(hip, dparts, dposvel, rec, box, real, sub)
            bad = np.nonzero(~np.all(got_p.view(np.int32) == want_p.view(np.int32), axis=1))[0]
            assert bad.size == 0, f"{rec.shape[0]} particles: {bad.size} differ, first {bad[:5]}"
            assert same(got_pv, want_pv)
            assert not same(got_p, rec)   # it did advance
    finally:
        dparts.free()
        dposvel.free()


def test_noalloc_rejects_grids_beyond_its_scratch(hip):
    L = _capi.load()
    dparts = DeviceBuffer(hip, 80)
    dposvel = DeviceBuffer(hip, 32)
    try:
        box = [65.0, 64.0, 64.0]   # 65*64*64 cells > 64^3
        rc = L.mpm_step_noalloc(None, dparts.ptr, dposvel.ptr, 1, _capi.f3(box), _capi.f3(box),
                                1, None)
        assert rc == 2 and b"64^3" in L.mpm_last_error()
        rc = L.mpm_step_noalloc(None, None, None, 1, _capi.f3([8, 8, 8]), _capi.f3([8, 8, 8]),
                                1, None)
        assert rc == 1
        # nothing to do is not an error
        assert L.mpm_step_noalloc(None, None, None, 0, _capi.f3([8, 8, 8]),
                                  _capi.f3([8, 8, 8]), 2, None) == 0
    finally:
        dparts.free()
        dposvel.free()
