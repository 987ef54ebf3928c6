"""FluidRenderer: the Python mirror of render/fluidRender.ts's class over the
C ABI of include/render.h (gfx950 kernels in libmpm355.so)."""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _capi
from .camera import synthetic_cubemap


class FluidRenderer:
    """new FluidRenderer(device, canvas, format, radius, fov, posvelBuffer,
    renderUniformBuffer, cubemapTextureView) (fluidRender.ts:35); the PosVel
    buffer and the uniforms are passed per execute() instead of being bound."""

    def __init__(self, width: int, height: int, radius: float = 0.6,
                 fov: float = 45 * math.pi / 180, cubemap: np.ndarray | None = None):
        self._L = _capi.load()
        self.width, self.height = int(width), int(height)
        self.radius, self.fov = float(radius), float(fov)
        cube = synthetic_cubemap() if cubemap is None else cubemap
        cube = np.ascontiguousarray(cube, dtype=np.float32)
        if cube.ndim != 4 or cube.shape[0] != 6 or cube.shape[1] != cube.shape[2] or cube.shape[3] != 4:
            raise ValueError("cubemap must be (6, size, size, 4) float32")
        self.cube_size = cube.shape[1]
        prm = _capi.Params(self.width, self.height, self.radius, self.fov)
        h = ctypes.c_void_p()
        _capi.check(self._L.render_create(ctypes.byref(prm), cube.ctypes.data_as(ctypes.c_void_p),
                                          self.cube_size, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.render_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def execute(self, posvel, num_particles: int, uniforms: np.ndarray, sphere_render: bool = False):
        """FluidRenderer.execute(context, encoder, numParticles, sphereRenderFl)
        (fluidRender.ts:332).  posvel: a device address (int, e.g.
        MLSMPMSimulator.posvel_device_ptr()) or host records (n, 8) float32."""
        u = np.ascontiguousarray(uniforms, dtype=np.float32)
        if u.size != 68:
            raise ValueError("uniforms must be the 68-float render uniform buffer")
        up = u.ctypes.data_as(ctypes.c_void_p)
        if isinstance(posvel, np.ndarray):
            pv = np.ascontiguousarray(posvel, dtype=np.float32).reshape(-1, 8)
            if pv.shape[0] < num_particles:
                raise ValueError("fewer PosVel records than num_particles")
            _capi.check(self._L.render_execute_host(self._h, pv.ctypes.data_as(ctypes.c_void_p),
                                                    num_particles, up, int(sphere_render)))
        else:
            _capi.check(self._L.render_execute(self._h, ctypes.c_void_p(int(posvel)),
                                               num_particles, up, int(sphere_render)))

    def read_image(self, name: str = "color") -> np.ndarray:
        """An image of the last frame: (height, width) float32, or (height,
        width, 4) for "color"."""
        shape = (self.height, self.width, 4) if name == "color" else (self.height, self.width)
        out = np.empty(shape, dtype=np.float32)
        _capi.check(self._L.render_read_image(self._h, _capi.IMAGES[name],
                                              out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def color_device_ptr(self) -> int:
        """Device address of the output image (RGBA f32, rows top to bottom)."""
        p = ctypes.c_void_p()
        _capi.check(self._L.render_color_device_ptr(self._h, ctypes.byref(p)))
        return p.value or 0

    def time_frames(self, posvel_device: int, num_particles: int, uniforms: np.ndarray,
                    frames: int, sphere_render: bool = False) -> float:
        """Milliseconds of `frames` frames on the device (events)."""
        u = np.ascontiguousarray(uniforms, dtype=np.float32)
        ms = ctypes.c_float()
        _capi.check(self._L.render_time_frames(self._h, ctypes.c_void_p(int(posvel_device)),
                                               num_particles, u.ctypes.data_as(ctypes.c_void_p),
                                               int(sphere_render), frames, ctypes.byref(ms)))
        return float(ms.value)
