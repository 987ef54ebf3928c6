"""Screen-space fluid renderer (render/fluidRender.ts) on gfx950: the
FluidRenderer mirror, the camera and the render uniforms."""
from .camera import Camera, look_at, perspective, synthetic_cubemap  # noqa: F401
from .renderer import FluidRenderer  # noqa: F401
from . import _capi  # noqa: F401


def load():
    return _capi.load()
