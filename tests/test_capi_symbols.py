"""The C-ABI library loads and exports every entry point include/mpm.h
declares; without a GPU the library reports MPM_ERR_NO_DEVICE instead of
falling back to anything.  No compute runs here."""
import ctypes
import os
import re
import subprocess

import pytest

import mlsmpm
from mlsmpm import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mpm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(_capi.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = mlsmpm.load()
    for name in declared_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", mlsmpm.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(mpm_\w+)", out))
    assert set(declared_functions()) <= exported


def test_abi_version_and_default_params():
    L = mlsmpm.load()
    assert L.mpm_abi_version() == 1
    p = mlsmpm.Params()
    assert L.mpm_default_params(ctypes.byref(p)) == 0
    assert (p.stiffness, p.rest_density, p.dt, p.fixed_point_multiplier) == (3.0, 4.0,
                                                                             ctypes.c_float(0.2).value, 1e7)
    assert abs(p.dynamic_viscosity - 0.1) < 1e-7


def test_null_arguments_are_rejected_with_status_not_crash():
    L = mlsmpm.load()
    assert L.mpm_create(None, None, 0, _capi.f3([8, 8, 8]), None) == 1
    assert b"NULL" in L.mpm_last_error()
    assert L.mpm_step(None, 2) == 1
    assert L.mpm_destroy(None) == 0
    assert L.mpm_default_params(None) == 1


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_is_reported_not_faked():
    with pytest.raises(mlsmpm.MPMError) as ei:
        mlsmpm.MLSMPMSimulator(init_box_size=(8, 8, 8))
    assert ei.value.status == 5


SPH_HEADER = os.path.join(ROOT, "include", "sph.h")


def sph_declared_functions():
    text = open(SPH_HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z_0-9]+)\s*\(", text)))


def test_sph_header_binding_and_library_agree():
    import sph
    from sph import _capi as sph_capi
    assert sph_declared_functions() == sorted(sph_capi.EXPORTED)
    L = sph.load()
    out = subprocess.run(["nm", "-D", "--defined-only", sph.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(sph_\w+)", out))
    for name in sph_declared_functions():
        assert hasattr(L, name) and name in exported, name


def test_sph_default_params_and_null_arguments():
    import sph
    p = sph.default_params()
    assert (p.kernel_radius, p.mass, p.dt, p.stiffness, p.near_stiffness, p.rest_density,
            p.viscosity) == (0.07, 1.0, 0.006, 20.0, 1.0, 15000.0, 100.0)
    L = sph.load()
    assert L.sph_create(None, None, 0, None, None) == 1
    assert b"NULL" in L.sph_last_error()
    assert L.sph_step(None, 2) == 1
    assert L.sph_destroy(None) == 0
    assert L.sph_default_params(None) == 1


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_sph_no_device_is_reported_not_faked():
    import sph
    with pytest.raises(sph.SPHError) as ei:
        sph.SPHSimulator()
    assert ei.value.status == 5


RENDER_HEADER = os.path.join(ROOT, "include", "render.h")


def render_declared_functions():
    text = open(RENDER_HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(render_[a-z_0-9]+)\s*\(", text)))


def test_render_header_binding_and_library_agree():
    import render
    from render import _capi as render_capi
    assert render_declared_functions() == sorted(render_capi.EXPORTED)
    L = render.load()
    out = subprocess.run(["nm", "-D", "--defined-only", render_capi.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(render_\w+)", out))
    for name in render_declared_functions():
        assert hasattr(L, name) and name in exported, name


def test_render_arguments_are_checked_before_the_device():
    import render
    from render import _capi as render_capi
    L = render.load()
    out = ctypes.c_void_p()
    prm = render_capi.Params(64, 48, 0.6, 0.785)
    assert L.render_create(None, None, 0, ctypes.byref(out)) == 1
    assert b"NULL" in L.render_last_error()
    assert L.render_create(ctypes.byref(render_capi.Params(0, 48, 0.6, 0.785)), None, 0,
                           ctypes.byref(out)) == 1
    assert L.render_create(ctypes.byref(render_capi.Params(64, 48, 0.6, 4.0)), None, 0,
                           ctypes.byref(out)) == 1
    assert L.render_create(ctypes.byref(prm), None, 0, ctypes.byref(out)) == 1   # no cube map
    assert L.render_execute(None, None, 0, None, 0) == 1
    assert L.render_read_image(None, 0, None) == 1
    assert L.render_destroy(None) == 0


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_render_no_device_is_reported_not_faked():
    import numpy as np
    import render
    with pytest.raises(render._capi.RenderError) as ei:
        render.FluidRenderer(64, 48, cubemap=np.ones((6, 2, 2, 4), np.float32))
    assert ei.value.status == 5
