"""GPU parity of the screen-space fluid renderer (include/render.h, the
render_* entry points of libmpm355) against the CPU oracle
(oracle/render_oracle.c):

  depth map            (same fragments, same winner, same f32 math)
  thickness (r16float) (the same blends in the same draw order)
  filtered depth       (4 x bilateral; exp is one formula on both sides)
  filtered thickness   (the gaussian, r16float)
  colour               (the fluid pass; exp and pow one formula each)
  sphere mode colour
each bit for bit, pass by pass on the GPU's own previous pass and end to
end.

Scenes: the reference's startup (MLS-MPM dam break of 40,000 and 70,000
particles with its camera, main.ts:128-196) and its SPH settings (the
second renderer, radius 0.04), rendered straight from the simulators' device
PosVel buffers after some frames; a random cloud; other projections; crowded
tiles; empty and off-screen / behind-the-camera / NaN particles."""
import math

import numpy as np
import pytest

from oracle import render as R

pytestmark = pytest.mark.gpu

FOV = 45 * math.pi / 180
RADIUS = 0.6


@pytest.fixture(scope="module")
def rmod():
    import render
    render.load()
    return render


def reference_camera(rmod, box, distance, w, h):
    cam = rmod.Camera()
    cam.reset(w, h, distance, [box[0] / 2, box[1] / 4, box[2] / 2], FOV)
    return cam.uniforms(w, h, 2 * RADIUS)


def _same(got, want, what):
    g = np.ascontiguousarray(got, dtype=np.float32).view(np.uint32)
    w = np.ascontiguousarray(want, dtype=np.float32).view(np.uint32)
    bad = np.count_nonzero(g != w)
    if bad:
        i = np.unravel_index(np.argmax(g != w), g.shape)
        raise AssertionError(f"{what}: {bad} values differ, first at {i}: "
                             f"{np.asarray(got)[i]} vs {np.asarray(want)[i]}")


def check_frame(r, pv, n, u, cube, sphere=False, what=""):
    """The GPU frame in `r` (already executed) against the oracle, pass by
    pass on the GPU's own previous pass (so a difference names its pass),
    then end to end: every image bit for bit."""
    w, h = r.width, r.height
    c = R.consts(w, h, r.radius, r.fov)
    want = R.render_frame(pv, n, u, c, cube, sphere, threads=16)
    depth = r.read_image("depth")
    _same(depth, want["depth"], what + " depth")
    color = r.read_image("color")
    if sphere:
        _same(color, want["color"], what + " sphere colour")
        return None
    thick = r.read_image("thickness")
    _same(thick, want["thickness"], what + " thickness")
    fd = r.read_image("filtered_depth")
    ref = depth
    for _ in range(4):
        ref = R.bilateral(R.bilateral(ref, (1, 0), c, 16), (0, 1), c, 16)
    _same(fd, ref, what + " filtered depth")
    ft = r.read_image("filtered_thickness")
    _same(ft, R.gaussian(R.gaussian(thick, (1, 0), 16), (0, 1), 16), what + " filtered thickness")
    _same(color, R.fluid(u, fd, ft, cube, 16), what + " colour")
    _same(fd, want["filtered_depth"], what + " filtered depth, end to end")
    _same(color, want["color"], what + " colour, end to end")
    return depth


@pytest.mark.parametrize("preset", [0, 1])
def test_reference_dam_break_rendered_from_the_simulator(rmod, preset):
    from mlsmpm import MLSMPMSimulator
    n = [40_000, 70_000][preset]
    box = [[35, 25, 55], [40, 30, 60]][preset]
    dist = [60, 70][preset]
    w, h = 400, 240
    cube = rmod.synthetic_cubemap(32)
    u = reference_camera(rmod, box, dist, w, h)
    with MLSMPMSimulator(init_box_size=box) as sim, \
            rmod.FluidRenderer(w, h, RADIUS, FOV, cube) as r:
        sim.reset(n, box)
        for _ in range(20):
            sim.execute(2)
        sim.synchronize()
        pv = sim.read_posvel()
        r.execute(sim.posvel_device_ptr(), n, u)
        depth = check_frame(r, pv, n, u, cube, what=f"dam break {n}")
        assert np.count_nonzero(depth) > 0.05 * w * h
        r.execute(sim.posvel_device_ptr(), n, u, sphere_render=True)
        check_frame(r, pv, n, u, cube, sphere=True, what=f"dam break {n} spheres")


@pytest.mark.parametrize("setting", [0, 3])
def test_sph_frame_rendered_from_the_simulator(rmod, setting):
    """The reference's second renderer instance (radius 0.04, main.ts:142-149)
    over the SPH simulator's device PosVel buffer, its camera as on the SPH
    buttons (distance 2.6-3.8, target (0, -half_y + 0.1, 0), main.ts:231-234),
    sphere size = the SPH diameter 0.08."""
    import sph
    half = sph.scenes.INIT_HALF_BOXES[setting]
    dist = [2.6, 3.0, 3.4, 3.8][setting]
    rec = sph.scenes.dam_break(setting, seed=1)
    w, h = 320, 200
    cube = rmod.synthetic_cubemap(16)
    cam = rmod.Camera()
    cam.reset(w, h, dist, [0.0, -half[1] + 0.1, 0.0], FOV)
    u = cam.uniforms(w, h, 0.08)
    with sph.SPHSimulator(particles=rec, init_half_box_size=half) as sim, \
            rmod.FluidRenderer(w, h, 0.04, FOV, cube) as r:
        for _ in range(20):
            sim.execute()
        sim.synchronize()
        pv = sim.read_posvel()
        n = len(pv)
        r.execute(sim.posvel_device_ptr(), n, u)
        depth = check_frame(r, pv, n, u, cube, what=f"sph setting {setting}")
        assert np.count_nonzero(depth) > 0.02 * w * h
        r.execute(sim.posvel_device_ptr(), n, u, sphere_render=True)
        check_frame(r, pv, n, u, cube, sphere=True, what=f"sph setting {setting} spheres")


def test_random_cloud_and_camera_orbit(rmod):
    rng = np.random.default_rng(5)
    n = 6000
    pv = np.zeros((n, 8), dtype=np.float32)
    pv[:, 0:3] = rng.uniform([2, 2, 2], [30, 20, 45], (n, 3))
    pv[:, 4:7] = rng.normal(0, 1.0, (n, 3))
    w, h = 320, 200
    cube = rmod.synthetic_cubemap(16)
    cam = rmod.Camera()
    cam.reset(w, h, 60, [17.5, 6.25, 27.5], FOV)
    with rmod.FluidRenderer(w, h, RADIUS, FOV, cube) as r:
        for dx, dy in ((0, 0), (180, -40), (-400, 150)):
            cam.orbit(dx, dy)
            u = cam.uniforms(w, h, 2 * RADIUS)
            r.execute(pv, n, u)
            check_frame(r, pv, n, u, cube, what=f"cloud orbit {dx},{dy}")
            r.execute(pv, n, u, sphere_render=True)
            check_frame(r, pv, n, u, cube, sphere=True, what=f"cloud spheres {dx},{dy}")


def test_orthographic_and_off_centre_projections(rmod):
    """Projections outside the perspective form: an orthographic one (clip w
    = 1) and an off-centre, sheared perspective (x and y depend on z)."""
    rng = np.random.default_rng(11)
    n = 4000
    pv = np.zeros((n, 8), dtype=np.float32)
    pv[:, 0:3] = rng.uniform([-12, -8, -60], [12, 8, -20], (n, 3))
    w, h = 240, 160
    cube = rmod.synthetic_cubemap(16)
    cam = rmod.Camera()
    cam.view = np.eye(4).reshape(16)
    near, far = 0.1, 500.0
    ortho = np.zeros(16)
    ortho[0], ortho[5] = 1 / 15.0, 1 / 10.0
    ortho[10], ortho[14], ortho[15] = 1 / (near - far), near / (near - far), 1.0
    sheared = rmod.perspective(FOV, w / h, near, far)
    sheared[8], sheared[9] = 0.15, -0.1
    with rmod.FluidRenderer(w, h, RADIUS, FOV, cube) as r:
        for name, P in (("orthographic", ortho), ("off-centre", sheared)):
            cam.projection = P
            u = cam.uniforms(w, h, 2 * RADIUS)
            for sphere in (False, True):
                r.execute(pv, n, u, sphere_render=sphere)
                d = check_frame(r, pv, n, u, cube, sphere=sphere, what=f"{name} {sphere}")
                if d is not None:
                    assert np.count_nonzero(d) > 0.05 * w * h


def test_close_up_filter_sizes(rmod):
    """Particles close to the camera (filter sizes up to the cap of 100
    taps) beside far ones (about 10 taps): the bilateral filter's blocks of
    64 x 4 pixels whose largest filter exceeds the 40-pixel LDS apron gather
    their taps from global memory, the others read them from LDS; both
    kinds occur in this frame, and in blocks straddling the image edges."""
    rng = np.random.default_rng(13)
    n = 600
    pv = np.zeros((n, 8), dtype=np.float32)
    near = np.arange(n) < 12  # a few, in the upper left of the image
    z = np.where(near, rng.uniform(-3.0, -1.5, n), rng.uniform(-14.0, -8.0, n))
    t = math.tan(FOV / 2)
    w, h = 240, 160
    sx = np.where(near, rng.uniform(-1.1, -0.6, n), rng.uniform(-1.1, 1.1, n))
    sy = np.where(near, rng.uniform(0.5, 1.1, n), rng.uniform(-1.1, 1.1, n))
    pv[:, 0] = sx * -z * t * w / h
    pv[:, 1] = sy * -z * t
    pv[:, 2] = z
    pv[:, 4:7] = rng.normal(0, 1.0, (n, 3))
    cube = rmod.synthetic_cubemap(16)
    cam = rmod.Camera()
    cam.reset(w, h, 60, [0.0, 0.0, 0.0], FOV)
    cam.view = np.eye(4).reshape(16)
    u = cam.uniforms(w, h, 2 * RADIUS)
    with rmod.FluidRenderer(w, h, RADIUS, FOV, cube) as r:
        r.execute(pv, n, u)
        d = check_frame(r, pv, n, u, cube, what="close-up")
    ppc = 12 * 2 * RADIUS * 0.05 * (h / 2) / math.tan(FOV / 2)
    depth = np.abs(d)
    act = (depth > 0) & (depth < 1e4)
    fs = np.where(act, np.minimum(100, np.ceil(ppc / np.where(act, depth, 1))), -1)
    blocks = np.array([fs[i:i + 4, j:j + 64].max() for i in range(0, h, 4) for j in range(0, w, 64)])
    assert np.any(blocks > 40) and np.any((blocks > 0) & (blocks <= 40))


@pytest.mark.parametrize("n", [6000, 20000])
def test_crowded_tiles(rmod, n):
    """Thousands of impostors over a few screen tiles: tile lists longer than
    the small and the large LDS list (the latter taken in windows of particle
    indices), still in draw order."""
    rng = np.random.default_rng(n)
    w, h = 160, 120
    cube = rmod.synthetic_cubemap(8)
    cam = rmod.Camera()
    cam.reset(w, h, 60, [0.0, 0.0, 0.0], FOV)
    u = cam.uniforms(w, h, 2 * RADIUS)
    pv = np.zeros((n, 8), dtype=np.float32)
    d = rng.normal(0, 1, (n, 3))
    pv[:, 0:3] = 0.4 * d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0, 1, (n, 1))
    pv[:, 4:7] = rng.normal(0, 1, (n, 3))
    with rmod.FluidRenderer(w, h, RADIUS, FOV, cube) as r:
        r.execute(pv, n, u)
        check_frame(r, pv, n, u, cube, what=f"crowded {n}")
        assert r.read_image("thickness").max() > 10
        r.execute(pv, n, u, sphere_render=True)
        check_frame(r, pv, n, u, cube, sphere=True, what=f"crowded {n} spheres")


def test_no_particles_and_particles_that_draw_nothing(rmod):
    w, h = 128, 96
    cube = rmod.synthetic_cubemap(8)
    cam = rmod.Camera()
    cam.reset(w, h, 50, [0, 0, 0], FOV)
    u = cam.uniforms(w, h, 2 * RADIUS)
    V = cam.view.reshape(4, 4).T
    eye = (np.linalg.inv(V) @ np.array([0, 0, 0, 1.0]))[:3]
    fwd = -V[2, :3]
    pos = [eye - 5 * fwd,                  # behind the camera
           eye + 0.05 * fwd,               # nearer than the near plane
           eye + 600 * fwd,                # beyond the far plane
           [1e4, 0, 0], [0, -1e4, 0],      # off screen
           [np.nan, 0, 0], [np.inf, 1, 1]]
    pv = np.zeros((len(pos), 8), dtype=np.float32)
    pv[:, 0:3] = np.array(pos, dtype=np.float32)
    with rmod.FluidRenderer(w, h, RADIUS, FOV, cube) as r:
        for n in (0, len(pos)):
            r.execute(pv, n, u)
            assert not r.read_image("depth").any() and not r.read_image("thickness").any()
            col = r.read_image("color")
            assert (col[..., :3] == np.float32(0.8)).all() and (col[..., 3] == 1).all()
            check_frame(r, pv, n, u, cube, what=f"nothing drawn n={n}")
        # one visible particle among them
        pv2 = np.concatenate([pv, np.zeros((1, 8), np.float32)])
        r.execute(pv2, len(pv2), u)
        assert np.count_nonzero(r.read_image("depth")) > 0
        check_frame(r, pv2, len(pv2), u, cube, what="one visible")


def test_device_and_host_inputs_give_the_same_frame(rmod):
    """The same records from the simulator's device buffer and from host
    memory give the same frame; the output image keeps its device address."""
    from mlsmpm import MLSMPMSimulator
    box = [35, 25, 55]
    w, h = 200, 120
    cube = rmod.synthetic_cubemap(16)
    u = reference_camera(rmod, box, 60, w, h)
    with MLSMPMSimulator(init_box_size=box) as sim, \
            rmod.FluidRenderer(w, h, RADIUS, FOV, cube) as r:
        sim.reset(5000, box)
        sim.execute(2)
        sim.synchronize()
        pv = sim.read_posvel()
        r.execute(pv, len(pv), u)
        a = {k: r.read_image(k) for k in ("depth", "filtered_depth", "thickness", "color")}
        ptr = r.color_device_ptr()
        r.execute(sim.posvel_device_ptr(), len(pv), u)
        assert r.color_device_ptr() == ptr != 0
        for k, v in a.items():   # deterministic: the same bits
            np.testing.assert_array_equal(r.read_image(k).view(np.uint32), v.view(np.uint32),
                                          err_msg=k)
        assert r._L.render_execute(r._h, sim.posvel_device_ptr() + 4, 1, u.ctypes.data, 0) == 1
        assert b"aligned" in r._L.render_last_error()


def test_invalid_arguments_are_errors_not_crashes(rmod):
    from render import _capi
    u = np.zeros(68, np.float32)
    with rmod.FluidRenderer(64, 48) as r:
        buf = np.zeros(64 * 48 * 4, np.float32)
        assert r._L.render_read_image(r._h, 9, buf.ctypes.data) == 1
        assert b"image" in r._L.render_last_error()
        assert r._L.render_execute(r._h, None, 5, u.ctypes.data, 0) == 1
        assert r._L.render_execute(r._h, None, 0, None, 0) == 1
        with pytest.raises(ValueError):
            r.execute(np.zeros((2, 8), np.float32), 3, u)
    with pytest.raises(_capi.RenderError):
        rmod.FluidRenderer(0, 10)
