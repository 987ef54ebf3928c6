"""The renderer's oracle (oracle/render_oracle.c) and camera host code
(render/camera.py) pinned by closed forms and independent implementations:
numpy's float16 conversion, the projection and view matrices' defining
properties, a lone impostor's hemisphere depth and thickness, the depth
test's order, the filters on constant and stepped images, the cube map's
face table, and the fluid pass on a flat surface facing the camera."""
import math

import numpy as np
import pytest

from oracle import render as R
from render import camera as C

W, H = 96, 64
FOV = 45 * math.pi / 180


def identity_uniforms(w=W, h=H, sphere_size=1.2, fov=FOV):
    """A camera at the origin looking down -z (view = identity)."""
    cam = C.Camera()
    cam.projection = C.perspective(fov, w / h, 0.1, 500)
    cam.view = np.eye(4).reshape(16)
    return cam.uniforms(w, h, sphere_size)


def records(pos, vel=None):
    pos = np.asarray(pos, dtype=np.float32).reshape(-1, 3)
    pv = np.zeros((len(pos), 8), dtype=np.float32)
    pv[:, 0:3] = pos
    if vel is not None:
        pv[:, 4:7] = np.asarray(vel, dtype=np.float32).reshape(-1, 3)
    return pv


def face_cube(size=8):
    """Cube map with a constant colour per face."""
    cube = np.zeros((6, size, size, 4), dtype=np.float32)
    for f in range(6):
        cube[f, :, :, 0] = 0.1 * (f + 1)
        cube[f, :, :, 1] = 0.05 * f
        cube[f, :, :, 2] = 0.9 - 0.1 * f
        cube[f, :, :, 3] = 1.0
    return cube


# ---- r16float rounding ------------------------------------------------------

def test_f16_round_equals_numpy_float16():
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2 ** 32, 200_000, dtype=np.uint64).astype(np.uint32)
    special = np.array([0.0, -0.0, 1.0, 65504.0, 65519.99, 65520.0, 1e-8, 2.0 ** -24,
                        2.0 ** -25, 3 * 2.0 ** -26, 2.0 ** -14, 0.5 + 2.0 ** -12,
                        0.5 + 3 * 2.0 ** -12, np.inf, -np.inf], dtype=np.float32)
    vals = np.concatenate([bits.view(np.float32), special,
                           rng.normal(0, 2, 50_000).astype(np.float32)])
    vals = vals[~np.isnan(vals)]
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).astype(np.float32)
    got = np.array([R.f16_round(float(v)) for v in vals], dtype=np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    assert math.isnan(R.f16_round(float("nan")))


# ---- exp and pow (one deterministic formula each, shared with the kernels) ---

def test_exp_is_within_wgsl_bounds_of_the_true_exp():
    """Within 1.5 ulp of exp on (-86, 0] (WGSL allows 3 + 2|x| ulp), exact at
    0, 0 below -86, inf above 88, NaN through."""
    rng = np.random.default_rng(3)
    x = np.concatenate([np.linspace(-85.999, 0, 40001), rng.uniform(-85.999, 0, 40000),
                        -np.logspace(-9, 1.93, 4000), rng.uniform(0, 88, 4000)]).astype(np.float32)
    got = np.array([R.exp(v) for v in x], dtype=np.float32)
    true = np.exp(x.astype(np.float64))
    ulps = np.abs(got - true) / np.spacing(true.astype(np.float32)).astype(np.float64)
    assert ulps.max() <= 1.5, (ulps.max(), x[ulps.argmax()])
    assert R.exp(0.0) == 1.0 and R.exp(-86.0) == 0.0 and R.exp(-1e4) == 0.0
    assert R.exp(100.0) == float("inf") and math.isnan(R.exp(float("nan")))


def test_log_pow5_and_pow250_are_within_wgsl_bounds():
    """log within 3 ulp; pow(x, 5) within 3 ulp; pow(x, 250) = exp(250 log x)
    within WGSL's bound for pow, inherited from exp2(y log2 x): log2 with an
    absolute error of 2^-21 on [0.5, 2] (scaled by y = 250 in the exponent)
    and exp2 within 3 + 2|a| ulp of its argument a, wherever x^250 >= 2^-124
    (below: 0, where exp flushes, or as tiny)."""
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(0, 1, 10000), rng.uniform(0.9, 1.1, 10000),
                        [1e-30, 0.5, 2.0, 3e10]]).astype(np.float32)
    got = np.array([R.log(v) for v in x], dtype=np.float32)
    true = np.log(x.astype(np.float64))
    nz = true != 0
    assert (np.abs(got - true)[nz] / np.spacing(np.abs(true).astype(np.float32))[nz]).max() <= 3
    assert R.log(1.0) == 0.0 and R.log(0.0) == float("-inf") and math.isnan(R.log(-1.0))
    x = rng.uniform(0, 2, 20000).astype(np.float32)
    got = np.array([R.pow5(v) for v in x], dtype=np.float32)
    true = x.astype(np.float64) ** 5
    assert (np.abs(got - true) / np.spacing(true.astype(np.float32))).max() <= 3
    x = np.concatenate([rng.uniform(0.5, 1, 20000), rng.uniform(0.99, 1, 5000), [1.0]]).astype(np.float32)
    got = np.array([R.pow250(v) for v in x], dtype=np.float32)
    true = x.astype(np.float64) ** 250
    big = true >= 2.0 ** -124
    a = np.abs(250 * np.log2(x.astype(np.float64)))
    allowed = (true * np.expm1(math.log(2) * 250 * 2.0 ** -21)
               + (3 + 2 * a) * np.spacing(true.astype(np.float32)))
    err = np.abs(got - true)
    assert np.all(err[big] <= allowed[big])
    # near x = 1 the bound is tightest (a chain of squarings exceeds it there)
    near1 = (x > 0.999) & big
    assert np.all(err[near1] <= allowed[near1])
    assert np.all(got[~big] <= 2.0 ** -123)   # flushed to 0 or as tiny
    assert R.pow250(1.0) == 1.0 and R.pow250(0.0) == 0.0


# ---- camera (camera.ts, wgpu-matrix 3.3.0) -----------------------------------

def test_perspective_maps_near_and_far_to_0_and_1():
    P = C.perspective(FOV, 16 / 9, 0.1, 500).reshape(4, 4).T
    for z, d in ((-0.1, 0.0), (-500.0, 1.0)):
        c = P @ np.array([0.0, 0.0, z, 1.0])
        assert c[2] / c[3] == pytest.approx(d, abs=1e-12)
    # the top of the view frustum at distance 1 maps to ndc y = 1
    c = P @ np.array([0.0, math.tan(FOV / 2), -1.0, 1.0])
    assert c[1] / c[3] == pytest.approx(1.0)
    c = P @ np.array([math.tan(FOV / 2) * 16 / 9, 0.0, -1.0, 1.0])
    assert c[0] / c[3] == pytest.approx(1.0)


def test_look_at_puts_the_eye_at_the_origin_looking_down_minus_z():
    eye, target = np.array([3.0, 4.0, 5.0]), np.array([-1.0, 2.0, 0.5])
    V = C.look_at(eye, target, [0, 1, 0]).reshape(4, 4).T
    np.testing.assert_allclose(V @ np.append(eye, 1), [0, 0, 0, 1], atol=1e-12)
    t = V @ np.append(target, 1)
    np.testing.assert_allclose(t[:2], [0, 0], atol=1e-12)
    assert t[2] == pytest.approx(-np.linalg.norm(eye - target))
    np.testing.assert_allclose(V[:3, :3] @ V[:3, :3].T, np.eye(3), atol=1e-12)
    # world up stays in the upper half of the image
    assert (V @ np.array([0, 1, 0, 0]))[1] > 0


def test_camera_reset_orbit_and_uniform_layout():
    cam = C.Camera()
    target = [20.0, 7.5, 30.0]
    cam.reset(320, 180, 70.0, target, FOV)
    V = cam.view.reshape(4, 4).T
    eye = np.linalg.inv(V) @ np.array([0, 0, 0, 1.0])
    assert np.linalg.norm(eye[:3] - target) == pytest.approx(70.0)
    # y_theta = -pi/12 below the horizon's mirror: the camera is above the target
    assert eye[1] - target[1] == pytest.approx(70.0 * math.sin(math.pi / 12))
    # x_theta = pi/4: equal x and z offsets
    assert eye[0] - target[0] == pytest.approx(eye[2] - target[2])
    u = cam.uniforms(320, 180, 1.2)
    assert u.dtype == np.float32 and u.size == 68
    np.testing.assert_array_equal(u[0:4], np.float32([1 / 320, 1 / 180, 1.2, 0]))
    for a, b in ((4, 20), (36, 52)):
        m = u[a:a + 16].reshape(4, 4).T.astype(np.float64)
        n = u[b:b + 16].reshape(4, 4).T.astype(np.float64)
        np.testing.assert_allclose(m @ n, np.eye(4), atol=2e-5)
    cam.orbit(0.0, -1e6)                    # dragged up: clamped at the horizon
    assert cam.y_theta == 0.0
    cam.orbit(0.0, 1e6)                     # dragged down: clamped above the zenith
    assert cam.y_theta == pytest.approx(-0.99 * math.pi / 2)


# ---- impostors, depth test, thickness ------------------------------------------

def lone_particle(D=20.0, x=0.0, y=0.0):
    u = identity_uniforms()
    c = R.consts(W, H, 0.6, FOV)
    return u, c, R.render_frame(records([x, y, -D]), 1, u, c, face_cube())


def test_lone_impostor_is_a_hemisphere():
    D, size = 20.0, 1.2
    u, c, im = lone_particle(D)
    P = u[20:36].reshape(4, 4).T.astype(np.float64)
    # screen rectangle of the impostor: the quad's corners through the projection
    half_w = 0.5 * size * P[0, 0] / D * W / 2
    half_h = 0.5 * size * P[1, 1] / D * H / 2
    xs, ys = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    nx = (xs - W / 2) / half_w
    ny = -(ys - H / 2) / half_h
    r2 = nx * nx + ny * ny
    inside = r2 < 1 - 1e-4
    outside = r2 > 1 + 1e-4
    covered = im["depth"] != 0
    assert covered[inside].all() and not covered[outside].any()
    n_cov = covered.sum()
    assert abs(n_cov - math.pi * half_w * half_h) < 2 * math.pi * max(half_w, half_h)
    # depth map: view-space z of the sphere's front surface
    want = -D + 0.5 * size * np.sqrt(np.clip(1 - r2, 0, None))
    np.testing.assert_allclose(im["depth"][inside], want[inside], rtol=0, atol=2e-4)
    # thickness 0.05 sqrt(1 - r^2) as r16float; its sum the hemisphere's volume
    t_want = (0.05 * np.sqrt(np.clip(1 - r2, 0, None))).astype(np.float16).astype(np.float32)
    np.testing.assert_allclose(im["thickness"][inside], t_want[inside], atol=2e-4)
    assert im["thickness"].sum() == pytest.approx(0.05 * 2 / 3 * math.pi * half_w * half_h,
                                                  rel=0.05)
    # background elsewhere; filtered depth positive (the bilateral pass takes abs)
    assert (im["thickness"][~covered] == 0).all()
    assert (im["filtered_depth"][covered] > 0).all()
    assert (im["color"][~covered][:, :3] == np.float32(0.8)).all()


def test_particles_behind_the_camera_or_off_screen_draw_nothing():
    u = identity_uniforms()
    c = R.consts(W, H, 0.6, FOV)
    pos = [[0, 0, 5.0], [0, 0, 0.0], [0, 0, -0.05], [500, 0, -20], [0, -300, -20]]
    im = R.render_frame(records(pos), len(pos), u, c, face_cube())
    assert not im["depth"].any() and not im["thickness"].any()
    im0 = R.render_frame(records(np.zeros((0, 3))), 0, u, c, face_cube())
    assert not im0["depth"].any()
    assert (im0["color"][..., :3] == np.float32(0.8)).all()


def test_depth_test_keeps_the_nearest_and_ties_go_to_the_first():
    u = identity_uniforms()
    c = R.consts(W, H, 0.6, FOV)
    near, far = [0, 0, -15.0], [0, 0, -16.0]
    a = R.render_frame(records([far, near]), 2, u, c, face_cube())
    b = R.render_frame(records([near, far]), 2, u, c, face_cube())
    np.testing.assert_array_equal(a["depth"], b["depth"])
    centre = (H // 2, W // 2)
    # the near sphere's footprint holds the far one's and its surface is in front
    one = R.render_frame(records([near]), 1, u, c, face_cube())
    np.testing.assert_array_equal(a["depth"], one["depth"])
    # the thickness pass has no depth test: both add
    assert a["thickness"][centre] > one["thickness"][centre]
    # equal depths: the earlier instance's fragment (sphere mode shows its colour)
    v_slow, v_fast = [0, 0, 0], [3.0, 0, 0]
    s1 = R.render_frame(records([near, near], [v_slow, v_fast]), 2, u, c, face_cube(), True)
    s2 = R.render_frame(records([near, near], [v_fast, v_slow]), 2, u, c, face_cube(), True)
    solo_slow = R.render_frame(records([near], [v_slow]), 1, u, c, face_cube(), True)
    solo_fast = R.render_frame(records([near], [v_fast]), 1, u, c, face_cube(), True)
    np.testing.assert_array_equal(s1["color"], solo_slow["color"])
    np.testing.assert_array_equal(s2["color"], solo_fast["color"])
    assert not np.array_equal(solo_slow["color"], solo_fast["color"])


def test_thickness_blends_in_f16_one_fragment_at_a_time():
    """The r16float target rounds after every additive blend: K fragments of
    thickness c at one pixel give the sequential f16 sum, which falls short
    of f16(K c) once the f16 spacing nears c."""
    D, K = 5.0, 300
    u = identity_uniforms()
    c = R.consts(W, H, 0.6, FOV)
    im = R.render_frame(records(np.tile([[0.0, 0.0, -D]], (K, 1))), K, u, c, face_cube())
    P = u[20:36].reshape(4, 4).T.astype(np.float64)
    half_w = 0.5 * 1.2 * P[0, 0] / D * W / 2
    half_h = 0.5 * 1.2 * P[1, 1] / D * H / 2
    y, x = H // 2, W // 2
    r2 = (0.5 / half_w) ** 2 + (0.5 / half_h) ** 2
    frag = 0.05 * math.sqrt(1 - r2)

    def sequential(cval):
        t = np.float16(0)
        for _ in range(K):
            t = np.float16(np.float32(t) + np.float32(cval))
        return np.float32(t)

    want = sequential(frag)
    assert want == sequential(frag * (1 + 1e-6)) == sequential(frag * (1 - 1e-6))
    assert im["thickness"][y, x] == want
    assert want < np.float32(np.float16(K * frag)) - 0.25


def test_sphere_mode_colours_by_speed_with_lambert_shading():
    u = identity_uniforms()
    c = R.consts(W, H, 0.6, FOV)
    D = 20.0
    im = R.render_frame(records([[0, 0, -D]], [[0, 0, 0]]), 1, u, c, face_cube(), True)
    # speed 0 -> the colour ramp's first entry (0, 0.4, 0.8), times the
    # diffuse term dot(normal, normalize(1, 1, 1)) of the sphere's normal
    P = u[20:36].reshape(4, 4).T.astype(np.float64)
    half = 0.5 * 1.2 * P[1, 1] / D * H / 2
    ys, xs = np.mgrid[H // 2 - 1:H // 2 + 1, W // 2 - 1:W // 2 + 1]
    nx = (xs + 0.5 - W / 2) / half
    ny = -(ys + 0.5 - H / 2) / half
    nz = np.sqrt(1 - nx * nx - ny * ny)
    d = np.maximum((nx + ny + nz) / math.sqrt(3), 0)[..., None]
    col = im["color"][H // 2 - 1:H // 2 + 1, W // 2 - 1:W // 2 + 1, :3]
    np.testing.assert_allclose(col, np.array([0.0, 0.4, 0.8]) * d, atol=1e-4)
    fast = R.render_frame(records([[0, 0, -D]], [[0, 0, 1.5 * 0.625]]), 1, u, c, face_cube(), True)
    # speed / 1.5 = 0.625: halfway along the ramp's third segment
    mid = (np.array([95, 254, 150]) + np.array([243, 250, 49])) / 2 / 256
    np.testing.assert_allclose(fast["color"][H // 2 - 1:H // 2 + 1, W // 2 - 1:W // 2 + 1, :3],
                               mid * d, atol=1e-4)
    assert not im["thickness"].any()


# ---- filters -------------------------------------------------------------------

def test_bilateral_keeps_constants_background_and_steps():
    c = R.consts(W, H, 0.6, FOV)
    flat = np.full((H, W), -40.0, dtype=np.float32)
    for d in ((1, 0), (0, 1)):
        out = R.bilateral(flat, d, c)
        np.testing.assert_allclose(out, 40.0, rtol=2e-6)
    img = np.zeros((H, W), dtype=np.float32)
    img[:, W // 2:] = -40.0
    img[:, W // 2 + 10:] = -60.0     # a step of 20 >> depth_threshold (6)
    out = R.bilateral(img, (1, 0), c)
    assert (out[:, :W // 2] == 0).all()
    np.testing.assert_allclose(out[:, W // 2:W // 2 + 10], 40.0, rtol=1e-4)
    np.testing.assert_allclose(out[:, W // 2 + 10:], 60.0, rtol=1e-4)
    # along y the columns are constant: nothing changes
    np.testing.assert_allclose(R.bilateral(img, (0, 1), c), np.abs(img), rtol=2e-6)


def test_bilateral_smooths_a_small_step_symmetrically():
    c = R.consts(W, H, 0.6, FOV)
    img = np.full((H, W), -40.0, dtype=np.float32)
    img[:, W // 2:] = -41.0          # within the depth threshold: blurred
    out = R.bilateral(img, (1, 0), c)
    row = out[H // 2]
    assert 40.0 < row[W // 2 - 1] < 40.5 < row[W // 2] < 41.0
    assert row[W // 2 - 1] - 40.0 == pytest.approx(41.0 - row[W // 2], rel=1e-3)


def test_gaussian_keeps_constants_and_zeros_and_conserves_a_line():
    flat = np.full((H, W), 0.25, dtype=np.float32)
    np.testing.assert_array_equal(R.gaussian(flat, (1, 0)), flat)
    img = np.zeros((H, W), dtype=np.float32)
    img[:, 40] = 1.0
    out = R.gaussian(img, (1, 0))
    # only covered pixels are written
    assert (out[:, :40] == 0).all() and (out[:, 41:] == 0).all()
    # the line's value: exp(0) over the weights of all 61 taps, sigma = 10
    k = np.arange(-30, 31)
    want = 1.0 / np.exp(-k * k / 200.0).sum()
    assert np.all(np.abs(out[:, 40] - want) <= np.spacing(np.float16(want)))
    wide = np.zeros((H, W), dtype=np.float32)
    wide[:, 30:61] = 0.5
    wide[:, 45] = 1.0
    g = R.gaussian(wide, (1, 0))
    assert g[0, 45] > g[0, 44] > g[0, 40] > g[0, 31]
    assert g[0, 44] == pytest.approx(g[0, 46], rel=1e-3)


# ---- cube map -------------------------------------------------------------------

def test_cube_map_face_table():
    cube = face_cube()
    for f, d in enumerate(([1, 0.2, -0.1], [-1, 0.1, 0.2], [0.2, 1, 0.1], [0.1, -1, 0.2],
                           [0.2, -0.1, 1], [-0.1, 0.2, -1])):
        np.testing.assert_allclose(R.cube_sample(cube, d), cube[f, 0, 0, :3], rtol=1e-6)


@pytest.mark.parametrize("face,d,sc,tc", [
    (0, (1.0, 0.5, -0.25), 0.25, -0.5), (1, (-1.0, 0.5, -0.25), -0.25, -0.5),
    (2, (0.25, 1.0, 0.5), 0.25, 0.5), (3, (0.25, -1.0, 0.5), 0.25, -0.5),
    (4, (0.25, 0.5, 1.0), 0.25, -0.5), (5, (0.25, 0.5, -1.0), -0.25, -0.5)])
def test_cube_map_orientation_and_bilinear_filter(face, d, sc, tc):
    """Face coordinates of the cube map table (sc, tc over the major axis);
    a face holding s + 2 t as its texel value returns that linear function
    at the sample point, so the filter is bilinear."""
    n = 16
    cube = np.zeros((6, n, n, 4), dtype=np.float32)
    s = (np.arange(n) + 0.5) / n
    cube[face, :, :, 0] = s[None, :] + 2 * s[:, None]
    s_want, t_want = (sc + 1) / 2, (tc + 1) / 2
    got = R.cube_sample(cube, d)[0]
    assert got == pytest.approx(s_want + 2 * t_want, abs=1e-5)


# ---- fluid pass -----------------------------------------------------------------

def test_fluid_pass_on_a_plane_facing_the_camera():
    """A flat depth map: the normal from the finite differences is (0, 0, 1),
    so the colour follows from the view ray alone: fresnel F0 + (1 - F0)
    (1 + ray.z)^5, transmittance exp(-1.5 t (1 - diffuse colour)), the
    reflected ray (rx, ry, -rz) hitting the +Z face, no specular highlight."""
    D, t = 30.0, 0.2
    u = identity_uniforms()
    depth = np.full((H, W), D, dtype=np.float32)
    thick = np.full((H, W), t, dtype=np.float32)
    cube = face_cube()
    col = R.fluid(u, depth, thick, cube)
    P = u[20:36].reshape(4, 4).T.astype(np.float64)
    ys, xs = np.mgrid[1:H - 1, 1:W - 1]
    ndc_x = (xs + 0.5) / W * 2 - 1
    ndc_y = 1 - (ys + 0.5) / H * 2
    view = np.stack([ndc_x * D / P[0, 0], ndc_y * D / P[1, 1], -np.full(xs.shape, D)], -1)
    ray = view / np.linalg.norm(view, axis=-1, keepdims=True)
    fr = np.clip(0.02 + 0.98 * (1 + ray[..., 2]) ** 5, 0, 1)[..., None]
    refr = 0.8 * np.exp(-1.5 * t * (1 - np.array([0.085, 0.6375, 0.9])))
    want = refr * (1 - fr) + cube[4, 0, 0, :3] * fr
    np.testing.assert_allclose(col[1:H - 1, 1:W - 1, :3], want, atol=2e-5)
    assert (col[..., 3] == 1).all()


def test_fluid_pass_background_and_isolated_pixels():
    u = identity_uniforms()
    depth = np.zeros((H, W), dtype=np.float32)
    depth[10, 10] = 30.0              # both horizontal neighbours background
    col = R.fluid(u, depth, np.zeros_like(depth), face_cube())
    assert (col[depth == 0][:, :3] == np.float32(0.8)).all()
    # infinite view positions make the colour NaN, which the unorm canvas stores as 0
    np.testing.assert_array_equal(col[10, 10], [0, 0, 0, 1])
