"""Pins of the SPH oracle (oracle/sph_oracle.c) on the CPU.

The reference ships no SPH tests or fixtures, so the oracle is pinned by:
uniform values computed by a JavaScript engine from the reference's host
constants (tests/golden/sph_uniforms.json); the smoothing kernels'
normalisation and derivative identities; a brute-force O(n^2) neighbour sum
(the grid walk must find every neighbour and no other); two-particle closed
forms (signs and magnitudes of pressure, near pressure and viscosity); the
wall and gravity terms of integrate; and the binning as a counting sort.
"""
import ctypes
import json
import math
import os

import numpy as np
import pytest

from oracle import sph as osph

HERE = os.path.dirname(os.path.abspath(__file__))
F = np.float32


@pytest.fixture(scope="module")
def eu():
    return osph.setup()


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_uniforms_match_the_javascript_engine(eu):
    env, uni = eu
    g = json.load(open(os.path.join(HERE, "golden", "sph_uniforms.json")))
    assert (env.xGrids, env.yGrids, env.zGrids) == (g["xGrids"],) * 3
    assert env.xGrids ** 3 == g["gridCount"]
    assert F(env.cellSize) == F(g["cellSize"]) and F(env.offset) == F(g["offset"])
    assert F(env.xHalf) == F(env.yHalf) == F(env.zHalf) == F(g["half"])
    for k in ("mass", "kernelRadius", "kernelRadiusPow2", "kernelRadiusPow5", "kernelRadiusPow6",
              "kernelRadiusPow9", "dt", "stiffness", "nearStiffness", "restDensity", "viscosity"):
        assert F(getattr(uni, k)) == F(g[k]), k


def _radii(h, m=20001):
    return np.linspace(0.0, h, m)


def test_density_kernels_are_normalised(eu):
    # poly6 (315/(64 pi h^9))(h^2-r^2)^3 and spiky 15/(pi h^6)(h-r)^3 both
    # integrate to 1 over the ball of radius h.
    _, uni = eu
    h = float(uni.kernelRadius)
    r = _radii(h)
    for name in ("density_kernel", "near_density_kernel"):
        w = np.array([osph.kernel(name, uni, float(x)) for x in r], dtype=np.float64)
        integral = np.trapezoid(4 * math.pi * r * r * w, r)
        assert abs(integral - 1.0) < 2e-4, (name, integral)


def test_kernels_vanish_at_the_support_edge_and_peak_at_zero(eu):
    _, uni = eu
    h = float(np.float32(uni.kernelRadius))
    for name in ("density_kernel", "near_density_kernel", "density_gradient",
                 "near_density_gradient", "viscosity_laplacian"):
        assert abs(osph.kernel(name, uni, h)) < 1e-3 * abs(osph.kernel(name, uni, 0.0)), name
    w0 = osph.kernel("density_kernel", uni, 0.0)
    assert w0 == pytest.approx(315.0 / (64 * math.pi * h ** 3), rel=2e-6)
    s0 = osph.kernel("near_density_kernel", uni, 0.0)
    assert s0 == pytest.approx(15.0 / (math.pi * h ** 3), rel=2e-6)


def test_pressure_gradient_is_minus_the_spiky_derivative(eu):
    _, uni = eu
    h = float(uni.kernelRadius)
    d = 1e-3 * h
    for frac in (0.1, 0.3, 0.5, 0.7, 0.9):
        r = frac * h
        num = -(osph.kernel("near_density_kernel", uni, r + d) -
                osph.kernel("near_density_kernel", uni, r - d)) / (2 * d)
        assert osph.kernel("density_gradient", uni, r) == pytest.approx(num, rel=1e-3)


def test_near_gradient_uses_h5_and_laplacian_is_gradient_over_distance(eu):
    # force.wgsl:47-52 scales the near gradient by 1/h^5 (not 1/h^6): it is h
    # times the pressure gradient.  viscosityKernelLaplacian * (h - r) is the
    # pressure gradient.
    _, uni = eu
    h = float(uni.kernelRadius)
    for frac in (0.05, 0.25, 0.5, 0.75, 0.95):
        r = frac * h
        g = osph.kernel("density_gradient", uni, r)
        assert osph.kernel("near_density_gradient", uni, r) == pytest.approx(h * g, rel=3e-6)
        assert osph.kernel("viscosity_laplacian", uni, r) * (h - r) == pytest.approx(g, rel=3e-6)


def _centre(env, i):
    return (i + 0.5) * 0.07 - 2.0 - 0.14


def test_cell_id_of_cell_centres(eu):
    env, _ = eu
    rng = np.random.default_rng(5)
    cells = [(0, 0, 0), (61, 61, 61), (0, 61, 0), (61, 0, 61)] + \
        [tuple(int(x) for x in rng.integers(0, 62, 3)) for _ in range(40)]
    for i, j, k in cells:
        pos = (_centre(env, i), _centre(env, j), _centre(env, k))
        assert osph.cell_position(env, pos) == (i, j, k)
        assert osph.cell_id(env, pos) == i + 62 * j + 62 * 62 * k
    # i32 arithmetic: x coordinate -1 in row y = 1 lands on linear id 61
    assert osph.cell_id(env, (_centre(env, -1), _centre(env, 1), _centre(env, 0))) == 61
    assert osph.cell_id(env, (_centre(env, -13), _centre(env, 0), _centre(env, 0))) == -13


def _scene(pos, vel=None, half=(2.0, 2.0, 2.0)):
    rec = np.zeros((len(pos), 16), dtype=np.float32)
    rec[:, 0:3] = pos
    if vel is not None:
        rec[:, 4:7] = vel
    return osph.Scene(rec, half)


def test_binning_is_a_counting_sort_by_cell_with_decreasing_index(eu):
    env, _ = eu
    rng = np.random.default_rng(11)
    pos = rng.uniform(-0.4, 0.4, size=(3000, 3)).astype(np.float32)
    pos[:5] = [-3.0, -3.0, -3.0]               # negative linear ids: not binned
    pos[5] = [_centre(env, -1), _centre(env, 1), _centre(env, 0)]   # linear id 61
    sc = _scene(pos)
    start, order, _ = sc.bin()
    ids = np.array([osph.cell_id(env, p) for p in pos], dtype=np.int64)
    binned = (ids >= 0) & (ids < sc.G)
    assert binned.sum() == len(pos) - 5 and binned[5]
    counts = np.bincount(ids[binned], minlength=sc.G + 1)
    np.testing.assert_array_equal(start, np.concatenate([[0], np.cumsum(counts)[:-1]]))
    idx = np.nonzero(binned)[0]
    expect = idx[np.lexsort((-idx, ids[idx]))]
    np.testing.assert_array_equal(order, expect)


def _f32_r2(a, b):
    d = (a - b).astype(np.float32)
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


def _fluid_block(n, seed, spacing=0.035, amp=0.01):
    rng = np.random.default_rng(seed)
    side = int(math.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    return (g * spacing - 0.3 + rng.uniform(-amp, amp, g.shape)).astype(np.float32)


def test_density_matches_a_brute_force_sum_over_all_pairs(eu):
    _, uni = eu
    pos = _fluid_block(1728, 3)
    sc = _scene(pos)
    sc.substep()
    h, h2, m = float(uni.kernelRadius), np.float32(uni.kernelRadiusPow2), float(uni.mass)
    r2 = _f32_r2(pos[:, None, :], pos[None, :, :])
    inside = r2 < h2
    r = np.sqrt(r2.astype(np.float64))
    poly6 = 315.0 / (64 * math.pi * h ** 9) * (h * h - r * r) ** 3
    spiky = 15.0 / (math.pi * h ** 6) * (h - r) ** 3
    dens = (m * poly6 * inside).sum(1)
    near = (m * spiky * inside).sum(1)
    np.testing.assert_allclose(sc.particles[:, 11], dens, rtol=2e-5)
    np.testing.assert_allclose(sc.particles[:, 12], near, rtol=2e-5)


def _force_with_densities(sc, dens, near):
    """The oracle's computeForce on the scene's positions/velocities with the
    given densities (bypassing computeDensity)."""
    sc.particles[:, 11] = dens
    sc.particles[:, 12] = near
    start, order, _ = sc.bin()
    srt = np.ascontiguousarray(sc.particles[order])
    osph.lib().sph_oracle_force(_ptr(sc.particles), sc.n, _ptr(srt), _ptr(start),
                                ctypes.byref(sc.env), ctypes.byref(sc.uni))
    return sc.particles[:, 8:11].copy()


def test_force_matches_a_brute_force_sum_over_all_pairs(eu):
    _, uni = eu
    pos = _fluid_block(1000, 4)
    rng = np.random.default_rng(6)
    vel = rng.normal(0, 0.3, pos.shape).astype(np.float32)
    dens = rng.uniform(12000, 22000, len(pos)).astype(np.float32)
    near = rng.uniform(20000, 30000, len(pos)).astype(np.float32)
    sc = _scene(pos, vel)
    got = _force_with_densities(sc, dens, near)
    h, m = float(uni.kernelRadius), float(uni.mass)
    k, kn, rho0, mu = (float(uni.stiffness), float(uni.nearStiffness), float(uni.restDensity),
                       float(uni.viscosity))
    r2 = _f32_r2(pos[:, None, :], pos[None, :, :])
    pair = (r2 < np.float32(uni.kernelRadiusPow2)) & (r2 > 0)
    d = pos[None, :, :].astype(np.float64) - pos[:, None, :]           # x_j - x_i
    r = np.sqrt((d * d).sum(-1))
    rs = np.where(pair, r, 1.0)
    unit = d / rs[..., None]
    grad = 45.0 / (math.pi * h ** 6) * (h - rs) ** 2
    ngrad = 45.0 / (math.pi * h ** 5) * (h - rs) ** 2
    lap = 45.0 / (math.pi * h ** 6) * (h - rs)
    P = k * (dens.astype(np.float64) - rho0)
    NP = kn * near.astype(np.float64)
    shared = (P[:, None] + P[None, :]) / 2
    nshared = (NP[:, None] + NP[None, :]) / 2
    rho_j = dens.astype(np.float64)[None, :]
    nrho_j = near.astype(np.float64)[None, :]
    terms = (-m * shared / rho_j * grad)[..., None] * unit + \
        (-m * nshared / nrho_j * ngrad)[..., None] * unit + \
        mu * (m * lap / rho_j)[..., None] * (vel[None, :, :] - vel[:, None, :])
    terms = np.where(pair[..., None], terms, 0.0)
    want = terms.sum(1) + dens.astype(np.float64)[:, None] * np.array([0.0, -9.8, 0.0])
    scale = np.abs(terms).sum(1) + np.abs(dens)[:, None] * 9.8
    assert np.all(np.abs(got - want) <= 1e-5 * scale + 1e-3)


def test_two_particles_repel_along_their_axis_with_the_closed_form_magnitude(eu):
    _, uni = eu
    h = float(uni.kernelRadius)
    dist = 0.5 * h
    pos = np.array([[0.1, 0.2, 0.3], [0.1 + dist, 0.2, 0.3]], dtype=np.float32)
    vel = np.array([[0.0, 0.0, 0.0], [0.0, 1.0, 0.0]], dtype=np.float32)
    rho, nrho = 20000.0, 25000.0
    sc = _scene(pos, vel)
    f = _force_with_densities(sc, np.full(2, rho, np.float32), np.full(2, nrho, np.float32))
    r = float(np.float32(pos[1, 0]) - np.float32(pos[0, 0]))
    P = 20.0 * (rho - 15000.0)
    press = P / rho * 45.0 / (math.pi * h ** 6) * (h - r) ** 2 + \
        nrho / nrho * 45.0 / (math.pi * h ** 5) * (h - r) ** 2
    visc = 100.0 * 1.0 / rho * 45.0 / (math.pi * h ** 6) * (h - r)
    assert f[0, 0] < 0 < f[1, 0]                         # pushed apart
    assert f[0, 0] == pytest.approx(-press, rel=1e-5)
    assert f[1, 0] == pytest.approx(press, rel=1e-5)
    # viscosity drags particle 0 toward particle 1's velocity (+y) and back
    assert f[0, 1] == pytest.approx(visc - 9.8 * rho, rel=1e-5)
    assert f[1, 1] == pytest.approx(-visc - 9.8 * rho, rel=1e-5)
    assert f[0, 2] == 0.0 and f[1, 2] == 0.0


def test_isolated_particle_has_its_self_density_and_falls_freely(eu):
    _, uni = eu
    sc = _scene(np.array([[0.3, 0.5, -0.2]], dtype=np.float32))
    sc.substep()
    p = sc.particles[0]
    w0 = np.float32(uni.mass) * np.float32(osph.kernel("density_kernel", uni, 0.0))
    assert p[11] == w0
    assert tuple(p[8:11]) == (0.0, np.float32(w0 * np.float32(-9.8)), 0.0)
    dt = np.float32(uni.dt)
    vy = dt * (np.float32(w0 * np.float32(-9.8)) / w0)
    assert p[5] == vy and p[4] == 0.0 and p[6] == 0.0
    assert p[1] == np.float32(0.5) + dt * vy
    np.testing.assert_array_equal(sc.posvel[0, [0, 1, 2, 4, 5, 6]], p[[0, 1, 2, 4, 5, 6]])


def test_wall_force_pushes_back_by_stiffness_times_penetration(eu):
    _, uni = eu
    half = (0.5, 2.0, 2.0)
    rec = np.zeros((1, 16), dtype=np.float32)
    rec[0, 0:3] = (0.52, 0.0, 0.0)       # 0.02 beyond the +x wall
    rec[0, 8:11] = (0.0, 0.0, 0.0)
    rec[0, 11] = 1000.0
    sc = osph.Scene(rec, half)
    osph.lib().sph_oracle_integrate(_ptr(sc.particles), 1, osph._f3(half), ctypes.byref(sc.uni))
    dt = np.float32(uni.dt)
    pen = np.float32(np.float32(0.5) - np.float32(0.52))
    ax = np.float32(8000.0) * pen
    assert ax < 0
    assert sc.particles[0, 4] == dt * ax
    assert sc.particles[0, 0] == np.float32(0.52) + dt * (dt * ax)
    # the -x wall mirrors it
    rec[0, 0] = -0.52
    sc = osph.Scene(rec, half)
    osph.lib().sph_oracle_integrate(_ptr(sc.particles), 1, osph._f3(half), ctypes.byref(sc.uni))
    assert sc.particles[0, 4] == -(dt * ax)


def test_zero_density_particles_are_not_integrated(eu):
    rec = np.zeros((1, 16), dtype=np.float32)
    rec[0, 0:3] = (5.0, 5.0, 5.0)
    rec[0, 4:7] = (1.0, 2.0, 3.0)
    sc = osph.Scene(rec, (0.5, 0.5, 0.5))
    sc.substep()
    np.testing.assert_array_equal(sc.particles[0, 0:7], rec[0, 0:7])
    assert sc.particles[0, 11] == 0.0 and np.all(sc.particles[0, 8:11] == 0.0)


def test_thread_count_does_not_change_results():
    import sph
    p = sph.scenes.dam_break(0, seed=3)
    a = osph.Scene(p, sph.scenes.INIT_HALF_BOXES[0])
    b = osph.Scene(p, sph.scenes.INIT_HALF_BOXES[0])
    a.execute(4, threads=1)
    b.execute(4, threads=8)
    np.testing.assert_array_equal(a.particles, b.particles)
    np.testing.assert_array_equal(a.posvel, b.posvel)


def test_dambreak_matches_the_reference_loop_nest():
    import sph
    seq = iter([0.5] * 100000)
    p = sph.scenes.init_dambreak((1.0, 2.0, 1.0), 5000, rand=lambda: next(seq))
    assert p.shape == (5000, 16)
    # first particle: (-0.95, -1.9, -0.95) + 0.0005 on every axis
    np.testing.assert_array_equal(p[0, :3], np.float32([-0.95 + 0.0005, -1.9 + 0.0005,
                                                        -0.95 + 0.0005]))
    # z runs fastest over [-0.95, 0) in steps of 0.035, then x, then y
    nz = sum(1 for k in range(100) if -0.95 + k * 0.035 < 0)
    assert p[1, 2] == np.float32(-0.95 + 0.035 + 0.0005)
    assert p[nz, 0] == np.float32(-0.95 + 0.035 + 0.0005) and p[nz, 2] == p[0, 2]
    assert np.all(p[:, 2] < 0.001) and np.all(p[:, 3:] == 0)
