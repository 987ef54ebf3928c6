"""Differentiable rollout (mpm_backward): the hand-written adjoint against
FP64 central finite differences of the unquantized model.

The FP64 rollout in oracle/mpm_oracle_f64.c is itself pinned on the CPU
against the FP32 oracle (same trajectory to fixed-point / rounding noise),
and its central differences are checked to converge in h."""
import numpy as np
import pytest

import oracle
from mlsmpm import MLSMPMSimulator, scenes

SUBSTEPS, CKPT, H = 64, 8, 1e-4


def fd_scene():
    return scenes.config4(n=1000, edge=16.0)


# --- CPU: pin the FP64 forward and the FD reference -----------------------

def test_f64_rollout_tracks_the_f32_oracle():
    rec, box, target = fd_scene()
    ref = oracle.Scene(rec, box)
    ref.execute(4)
    _, st64 = oracle.rollout_loss_f64(rec, box, 4, target, return_state=True)
    assert np.max(np.abs(st64[:, 0:3] - ref.particles[:, 0:3])) < 1e-4
    assert np.max(np.abs(st64[:, 4:7] - ref.particles[:, 4:7])) < 1e-3
    # the loss definition: |com - target|^2 + w * mean|v|^2
    loss = oracle.rollout_loss_f64(rec, box, 4, target, ke_weight=0.5)
    p = st64.astype(np.float64)
    want = np.sum((p[:, 0:3].mean(axis=0) - target) ** 2) + 0.5 * np.mean(np.sum(p[:, 4:7] ** 2, axis=1))
    assert abs(loss - want) < 1e-5 * abs(want)


def test_central_differences_converge_in_h():
    rec, box, target = fd_scene()
    v0 = rec[:, 4:7].astype(np.float64)
    u = np.random.default_rng(1).normal(size=v0.shape)
    d = []
    for h in (1e-3, 1e-4, 1e-5):
        lp = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, v0=v0 + h * u)
        lm = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, v0=v0 - h * u)
        d.append((lp - lm) / (2 * h))
    assert abs(d[1] - d[2]) < 5e-4 * abs(d[2])   # h = 1e-4 is converged well inside 1e-3
    assert abs(d[0] - d[2]) < 1e-2 * abs(d[2])


# Positions: the loss is only C1 in x0 (the quadratic B-spline's second
# derivative jumps at its knots, and g2p clamps positions), so central
# differences converge in O(h) rather than O(h^2): at h = 1e-4 they are still
# ~0.7% off on this scene.  The position check therefore uses H_X = 1e-6,
# where the differences have converged to ~1e-7 relative.
H_X = 1e-6


def test_position_central_differences_converge_in_h():
    rec, box, target = fd_scene()
    x0 = rec[:, 0:3].astype(np.float64)
    u = np.random.default_rng(4).normal(size=x0.shape)
    d = []
    for h in (1e-4, 1e-5, 3e-6, H_X):
        lp = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, x0=x0 + h * u)
        lm = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, x0=x0 - h * u)
        d.append((lp - lm) / (2 * h))
    assert abs(d[2] - d[3]) < 1e-5 * abs(d[3])   # converged at H_X
    assert abs(d[1] - d[3]) < 1e-4 * abs(d[3])
    assert abs(d[0] - d[3]) < 2e-2 * abs(d[3])   # h = 1e-4: first-order error
    # the override itself: x0 equal to the records' positions changes nothing
    assert oracle.rollout_loss_f64(rec, box, 4, target, x0=x0) == \
        oracle.rollout_loss_f64(rec, box, 4, target)


# --- GPU: the adjoint ------------------------------------------------------

def fd_dir(rec, box, target, u, h=H):
    v0 = rec[:, 4:7].astype(np.float64)
    lp = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, v0=v0 + h * u)
    lm = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, v0=v0 - h * u)
    return (lp - lm) / (2 * h)


@pytest.mark.gpu
def test_adjoint_matches_fp64_central_differences():
    rec, box, target = fd_scene()
    with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
        g = sim.backward(SUBSTEPS, target, ke_weight=1e-3, checkpoint_every=CKPT)
    loss64 = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target)
    assert abs(g["loss"] - loss64) <= 1e-3 * abs(loss64)
    rng = np.random.default_rng(2)
    tol = 1e-3   # relative, FP32 adjoint on the quantized trajectory vs FP64 FD
    for trial in range(3):
        u = rng.normal(size=(rec.shape[0], 3))
        fd = fd_dir(rec, box, target, u)
        ad = float(np.sum(g["grad_v0"].astype(np.float64) * u))
        assert abs(ad - fd) <= tol * abs(fd), f"direction {trial}: adjoint {ad} fd {fd}"
    for name, kw, base in (("grad_stiffness", "stiffness", 3.0),
                           ("grad_viscosity", "viscosity", 0.1)):
        lp = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, **{kw: base + H})
        lm = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, **{kw: base - H})
        fd = (lp - lm) / (2 * H)
        assert abs(g[name] - fd) <= tol * abs(fd), f"{name}: adjoint {g[name]} fd {fd}"
    # the largest per-particle velocity gradients, component by component
    flat = np.argsort(-np.abs(g["grad_v0"]).ravel())[:4]
    for f in flat:
        u = np.zeros(rec.shape[0] * 3)
        u[f] = 1.0
        fd = fd_dir(rec, box, target, u.reshape(-1, 3))
        ad = float(g["grad_v0"].ravel()[f])
        assert abs(ad - fd) <= 5 * tol * abs(fd), f"component {f}: adjoint {ad} fd {fd}"


def fd_dir_x(rec, box, target, u, h=H_X):
    x0 = rec[:, 0:3].astype(np.float64)
    lp = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, x0=x0 + h * u)
    lm = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, x0=x0 - h * u)
    return (lp - lm) / (2 * h)


@pytest.mark.gpu
def test_position_adjoint_matches_fp64_central_differences():
    # d loss / d x0 runs through the weight derivatives, the -w term of
    # d(dist)/dx and the clamp and wall masks of g2p.
    rec, box, target = fd_scene()
    with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
        g = sim.backward(SUBSTEPS, target, ke_weight=1e-3, checkpoint_every=CKPT)
    gx = g["grad_x0"].astype(np.float64)
    assert np.all(np.isfinite(gx)) and np.any(gx != 0)
    rng = np.random.default_rng(3)
    tol = 1e-3
    for trial in range(3):
        u = rng.normal(size=(rec.shape[0], 3))
        fd = fd_dir_x(rec, box, target, u)
        ad = float(np.sum(gx * u))
        assert abs(ad - fd) <= tol * abs(fd), f"direction {trial}: adjoint {ad} fd {fd}"
    flat = np.argsort(-np.abs(gx).ravel())[:4]
    for f in flat:
        u = np.zeros(rec.shape[0] * 3)
        u[f] = 1.0
        fd = fd_dir_x(rec, box, target, u.reshape(-1, 3))
        ad = float(gx.ravel()[f])
        assert abs(ad - fd) <= 5 * tol * abs(fd), f"component {f}: adjoint {ad} fd {fd}"


@pytest.mark.gpu
def test_checkpoint_interval_does_not_change_gradients():
    rec, box, target = fd_scene()
    res = []
    # 7 and 48 leave a shorter last segment (64 = 9 * 7 + 1 = 48 + 16); that
    # segment is taped during the forward rollout
    for k in (1, 8, 7, 48, 64):
        with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
            res.append(sim.backward(SUBSTEPS, target, checkpoint_every=k))
    # The forward recomputation is bit-exact; the adjoint's float scatters add
    # in an order that varies between runs, so per-component noise of ~1e-5
    # of the largest gradient component is allowed beside the relative bound.
    for r in res[1:]:
        assert r["loss"] == res[0]["loss"]
        for key in ("grad_v0", "grad_x0"):
            atol = 2e-5 * float(np.max(np.abs(res[0][key])))
            np.testing.assert_allclose(r[key], res[0][key], rtol=1e-4, atol=atol)
        assert abs(r["grad_stiffness"] - res[0]["grad_stiffness"]) <= 1e-4 * abs(res[0]["grad_stiffness"])


@pytest.mark.gpu
def test_backward_leaves_the_rollout_final_state():
    rec, box, target = fd_scene()
    with MLSMPMSimulator(particles=rec, init_box_size=box) as a, \
            MLSMPMSimulator(particles=rec, init_box_size=box) as b:
        a.backward(24, target, checkpoint_every=5)
        b.execute(24)
        pa, pb = a.read_particles(), b.read_particles()
        assert np.array_equal(pa.view(np.int32), pb.view(np.int32))
        assert np.array_equal(a.read_posvel().view(np.int32), b.read_posvel().view(np.int32))


@pytest.mark.gpu
def test_backward_from_8cube_tiles_runs_on_4cube_tiles_and_restores_the_tuning():
    rec, box, target = fd_scene()
    with MLSMPMSimulator(particles=rec, init_box_size=box) as a, \
            MLSMPMSimulator(particles=rec, init_box_size=box) as b:
        a.set_tuning(tile_edge=8)
        ga = a.backward(16, target, checkpoint_every=4)
        gb = b.backward(16, target, checkpoint_every=4)
        assert a.tuning["tile_edge"] == 8
        assert ga["loss"] == gb["loss"]
        for key in ("grad_v0", "grad_x0"):
            atol = 2e-5 * float(np.max(np.abs(gb[key])))
            np.testing.assert_allclose(ga[key], gb[key], rtol=1e-4, atol=atol)
        # and the 8^3 forward path continues from the rollout's final state
        a.execute(2)
        b.execute(2)
        assert np.array_equal(a.read_posvel().view(np.int32), b.read_posvel().view(np.int32))


@pytest.mark.gpu
def test_config4_full_size_gradients_match_fp64_differences_and_checkpoints():
    """The adjoint at the full config-4 size (100,000 particles, 64^3, 64
    substeps), where each 256-particle round of the i32 round-scaled scatter
    holds far larger value ranges than the 1,000-particle instance: finite,
    independent of the checkpoint interval, and equal to FP64 central
    differences of the unquantized rollout in random directions of v0 and x0
    and in the stiffness."""
    rec, box, target = scenes.config4()
    res = {}
    for k in (8, 16):
        with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
            res[k] = sim.backward(SUBSTEPS, target, ke_weight=1e-3, checkpoint_every=k)
    g = res[8]
    for key in ("grad_v0", "grad_x0"):
        assert np.all(np.isfinite(g[key])) and np.any(g[key] != 0), key
        atol = 2e-5 * float(np.max(np.abs(g[key])))
        np.testing.assert_allclose(res[16][key], g[key], rtol=1e-4, atol=atol)
    assert res[16]["loss"] == g["loss"]
    loss64 = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target)
    assert abs(g["loss"] - loss64) <= 1e-3 * abs(loss64)
    tol = 1e-3
    rng = np.random.default_rng(11)
    for trial in range(2):
        u = rng.normal(size=(rec.shape[0], 3))
        fd = fd_dir(rec, box, target, u)
        ad = float(np.sum(g["grad_v0"].astype(np.float64) * u))
        assert abs(ad - fd) <= tol * abs(fd), f"v0 direction {trial}: adjoint {ad} fd {fd}"
    u = rng.normal(size=(rec.shape[0], 3))
    fd = fd_dir_x(rec, box, target, u)
    ad = float(np.sum(g["grad_x0"].astype(np.float64) * u))
    assert abs(ad - fd) <= tol * abs(fd), f"x0 direction: adjoint {ad} fd {fd}"
    lp = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, stiffness=3.0 + H)
    lm = oracle.rollout_loss_f64(rec, box, SUBSTEPS, target, stiffness=3.0 - H)
    fd = (lp - lm) / (2 * H)
    assert abs(g["grad_stiffness"] - fd) <= tol * abs(fd), \
        f"stiffness: adjoint {g['grad_stiffness']} fd {fd}"
