"""GPU parity: libmpm355 (through its C ABI) against the CPU oracle.

Bar: i32 grid accumulators bit-exact; particle state and PosVel bit-exact too
(both sides evaluate the same IEEE binary32 expressions in the same order,
with no contraction), so any difference is a bug rather than a tolerance
question.
"""
import numpy as np
import pytest

import oracle
from mlsmpm import MLSMPMSimulator, MPMError, scenes

pytestmark = pytest.mark.gpu


def pair(rec, box):
    return MLSMPMSimulator(particles=rec, init_box_size=box), oracle.Scene(rec, box)


def bits_equal(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Element-wise bit equality of float32 arrays, except that every NaN
    equals every NaN (IEEE 754 leaves NaN payload and sign propagation open)."""
    return (a.view(np.int32) == b.view(np.int32)) | (np.isnan(a) & np.isnan(b))


def assert_same(sim, ref, what=("grid", "particles", "posvel"), ctx=""):
    if "grid" in what:
        g = sim.read_grid()
        bad = np.nonzero(np.any(g != ref.cells, axis=1))[0]
        assert bad.size == 0, f"{ctx}: {bad.size} cells differ, first {bad[:5]}: " \
                              f"{g[bad[:3]].tolist()} vs {ref.cells[bad[:3]].tolist()}"
    if "particles" in what:
        p = sim.read_particles()
        bad = np.nonzero(~np.all(bits_equal(p, ref.particles), axis=1))[0]
        assert bad.size == 0, f"{ctx}: {bad.size} particles differ, first {bad[:5]}: " \
                              f"{p[bad[:2]].tolist()} vs {ref.particles[bad[:2]].tolist()}"
    if "posvel" in what:
        pv = sim.read_posvel()
        assert np.all(bits_equal(pv, ref.posvel)), ctx


SMALL = [
    # n, box, seed, velocity, affine
    (1, [8.0, 8.0, 8.0], 0, 0.0, 0.0),
    (65, [16.0, 16.0, 16.0], 1, 1.0, 0.5),
    (1000, [24.0, 20.0, 28.0], 2, 2.0, 0.3),
    (5000, [32.0, 32.0, 32.0], 3, 0.5, 0.1),
]


@pytest.mark.parametrize("n,box,seed,vel,aff", SMALL)
def test_each_reference_stage_bit_exact(n, box, seed, vel, aff):
    rec = scenes.small_scene(n, box, seed=seed, velocity=vel, affine=aff) if n > 1 else \
        scenes.records(np.array([[3.3, 4.6, 2.2]]), np.array([[0.5, -0.25, 1.0]]))
    sim, ref = pair(rec, box)
    with sim:
        for sub in range(3):
            for stage, what in [(0, ("grid",)), (1, ("grid",)), (2, ("grid",)),
                                (3, ("grid",)), (4, ("particles",)), (5, ("posvel",))]:
                sim.run_stage(stage)
                ref.run_stage(stage)
                assert_same(sim, ref, what, ctx=f"substep {sub} stage {stage}")


@pytest.mark.parametrize("n,box,seed,vel,aff", SMALL)
def test_execute_matches_oracle_substeps(n, box, seed, vel, aff):
    rec = scenes.small_scene(max(n, 2), box, seed=seed, velocity=vel, affine=aff)
    sim, ref = pair(rec, box)
    with sim:
        for frame in range(3):
            sim.execute()
            ref.execute(2)
            assert_same(sim, ref, ("particles", "posvel"), ctx=f"frame {frame}")


def test_config0_rollout_bit_exact_with_box_resize():
    rec, box = scenes.config0()
    sim, ref = pair(rec, box)
    with sim:
        for frame in range(12):
            if frame == 6:
                sim.change_box_size([64.0, 64.0, 48.5])
                ref.change_box_size([64.0, 64.0, 48.5])
            sim.execute()
            ref.execute(2)
        assert_same(sim, ref, ("particles", "posvel"), ctx="cfg0 frame 12")


def test_config1_full_size_substeps_bit_exact():
    rec, box = scenes.config1()
    sim, ref = pair(rec, box)
    with sim:
        for frame in range(2):
            sim.execute()
            ref.execute(2, threads=8)
        assert_same(sim, ref, ("particles", "posvel"), ctx="cfg1 frame 2")
        # stage path from the same state: P2G accumulators bit-exact at 300k
        for stage in (0, 1, 2):
            sim.run_stage(stage)
            ref.run_stage(stage)
        assert_same(sim, ref, ("grid",), ctx="cfg1 p2g grid")
        mass = int(ref.cells[:, 3].astype(np.int64).sum())
        n = rec.shape[0]
        assert n * 10**7 - 27 * n <= mass <= n * 10**7


def _production_step_matches_oracle_stages(sim, box, real_box, ctx, threads=0):
    """One production substep (mpm_step(1) on the default path) from the
    simulator's current state, checked against the oracle's stages from the
    same state: the raw P2G accumulators that mpm_read_grid returns (clear +
    p2g_1 + p2g_2, mls-mpm/p2g_1.wgsl:57-60, p2g_2.wgsl:87-89) bit for bit,
    the decoded updateGrid velocities (updateGrid.wgsl:26-43) bit for bit,
    then g2p + copyPosition."""
    ref = oracle.Scene(sim.read_particles(), box)
    ref.change_box_size(real_box)
    sim.execute(1)
    g = sim.read_grid()
    vel = sim.read_grid_velocity()
    for stage in (0, 1, 2):
        ref.run_stage(stage)
    bad = np.nonzero(np.any(g != ref.cells, axis=1))[0]
    assert bad.size == 0, f"{ctx}: {bad.size} raw accumulators differ, first {bad[:5]}: " \
                          f"{g[bad[:3]].tolist()} vs {ref.cells[bad[:3]].tolist()}"
    assert np.count_nonzero(g[:, 3]) > 0
    ref.run_stage(3)
    want = ref.cells[:, :3].astype(np.float32) / np.float32(1e7)   # decodeFixedPoint
    bad = np.nonzero(~np.all(bits_equal(vel[:, :3], want), axis=1))[0]
    assert bad.size == 0, f"{ctx}: {bad.size} updated velocities differ, first {bad[:5]}: " \
                          f"{vel[bad[:3]].tolist()} vs {want[bad[:3]].tolist()}"
    assert np.all(vel[:, 3] == 0)
    ref.run_stage(4)
    ref.run_stage(5)
    assert_same(sim, ref, ("particles", "posvel"), ctx=ctx)


@pytest.mark.parametrize("cfg,frames,resize,edge", [
    ("config0", 3, None, 4), ("config1", 2, None, 4), ("config1", 2, [64.0, 80.0, 80.0], 4),
    ("config2", 1, None, 4), ("config0", 3, None, 8), ("config1", 2, [64.0, 80.0, 80.0], 8),
    ("config2", 1, None, 8)],
    ids=["40k", "300k", "300k-resized", "2M", "40k-8^3", "300k-resized-8^3", "2M-8^3"])
def test_production_grid_accumulators_bit_exact(cfg, frames, resize, edge):
    rec, box = getattr(scenes, cfg)()
    with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
        sim.set_tuning(tile_edge=edge)
        sim.execute(2 * frames)
        real = list(box)
        if resize is not None:
            sim.change_box_size(resize)
            real = resize
        for k in range(2):
            _production_step_matches_oracle_stages(sim, box, real, f"{cfg} substep {k}")


def test_rollouts_while_the_host_is_busy_match_the_oracle():
    """Regression: many-substep calls made while every host core is busy (the
    oracle's own threads) must still give the oracle's result.  A counting
    variant that published the next binning's counts with atomics from g2p
    failed exactly this way (lost particles in most such rollouts)."""
    import threading

    rec, box = scenes.config1()
    substeps = 200
    ref = oracle.Scene(rec, box)
    worker = threading.Thread(target=ref.execute, args=(substeps, 16))
    worker.start()
    finals = []
    try:
        while worker.is_alive() or not finals:
            with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
                for _ in range(substeps // 20):
                    sim.execute(20)
                    sim.synchronize()
                finals.append(sim.read_particles())
            if len(finals) >= 12:
                break
    finally:
        worker.join()
    for k, p in enumerate(finals):
        bad = np.nonzero(~np.all(bits_equal(p, ref.particles), axis=1))[0]
        assert bad.size == 0, f"rollout {k} of {len(finals)}: {bad.size} particles differ"


def test_config2_two_million_mass_and_one_substep():
    rec, box = scenes.config2()
    sim, ref = pair(rec, box)
    with sim:
        sim.run_stage(0)
        sim.run_stage(1)
        g = sim.read_grid()
        n = rec.shape[0]
        mass = int(g[:, 3].astype(np.int64).sum())
        assert n * 10**7 - 27 * n <= mass <= n * 10**7
        assert g[:, 3].max() < 2**31 - 1
        sim.write_particles(rec)
        sim.execute(1)
        ref.substep(threads=8)
        assert_same(sim, ref, ("particles", "posvel"), ctx="cfg2 substep 1")


def _dense_scene():
    rng = np.random.default_rng(11)
    box = [24.0, 20.0, 24.0]
    dense = rng.uniform([9.0, 6.0, 9.0], [11.0, 8.0, 11.0], size=(1800, 3))  # 8 cells
    plane = np.stack([np.full(700, 14.0), rng.uniform(4.0, 7.0, 700),
                      rng.uniform(4.0, 20.0, 700)], axis=1)  # one x plane, like a clamp
    spread = rng.uniform(3.5, 16.5, size=(500, 3))
    pos = np.concatenate([dense, plane, spread]).astype(np.float32)
    vel = rng.normal(0.0, 0.3, size=pos.shape).astype(np.float32)
    return scenes.records(pos, vel), box


@pytest.mark.parametrize("ppt", [1, 2, 4])
def test_dense_cells_match_the_oracle(ppt):
    """Hundreds of particles per cell (the pile-up at a wall after a box
    shrink): tiles split into several chunks, the P2G level order runs past
    its level-major part into the crowded cells' cell-major tail (whole waves
    of one cell, added through the window replicas), and the i32 mass
    accumulators wrap.  Integer adds wrap the same on both sides, so results
    stay bit-exact."""
    rec, box = _dense_scene()
    sim, ref = pair(rec, box)
    with sim:
        sim.set_tuning(particles_per_thread=ppt)
        for frame in range(2):
            sim.execute()
            ref.execute(2)
            assert_same(sim, ref, ("particles", "posvel"), ctx=f"ppt {ppt} frame {frame}")


@pytest.mark.parametrize("edge", [4, 8])
def test_dense_cells_production_grid_accumulators_bit_exact(edge):
    """The raw P2G accumulators of crowded chunks (cells of hundreds of
    particles: replicated LDS windows on 4^3 tiles) equal the oracle's, bit
    for bit, after a frame and on the next two substeps."""
    rec, box = _dense_scene()
    with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
        sim.set_tuning(tile_edge=edge)
        sim.execute(2)
        for k in range(2):
            _production_step_matches_oracle_stages(sim, box, box, f"dense edge {edge} substep {k}")


def test_particles_outside_the_grid_are_dropped_like_the_oracle():
    box = [12.0, 12.0, 12.0]
    pos = np.array([[0.2, 5.0, 5.0], [11.8, 5.0, 5.0], [5.0, 5.0, 11.9], [5.0, -3.0, 5.0],
                    [5.5, 5.5, 5.5], [40.0, 40.0, 40.0], [np.nan, 5.0, 5.0],
                    [5.0, np.inf, 5.0], [3e6, 5.0, 5.0], [5.2, 5.7, 6.1]])
    rec = scenes.records(pos, np.full((len(pos), 3), 0.25))
    sim, ref = pair(rec, box)
    with sim:
        for stage in range(6):
            sim.run_stage(stage)
            ref.run_stage(stage)
        assert_same(sim, ref, ("grid", "particles", "posvel"), ctx="out of grid")
        sim.execute()
        ref.execute(2)
        assert_same(sim, ref, ("particles", "posvel"), ctx="out of grid step")


@pytest.mark.parametrize("path", [0, 1])
def test_saturating_encode_and_non_finite_velocities(path):
    # |w * v * 1e7| beyond 2^31 saturates; NaN converts to 0 (i32(f32) in WGSL)
    box = [16.0, 16.0, 16.0]
    pos = np.array([[5.5, 5.5, 5.5], [5.7, 5.2, 5.9], [8.1, 8.3, 8.2], [8.4, 8.9, 8.6],
                    [10.2, 6.6, 7.7], [6.1, 9.5, 10.4], [7.3, 7.3, 7.3], [9.9, 9.1, 6.2]])
    vel = np.array([[1e6, -1e6, 300.0], [-300.0, 2e9, -2e9], [np.inf, 0.0, -np.inf],
                    [np.nan, 1.0, 1.0], [214.0, -214.5, 215.0], [-3e38, 3e38, 0.5],
                    [0.0, 0.0, 0.0], [1.0, np.nan, -1.0]])
    rec = scenes.records(pos, vel)
    rec[:, 8] = np.array([1e7, -1e7, 0, 0, 3e3, 0, np.nan, 0], dtype=np.float32)  # C00
    sim, ref = pair(rec, box)
    with sim:
        sim.set_tuning(path=path)
        for stage in range(6):
            sim.run_stage(stage)
            ref.run_stage(stage)
            assert_same(sim, ref, ("grid", "particles", "posvel"), ctx=f"stage {stage}")
        for frame in range(2):
            sim.execute()
            ref.execute(2)
            assert_same(sim, ref, ("particles", "posvel"), ctx=f"frame {frame}")


def test_empty_scene_and_reset():
    box = [8.0, 8.0, 8.0]
    with MLSMPMSimulator(init_box_size=box) as sim:
        assert sim.num_particles == 0 and sim.grid_count == 512
        sim.execute()
        sim.synchronize()
        rec = scenes.small_scene(300, [16.0, 16.0, 16.0], seed=9)
        sim.reset(300, [16.0, 16.0, 16.0], particles=rec)
        ref = oracle.Scene(rec, [16.0, 16.0, 16.0])
        sim.execute()
        ref.execute(2)
        assert_same(sim, ref, ("particles", "posvel"), ctx="after reset")
        assert sim.grid_count == 4096


def test_posvel_device_pointer_is_stable_and_holds_posvel():
    import ctypes

    box = [24.0, 24.0, 24.0]
    rec = scenes.small_scene(2000, box, seed=11, velocity=1.0, affine=0.3)
    sim, ref = pair(rec, box)
    with sim:
        hip = ctypes.CDLL("libamdhip64.so.7")  # the HIP runtime libmpm355 already loaded
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        hip.hipMemcpy.restype = ctypes.c_int
        p0 = sim.posvel_device_ptr()
        assert p0
        for f in range(3):
            if f == 1:  # a box change keeps the buffer
                sim.change_box_size([20.0, 24.0, 24.0])
                ref.change_box_size([20.0, 24.0, 24.0])
            sim.execute()
            ref.execute(2)
            sim.synchronize()
            assert sim.posvel_device_ptr() == p0
            out = np.zeros((2000, 8), dtype=np.float32)
            assert hip.hipMemcpy(out.ctypes.data, p0, out.nbytes, 2) == 0  # device to host
            assert np.all(bits_equal(out, ref.posvel)), f"frame {f}"


def test_errors_are_status_codes():
    with MLSMPMSimulator(init_box_size=(8, 8, 8)) as sim:
        with pytest.raises(MPMError) as ei:
            sim.reset(0, [2000.0, 2000.0, 2000.0], particles=np.zeros((0, 20), np.float32))
        assert ei.value.status == 2            # grid too large
        with pytest.raises(MPMError):
            sim.change_box_size([float("nan"), 8, 8])
        with pytest.raises(MPMError):
            sim.run_stage(7)
        # the handle stays usable after failed calls
        sim.reset(10, [8.0, 8.0, 8.0], particles=scenes.small_scene(10, [8.0, 8.0, 8.0], 1))
        sim.execute()
        assert sim.read_posvel().shape == (10, 8)


@pytest.mark.parametrize("fail_after", [0, 7, 21])
def test_failed_reset_leaves_a_handle_that_refuses_work(monkeypatch, fail_after):
    """A reset whose device allocation fails (forced from the host with the
    library's MPM355_FAIL_ALLOC_AFTER knob) releases everything: every later
    call reports MPM_ERR_NOT_READY instead of launching on missing buffers,
    until a reset succeeds."""
    box, big_box = [16.0, 16.0, 16.0], [32.0, 32.0, 32.0]
    rec = scenes.small_scene(500, box, seed=4)
    big = scenes.small_scene(5000, big_box, seed=5)   # larger buffers: reallocation
    with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
        sim.execute()
        monkeypatch.setenv("MPM355_FAIL_ALLOC_AFTER", str(fail_after))
        with pytest.raises(MPMError) as ei:
            sim.reset(5000, big_box, particles=big)
        assert ei.value.status == 3   # MPM_ERR_OUT_OF_MEMORY
        monkeypatch.delenv("MPM355_FAIL_ALLOC_AFTER")
        calls = [lambda: sim.execute(), lambda: sim.read_posvel(), lambda: sim.read_particles(),
                 lambda: sim.read_grid(), lambda: sim.change_box_size([8.0, 8.0, 8.0]),
                 lambda: sim.run_stage(0), lambda: sim.backward(2, [1.0, 1.0, 1.0]),
                 lambda: sim.posvel_device_ptr(), lambda: sim.time_stages(1),
                 lambda: sim.num_particles]
        for k, call in enumerate(calls):
            with pytest.raises(MPMError) as ei:
                call()
            assert ei.value.status == 6, f"call {k}"   # MPM_ERR_NOT_READY
        sim.synchronize()
        sim.reset(5000, big_box, particles=big)
        ref = oracle.Scene(big, big_box)
        sim.execute()
        ref.execute(2)
        assert_same(sim, ref, ("particles", "posvel"), ctx="after the failed reset")


TUNINGS = [
    dict(path=0, particles_per_thread=1, grid_blocks=0, xcd_remap=1),
    dict(path=0, particles_per_thread=2, grid_blocks=24, xcd_remap=1),
    dict(path=0, particles_per_thread=4, grid_blocks=3, xcd_remap=0),
    dict(path=0, particles_per_thread=1, grid_blocks=4096, xcd_remap=0),
    dict(path=1, particles_per_thread=2, grid_blocks=0, xcd_remap=1),
    dict(path=0, particles_per_thread=1, grid_blocks=0, xcd_remap=1, tile_edge=8),
    dict(path=0, particles_per_thread=2, grid_blocks=24, xcd_remap=0, tile_edge=8),
    dict(path=0, particles_per_thread=4, grid_blocks=3, xcd_remap=1, tile_edge=8),
]
# integral init box: tame particles take the tile-partial path; a non-integral
# one sends every particle down the global-atomic path
BOXES = [[40.0, 36.0, 44.0], [40.5, 36.0, 43.25]]


@pytest.mark.parametrize("box", BOXES, ids=["integral", "fractional"])
@pytest.mark.parametrize("tun", TUNINGS, ids=lambda t: "-".join(str(v) for v in t.values()))
def test_tunings_do_not_change_results(tun, box):
    rec = scenes.small_scene(20000, box, seed=21, velocity=1.5, affine=0.4)
    sim, ref = pair(rec, box)
    with sim:
        sim.set_tuning(**tun)
        for frame in range(4):
            if frame == 2:
                sim.change_box_size([33.5, 36.0, 41.0])
                ref.change_box_size([33.5, 36.0, 41.0])
            sim.execute()
            ref.execute(2)
            assert_same(sim, ref, ("particles", "posvel"), ctx=f"{tun} frame {frame}")
        # the per-stage hooks still work on the binned storage order
        for stage in range(6):
            sim.run_stage(stage)
            ref.run_stage(stage)
        assert_same(sim, ref, ("grid", "particles", "posvel"), ctx=f"{tun} stages")


def test_time_stages_reports_every_kernel_slot():
    rec, box = scenes.config0()
    with MLSMPMSimulator(particles=rec, init_box_size=box) as sim:
        t = sim.time_stages(8)
        assert set(t) == {"bin_gather", "bin_scan", "bin_scatter", "p2g_1", "reduce_1", "p2g_2",
                          "reduce_2", "g2p"}
        assert all(ms > 0 for ms, calls in t.values())
        assert t["p2g_1"][1] == 8
        ms = sim.time_steps(4)
        assert ms > 0


def test_switching_the_tile_edge_mid_rollout_keeps_results():
    """4^3 -> 8^3 -> 4^3 tiles between frames (the autotuner's switch): the
    key geometry and tile tables are rebuilt under the running state, and
    every frame still equals the oracle."""
    rec, box = scenes.config0()
    sim, ref = pair(rec, box)
    with sim:
        for frame, edge in enumerate((4, 8, 8, 4, 8)):
            sim.set_tuning(tile_edge=edge)
            assert sim.tuning["tile_edge"] == edge
            sim.execute()
            ref.execute(2)
            assert_same(sim, ref, ("particles", "posvel"), ctx=f"frame {frame} edge {edge}")


def test_autotune_restores_the_state_and_keeps_results():
    """mpm_autotune times graph-replayed candidates (both tile edges) on a
    snapshot and restores it: the simulation then continues exactly as the
    oracle's from the same state."""
    rec, box = scenes.config0()
    sim, ref = pair(rec, box)
    with sim:
        sim.execute()
        ref.execute(2)
        chosen = sim.autotune(3)
        assert chosen["tile_edge"] in (4, 8) and chosen["particles_per_thread"] in (1, 2, 4)
        assert_same(sim, ref, ("particles", "posvel"), ctx="after autotune")
        sim.execute()
        ref.execute(2)
        assert_same(sim, ref, ("particles", "posvel"), ctx="frame after autotune")
