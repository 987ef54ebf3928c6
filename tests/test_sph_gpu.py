"""GPU parity of the SPH path (libmpm355's sph_* entry points) against the
CPU oracle (oracle/sph_oracle.c), bit for bit: the neighbour grid (prefix
sums and sorted order), every particle field after one substep for each of
the reference's four settings (main.ts:131-132), a rollout with a box change,
and the edge cases (no particles, an isolated particle, particles outside
the grid, NaN and infinite positions)."""
import numpy as np
import pytest

from oracle import sph as osph

pytestmark = pytest.mark.gpu

FIELDS = [0, 1, 2, 4, 5, 6, 8, 9, 10, 11, 12]   # position, v, force, density, nearDensity
PV = [0, 1, 2, 4, 5, 6]


@pytest.fixture(scope="module")
def sphmod():
    import sph
    return sph


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _same(got, want, what):
    g, w = _bits(got), _bits(want)
    bad = np.nonzero(g != w)
    assert bad[0].size == 0, (f"{what}: {bad[0].size} words differ, first at {bad[0][0]}: "
                              f"{got.ravel()[np.ravel_multi_index(bad, got.shape)[0]]} vs "
                              f"{want.ravel()[np.ravel_multi_index(bad, want.shape)[0]]}")


def _compare(sim, sc, what):
    _same(sim.read_particles()[:, FIELDS], sc.particles[:, FIELDS], what + " particles")
    _same(sim.read_posvel()[:, PV], sc.posvel[:, PV], what + " posvel")


def test_neighbour_grid_equals_the_oracle(sphmod):
    p = sphmod.scenes.dam_break(1, seed=2)
    half = sphmod.scenes.INIT_HALF_BOXES[1]
    with sphmod.SPHSimulator(particles=p, init_half_box_size=half) as sim:
        for stage in ("count", "scan_reduce", "scan", "scatter", "sort_gather"):
            sim.run_stage(stage)
        start, order, _ = osph.Scene(p, half).bin()
        np.testing.assert_array_equal(sim.read_cell_start(), start)
        np.testing.assert_array_equal(sim.read_sorted_order(len(order)), order)


@pytest.mark.parametrize("setting", [0, 1, 2, 3])
def test_one_substep_is_bit_exact_for_each_reference_setting(sphmod, setting):
    p = sphmod.scenes.dam_break(setting, seed=10 + setting)
    half = sphmod.scenes.INIT_HALF_BOXES[setting]
    with sphmod.SPHSimulator(particles=p, init_half_box_size=half) as sim:
        assert sim.num_particles == sphmod.scenes.NUM_PARTICLES[setting]
        assert sim.grid_count == 62 ** 3
        sim.execute(1)
        sc = osph.Scene(p, half)
        sc.substep(threads=8)
        _compare(sim, sc, f"setting {setting}")


def test_stage_by_stage_equals_the_graph_replayed_substep(sphmod):
    p = sphmod.scenes.dam_break(0, seed=4)
    half = sphmod.scenes.INIT_HALF_BOXES[0]
    with sphmod.SPHSimulator(particles=p, init_half_box_size=half) as a, \
            sphmod.SPHSimulator(particles=p, init_half_box_size=half) as b:
        for _ in range(3):
            for stage in sphmod.STAGES:
                a.run_stage(stage)
            b.execute(1)
        _same(a.read_particles(), b.read_particles(), "stages vs graph")
        _same(a.read_posvel(), b.read_posvel(), "stages vs graph posvel")


def test_rollout_with_a_box_change_stays_bit_exact(sphmod):
    # main.ts:253-260 narrows the z half extent with the slider.
    p = sphmod.scenes.dam_break(1, seed=7)
    half = list(sphmod.scenes.INIT_HALF_BOXES[1])
    sc = osph.Scene(p, half)
    with sphmod.SPHSimulator(particles=p, init_half_box_size=half) as sim:
        for frame in range(40):
            if frame == 20:
                narrow = [half[0], half[1], half[2] * 0.7]
                sim.change_box_size(narrow)
                sc.change_box_size(narrow)
            sim.execute()
            sc.execute(2, threads=8)
            if frame % 10 == 9:
                _compare(sim, sc, f"frame {frame}")
        ke = 0.5 * (sim.read_posvel()[:, 4:7].astype(np.float64) ** 2).sum()
        assert np.isfinite(ke) and ke > 0


def test_edge_cases_outside_the_grid_nan_inf_isolated(sphmod):
    rec = np.zeros((8, 16), dtype=np.float32)
    rec[0, :3] = (0.3, 0.5, -0.2)                     # isolated particle
    rec[1, :3] = (-3.0, -3.0, -3.0)                   # negative linear id: not binned
    rec[2, :3] = (2.12, 0.0, 0.0)                     # last cell of the x row
    rec[3, :3] = (-2.2, -2.0, -2.0)                   # x coordinate -1, wraps to a valid id
    rec[4, :3] = (np.nan, 0.0, 0.0)
    rec[5, :3] = (np.inf, -np.inf, 0.0)
    rec[6, :3] = (0.3 + 0.01, 0.5, -0.2)              # neighbour of particle 0
    rec[7, :3] = (9.0, 9.0, 9.0)                      # beyond the last cell
    rec[:, 4:7] = np.float32([0.1, -0.2, 0.3])
    half = (1.0, 2.0, 1.0)
    with sphmod.SPHSimulator(particles=rec, init_half_box_size=half) as sim:
        sc = osph.Scene(rec, half)
        for _ in range(3):
            start, order, _ = sc.bin()        # the binning this substep starts from
            sim.execute(1)
            sc.substep()
            got, want = sim.read_particles(), sc.particles
            np.testing.assert_array_equal(np.isnan(got[:, FIELDS]), np.isnan(want[:, FIELDS]))
            _same(np.nan_to_num(got[:, FIELDS]), np.nan_to_num(want[:, FIELDS]), "edge")
            np.testing.assert_array_equal(sim.read_cell_start(), start)
            np.testing.assert_array_equal(sim.read_sorted_order(len(order)), order)


def test_no_particles_and_reset(sphmod):
    with sphmod.SPHSimulator(particles=None, init_half_box_size=(1, 2, 1)) as sim:
        assert sim.num_particles == 0
        sim.execute()
        sim.synchronize()
        assert sim.read_posvel().shape == (0, 8)
        p = sphmod.scenes.dam_break(0, seed=1, num_particles=3000)
        sim.reset(3000, (0.7, 2.0, 0.7), particles=p)
        assert sim.num_particles == 3000
        sim.execute()
        sc = osph.Scene(p, (0.7, 2.0, 0.7))
        sc.execute(2)
        _compare(sim, sc, "after reset")


def test_write_particles_round_trip_and_argument_errors(sphmod):
    p = sphmod.scenes.dam_break(0, seed=9, num_particles=2000)
    p[:, 8:13] = np.arange(2000 * 5, dtype=np.float32).reshape(2000, 5)
    with sphmod.SPHSimulator(particles=p, init_half_box_size=(0.7, 2, 0.7)) as sim:
        _same(sim.read_particles(), p, "round trip")
        q = p.copy()
        q[:, 0] += np.float32(0.01)
        sim.write_particles(q)
        _same(sim.read_particles(), q, "write")
        with pytest.raises(sphmod.SPHError):
            sim.write_particles(q[:10])
        with pytest.raises(sphmod.SPHError):
            sim.change_box_size((np.nan, 1, 1))
        with pytest.raises(sphmod.SPHError):
            sim.run_stage(7)
    bad = sphmod.default_params()
    bad.kernel_radius = 0.0
    with pytest.raises(sphmod.SPHError):
        sphmod.SPHSimulator(params=bad)
    bad.kernel_radius = 1e-3
    with pytest.raises(sphmod.SPHError) as ei:
        sphmod.SPHSimulator(params=bad)
    assert ei.value.status == 2


def test_timing_hooks_cover_every_stage(sphmod):
    p = sphmod.scenes.dam_break(3, seed=1)
    with sphmod.SPHSimulator(particles=p, init_half_box_size=sphmod.scenes.INIT_HALF_BOXES[3]) as sim:
        st = sim.time_stages(4)
        assert set(st) == set(sphmod.STAGES) and all(v > 0 for v in st.values())
        assert sim.time_steps(10) > 0


def test_particle_cap_200k_two_frames_bit_exact(sphmod):
    # numParticlesMax (common.ts:11) in the largest setting's box: long
    # neighbour lists (the force walk's list drains mid-stream)
    half = sphmod.scenes.INIT_HALF_BOXES[3]
    p = sphmod.scenes.dam_break(3, seed=21, num_particles=200000)
    with sphmod.SPHSimulator(particles=p, init_half_box_size=half) as sim:
        sc = osph.Scene(p, half)
        for frame in range(2):
            sim.execute()
            sc.execute(2, threads=16)
        _compare(sim, sc, "200k")
