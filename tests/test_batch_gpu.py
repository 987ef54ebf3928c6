"""Batched scenes: one launch per stage over many uneven scenes.  Each
scene's results must equal the oracle's (and a single simulator's) on the
same inputs, bit for bit, whatever the tuning the autotuner picks."""
import numpy as np
import pytest

import oracle
from mlsmpm import MLSMPMBatch, MLSMPMSimulator, Params, scenes

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return (a.view(np.int32) == b.view(np.int32)) | (np.isnan(a) & np.isnan(b))


def params_of(phys):
    p = Params(3.0, 4.0, 0.1, 0.2, 1e7)
    for k, v in phys.items():
        setattr(p, k, v)
    return p


def small_batch(seed=5, count=7):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(count):
        edge = float(rng.choice([12, 16, 20, 28]))
        box = [edge, float(rng.choice([12, 16, 24])), edge + 2]
        n = int(rng.integers(1, 3000)) if s else 0   # scene 0 is empty
        n = min(n, scenes.small_scene_capacity(box))
        rec = scenes.small_scene(max(n, 1), box, seed=seed + s, velocity=1.0, affine=0.3)[:n]
        phys = {"stiffness": float(rng.uniform(1, 5)),
                "dynamic_viscosity": float(rng.uniform(0.02, 0.2))}
        out.append((rec, box, phys))
    return out


def oracle_scene(rec, box, phys):
    return oracle.Scene(rec, box, oracle.default_params(**phys))


@pytest.mark.parametrize("tuning", [None, dict(particles_per_thread=1, grid_blocks=5, xcd_remap=0),
                                    dict(particles_per_thread=4, grid_blocks=64),
                                    dict(particles_per_thread=2, grid_blocks=40, tile_edge=8)])
def test_batch_matches_oracle_per_scene(tuning):
    sc = small_batch()
    refs = [oracle_scene(r, b, p) for r, b, p in sc]
    with MLSMPMBatch([(r, b, params_of(p)) for r, b, p in sc]) as bt:
        if tuning:
            bt.set_tuning(**tuning)
        for frame in range(3):
            if frame == 1:
                bt.change_box_size(3, [10.5, 12.0, 13.0])
                refs[3].change_box_size([10.5, 12.0, 13.0])
            bt.step(2)
            for r in refs:
                r.execute(2)
        for s, r in enumerate(refs):
            assert np.all(bits_equal(bt.read_posvel(s), r.posvel)), f"scene {s} posvel"
            assert np.all(bits_equal(bt.read_particles(s), r.particles)), f"scene {s} state"


def test_autotuned_batch_equals_scenes_run_alone():
    sc = scenes.config3(seed=3, scenes=12, n_lo=2000, n_hi=40000)
    with MLSMPMBatch([(r, b, params_of(p)) for r, b, p in sc]) as bt:
        chosen = bt.autotune(2)
        assert chosen["particles_per_thread"] in (1, 2, 4)
        assert chosen["tile_edge"] in (4, 8)
        bt.step(4)
        for s, (r, b, p) in enumerate(sc):
            with MLSMPMSimulator(params=params_of(p), particles=r, init_box_size=b) as sim:
                sim.execute(4)
                assert np.all(bits_equal(bt.read_posvel(s), sim.read_posvel())), f"scene {s}"


def test_config3_full_batch_equals_every_scene_alone_and_the_oracle():
    """The 64-scene mix (about 3.4M particles, uneven counts, boxes and
    parameters) in one batch: after 4 substeps every scene equals a single
    simulator run alone on it, and 8 scenes spread over the mix (the
    smallest and largest included) equal the oracle, bit for bit."""
    sc = scenes.config3()
    assert len(sc) == 64
    steps = 4
    with MLSMPMBatch([(r, b, params_of(p)) for r, b, p in sc]) as bt:
        bt.step(steps)
        for s, (r, b, p) in enumerate(sc):
            with MLSMPMSimulator(params=params_of(p), particles=r, init_box_size=b) as sim:
                sim.execute(steps)
                assert np.all(bits_equal(bt.read_posvel(s), sim.read_posvel())), f"scene {s}"
                assert np.all(bits_equal(bt.read_particles(s), sim.read_particles())), \
                    f"scene {s} state"
        sizes = [r.shape[0] for r, _, _ in sc]
        picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 5, 13, 27, 38, 50, 61})
        assert len(picks) >= 8
        for s in picks:
            r, b, p = sc[s]
            ref = oracle_scene(r, b, p)
            ref.execute(steps, threads=8)
            assert np.all(bits_equal(bt.read_posvel(s), ref.posvel)), f"scene {s} vs oracle"
            assert np.all(bits_equal(bt.read_particles(s), ref.particles)), f"scene {s} state"


def test_batch_errors():
    from mlsmpm import MPMError
    with pytest.raises(MPMError):
        MLSMPMBatch([(np.zeros((0, 20), np.float32), [5000.0, 5000.0, 5000.0], None)])
    sc = small_batch(count=2)
    with MLSMPMBatch([(r, b, params_of(p)) for r, b, p in sc]) as bt:
        with pytest.raises(MPMError):
            bt.read_posvel(5)
        with pytest.raises(MPMError):
            bt.set_tuning(path=1)        # reference-shaped path is single-scene only
        bt.step(1)


def test_large_key_space_scan_matches_oracle():
    """Six sparse 96^3 scenes: more than 65536 bin keys, so the binning scan
    takes its multi-block form."""
    sc = []
    for s in range(6):
        box = [96.0, 96.0, 96.0]
        rec = scenes.small_scene(1500 + 100 * s, box, seed=40 + s, velocity=1.0, affine=0.3)
        sc.append((rec, box, {}))
    refs = [oracle_scene(r, b, p) for r, b, p in sc]
    with MLSMPMBatch([(r, b, params_of(p)) for r, b, p in sc]) as bt:
        bt.step(2)
        bt.step(2)
        for r in refs:
            r.execute(4)
        for s, r in enumerate(refs):
            assert np.all(bits_equal(bt.read_posvel(s), r.posvel)), f"scene {s} posvel"
            assert np.all(bits_equal(bt.read_particles(s), r.particles)), f"scene {s} state"


def test_batched_runs_while_the_host_is_busy_are_identical():
    """Regression companion of the single-scene host-load test: 16 uneven
    scenes of the 64-scene mix, stepped in 20-substep calls while the oracle
    keeps every host core busy, give the same particles on every run."""
    import threading

    mix = scenes.config3()[:16]
    rec1, box1 = scenes.config1()
    burner = oracle.Scene(rec1, box1)
    stop = threading.Event()

    def burn():
        while not stop.is_set():
            burner.substep(threads=16)

    worker = threading.Thread(target=burn)
    worker.start()
    runs = []
    try:
        for _ in range(3):
            with MLSMPMBatch([(r, b, params_of(p)) for r, b, p in mix]) as bt:
                for _ in range(5):
                    bt.step(20)
                    bt.synchronize()
                runs.append([bt.read_particles(s) for s in range(len(mix))])
    finally:
        stop.set()
        worker.join()
    for k, run in enumerate(runs[1:], 1):
        for s, (a, b) in enumerate(zip(run, runs[0])):
            assert np.all(bits_equal(a, b)), f"run {k} scene {s} differs from run 0"
