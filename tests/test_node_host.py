"""The Node host: N-API addon + JS MLSMPMSimulator (the reference's host
language).  CPU tests check the addon's exports, the JS class surface and the
JS initDambreak; the GPU test drives headless.js (the reference frame loop of
main.ts:249-277) and compares its PosVel output with the oracle."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from mlsmpm import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE_DIR = os.path.join(ROOT, "webgpu-ocean_amd", "node")
ADDON = os.path.join(NODE_DIR, "mpm.node")

pytestmark = pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")


def node(script: str, *args, timeout=120):
    return subprocess.run(["node", "-e", script, *args], capture_output=True, text=True,
                          timeout=timeout, cwd=NODE_DIR)


def test_addon_exports_and_reference_method_names():
    assert os.path.exists(ADDON), "build mpm.node first (__graft_entry__.build())"
    r = node("""
const m = require('./mls-mpm.js');
const p = m.MLSMPMSimulator.prototype;
console.log(JSON.stringify({
  addon: Object.keys(m.addon).sort(),
  methods: Object.getOwnPropertyNames(p).sort(),
  size: m.mlsmpmParticleStructSize,
}));""")
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["size"] == 80
    for name in ["initDambreak", "reset", "execute", "changeBoxSize"]:
        assert name in out["methods"]
    assert {"create", "reset", "setBox", "step", "readPosvel"} <= set(out["addon"])


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_constructor_throws_without_device():
    r = node("""
const m = require('./mls-mpm.js');
try { new m.MLSMPMSimulator(null, null, 1.2, {}); console.log('no-throw'); }
catch (e) { console.log(e.message); }""")
    assert "mpm error 5" in r.stdout


def test_js_init_dambreak_matches_python_mirror():
    # Same deterministic random sequence on both sides.
    r = node("""
const m = require('./mls-mpm.js');
let k = 0;
const self = { numParticles: 0, random: () => ((k++ % 97) / 97) };
const ab = m.MLSMPMSimulator.prototype.initDambreak.call(self, [40, 30, 60], 5000);
process.stdout.write(Buffer.from(ab).toString('base64'));""")
    assert r.returncode == 0, r.stderr
    import base64
    js = np.frombuffer(base64.b64decode(r.stdout), dtype=np.float32).reshape(-1, 20)
    seq = iter([(k % 97) / 97 for k in range(100000)])
    py = scenes.init_dambreak([40, 30, 60], 5000, rand=lambda: next(seq))
    assert js.shape == py.shape == (5000, 20)
    assert np.array_equal(js, py)


@pytest.mark.gpu
@pytest.mark.parametrize("slider", [100, 40])
def test_headless_frame_loop_matches_oracle(tmp_path, slider):
    box = [36.0, 30.0, 40.0]
    rec = scenes.small_scene(6000, box, seed=31, velocity=0.5, affine=0.1)
    pfile, vfile = tmp_path / "p.bin", tmp_path / "pv.bin"
    rec.tofile(pfile)
    frames = 6
    r = subprocess.run(["node", os.path.join(NODE_DIR, "headless.js"), "--particles", str(pfile),
                        "--box", ",".join(str(b) for b in box), "--frames", str(frames),
                        "--slider", str(slider), "--posvel", str(vfile)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout)
    assert info["numParticles"] == 6000
    ref = oracle.Scene(rec, box)
    ratio = 1.0
    real = list(box)
    for _ in range(frames):
        ratio += max(slider / 200 + 0.5 - ratio, -0.007)
        real[2] = box[2] * ratio
        ref.change_box_size(real)
        ref.execute(2)
    pv = np.fromfile(vfile, dtype=np.float32).reshape(-1, 8)
    same = (pv.view(np.int32) == ref.posvel.view(np.int32)) | (np.isnan(pv) & np.isnan(ref.posvel))
    assert np.all(same)


# ---- SPH (sph.js over the same addon) -------------------------------------

def test_sph_class_surface_and_addon_exports():
    r = node("""
const m = require('./sph.js');
const p = m.SPHSimulator.prototype;
console.log(JSON.stringify({
  addon: Object.keys(m.addon).filter(k => k.startsWith('sph')).sort(),
  methods: Object.getOwnPropertyNames(p).sort(),
  size: m.sphParticleStructSize,
}));""")
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["size"] == 64
    for name in ["initDambreak", "reset", "execute", "changeBoxSize"]:
        assert name in out["methods"]
    assert {"sphCreate", "sphReset", "sphSetBox", "sphStep", "sphReadPosvel"} <= set(out["addon"])


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_sph_constructor_throws_without_device():
    r = node("""
const m = require('./sph.js');
try { new m.SPHSimulator(null, null, 0.08, {}); console.log('no-throw'); }
catch (e) { console.log(e.message); }""")
    assert "sph error 5" in r.stdout


def test_sph_js_init_dambreak_matches_python_mirror():
    import sph
    r = node("""
const m = require('./sph.js');
let k = 0;
const self = { numParticles: 0, kernelRadius: 0.07, random: () => ((k++ % 89) / 89) };
const ab = m.SPHSimulator.prototype.initDambreak.call(self, [1.2, 2.0, 1.2], 7000);
process.stdout.write(Buffer.from(ab).toString('base64'));""")
    assert r.returncode == 0, r.stderr
    import base64
    js = np.frombuffer(base64.b64decode(r.stdout), dtype=np.float32).reshape(-1, 16)
    seq = iter([(k % 89) / 89 for k in range(100000)])
    py = sph.scenes.init_dambreak([1.2, 2.0, 1.2], 7000, rand=lambda: next(seq))
    assert js.shape == py.shape == (7000, 16)
    assert np.array_equal(js, py)


@pytest.mark.gpu
@pytest.mark.parametrize("slider", [100, 30])
def test_sph_headless_frame_loop_matches_oracle(tmp_path, slider):
    import sph
    from oracle import sph as osph
    half = [1.0, 2.0, 1.0]
    rec = sph.scenes.dam_break(1, seed=5, num_particles=8000)
    pfile, vfile = tmp_path / "p.bin", tmp_path / "pv.bin"
    rec.tofile(pfile)
    frames = 8
    r = subprocess.run(["node", os.path.join(NODE_DIR, "headless.js"), "--solver", "sph",
                        "--particles", str(pfile), "--box", ",".join(str(b) for b in half),
                        "--frames", str(frames), "--slider", str(slider), "--posvel", str(vfile)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout)
    assert info["numParticles"] == 8000 and info["solver"] == "sph"
    ref = osph.Scene(rec, half)
    ratio = 1.0
    real = list(half)
    for _ in range(frames):
        ratio += max(slider / 200 + 0.5 - ratio, -0.015)   # main.ts:253
        real[2] = half[2] * ratio
        ref.change_box_size(real)
        ref.execute(2, threads=8)
    pv = np.fromfile(vfile, dtype=np.float32).reshape(-1, 8)
    cols = [0, 1, 2, 4, 5, 6]
    assert np.array_equal(pv[:, cols].view(np.int32), ref.posvel[:, cols].view(np.int32))
