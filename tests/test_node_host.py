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


# ---- the renderer on Node (fluid-render.js, camera.js) ----

def test_renderer_and_camera_surface():
    r = node("""
const f = require('./fluid-render.js');
const c = require('./camera.js');
console.log(JSON.stringify({
  addon: Object.keys(f.addon).filter(k => k.startsWith('render')).sort(),
  renderer: Object.getOwnPropertyNames(f.FluidRenderer.prototype).sort(),
  camera: Object.getOwnPropertyNames(c.Camera.prototype).sort(),
  bytes: c.renderUniformsValues.byteLength,
}));""")
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["addon"] == ["renderColorDevicePtr", "renderCreate", "renderDestroy",
                            "renderExecute", "renderReadImage", "renderTimeFrames"]
    assert {"execute", "readColor", "readImage", "destroy"} <= set(out["renderer"])
    assert {"reset", "recalculateView", "drag", "wheel"} <= set(out["camera"])
    assert out["bytes"] == 272


def test_js_camera_uniforms_match_the_python_camera():
    """camera.js and render/camera.py restate the same camera.ts and
    wgpu-matrix arithmetic: the 68 uniform floats agree to f32 rounding after
    a reset, two drags and two wheel events (JS drag(dx, dy) is a mouse move
    of (-dx, -dy))."""
    import math

    from render import camera
    r = node("""
const c = require('./camera.js');
const cam = new c.Camera(null);
const out = [];
cam.reset({width: 400, height: 240}, 70, [20, 7.5, 30], 45 * Math.PI / 180, 1.5);
c.renderUniformsViews.texel_size.set([1 / 400, 1 / 240]);
c.renderUniformsViews.sphere_size.set([1.2]);
out.push(Array.from(new Float32Array(c.renderUniformsValues)));
cam.drag(37, -12); out.push(Array.from(new Float32Array(c.renderUniformsValues)));
cam.drag(-400, 900); out.push(Array.from(new Float32Array(c.renderUniformsValues)));
cam.wheel(-3); cam.wheel(-3); out.push(Array.from(new Float32Array(c.renderUniformsValues)));
console.log(JSON.stringify(out));""")
    assert r.returncode == 0, r.stderr
    js = np.array(json.loads(r.stdout), dtype=np.float32)
    cam = camera.Camera()
    cam.reset(400, 240, 70, [20, 7.5, 30], 45 * math.pi / 180, 1.5)
    py = [cam.uniforms(400, 240, 1.2)]
    cam.orbit(-37, 12)
    py.append(cam.uniforms(400, 240, 1.2))
    cam.orbit(400, -900)
    py.append(cam.uniforms(400, 240, 1.2))
    cam.zoom(-3)
    cam.zoom(-3)
    py.append(cam.uniforms(400, 240, 1.2))
    np.testing.assert_allclose(js, np.array(py), rtol=1e-6, atol=1e-6)
    assert cam.y_theta == 0.0   # the second drag clamps at the horizon
    assert cam.distance == pytest.approx(67.0)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_renderer_throws_without_device():
    r = node("""
const f = require('./fluid-render.js');
const c = require('./camera.js');
try { new f.FluidRenderer(null, {width: 64, height: 48}, null, 0.6, 0.78, null,
                          c.renderUniformsValues, {size: 2, data: new Float32Array(96)});
      console.log('no-throw'); }
catch (e) { console.log(e.message); }""")
    assert "render error 5" in r.stdout


@pytest.mark.gpu
def test_node_frame_loop_renders_like_the_oracle(tmp_path):
    """main.ts's frame loop on Node: MLSMPMSimulator steps, the camera fills
    the uniform buffer, FluidRenderer renders from the simulator's device
    buffer (fluid, then spheres); the depth and thickness maps equal the
    renderer oracle's on the same PosVel and uniforms, bit for bit."""
    import math

    from oracle import render as orender
    from render import camera
    cube = camera.synthetic_cubemap(16)
    cube.tofile(str(tmp_path / "cube.f32"))
    w, h = 240, 160
    r = node("""
const fs = require('fs');
const m = require('./mls-mpm.js');
const f = require('./fluid-render.js');
const c = require('./camera.js');
const dir = process.argv[1];
let k = 0;
const sim = new m.MLSMPMSimulator(null, null, 1.2, {random: () => ((k++ % 89) / 89)});
sim.reset(40000, [35, 25, 55]);
for (let i = 0; i < 10; ++i) sim.execute(null);
const cam = new c.Camera(null);
cam.reset({width: WIDTH, height: HEIGHT}, 60, [35 / 2, 25 / 4, 55 / 2], 45 * Math.PI / 180, 1.5);
c.renderUniformsViews.texel_size.set([1 / WIDTH, 1 / HEIGHT]);
c.renderUniformsViews.sphere_size.set([1.2]);
const cubeData = new Float32Array(fs.readFileSync(dir + '/cube.f32').buffer.slice(0));
const ren = new f.FluidRenderer(null, {width: WIDTH, height: HEIGHT}, null, 0.6, 45 * Math.PI / 180,
                                sim, c.renderUniformsValues, {size: 16, data: cubeData});
ren.execute(null, null, sim.numParticles, false);
fs.writeFileSync(dir + '/depth.f32', Buffer.from(ren.readImage('depth').buffer));
fs.writeFileSync(dir + '/thick.f32', Buffer.from(ren.readImage('thickness').buffer));
fs.writeFileSync(dir + '/fdepth.f32', Buffer.from(ren.readImage('filteredDepth').buffer));
fs.writeFileSync(dir + '/fthick.f32', Buffer.from(ren.readImage('filteredThickness').buffer));
fs.writeFileSync(dir + '/fluid.f32', Buffer.from(ren.readColor().buffer));
ren.execute(null, null, sim.numParticles, true);
fs.writeFileSync(dir + '/sphere.f32', Buffer.from(ren.readColor().buffer));
fs.writeFileSync(dir + '/posvel.f32', Buffer.from(sim.readPositions()));
fs.writeFileSync(dir + '/uniforms.f32', Buffer.from(c.renderUniformsValues));
console.log(sim.numParticles);
ren.destroy();
sim.destroy();""".replace("WIDTH", str(w)).replace("HEIGHT", str(h)), str(tmp_path), timeout=300)
    assert r.returncode == 0, r.stderr
    n = int(r.stdout.strip())
    pv = np.fromfile(str(tmp_path / "posvel.f32"), dtype=np.float32).reshape(n, 8)
    u = np.fromfile(str(tmp_path / "uniforms.f32"), dtype=np.float32)
    c = orender.consts(w, h, 0.6, 45 * math.pi / 180)
    want = orender.render_frame(pv, n, u, c, cube, False, 16)
    depth = np.fromfile(str(tmp_path / "depth.f32"), dtype=np.float32).reshape(h, w)
    thick = np.fromfile(str(tmp_path / "thick.f32"), dtype=np.float32).reshape(h, w)
    assert np.count_nonzero(depth) > 0.05 * w * h
    np.testing.assert_array_equal(depth.view(np.uint32), want["depth"].view(np.uint32))
    np.testing.assert_array_equal(thick.view(np.uint32), want["thickness"].view(np.uint32))
    for name, key in (("fdepth", "filtered_depth"), ("fthick", "filtered_thickness"),
                      ("fluid", "color")):
        got = np.fromfile(str(tmp_path / (name + ".f32")), dtype=np.float32)
        np.testing.assert_array_equal(got.view(np.uint32), want[key].reshape(-1).view(np.uint32))
    sph = orender.render_frame(pv, n, u, c, cube, True, 16)
    got = np.fromfile(str(tmp_path / "sphere.f32"), dtype=np.float32).reshape(h, w, 4)
    np.testing.assert_array_equal(got.view(np.uint32), sph["color"].view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["mlsmpm", "sph"])
def test_headless_frame_loop_renders_every_frame(tmp_path, solver):
    """headless.js --render: main.ts's frame loop with the renderer's pass over
    the simulator's buffer each frame (fluid, then spheres); the last frame
    shows the fluid over the 0.8 grey background, every value in [0, 1]."""
    w, h = 200, 120
    for spheres in (0, 1):
        out = tmp_path / f"color{spheres}.f32"
        r = subprocess.run(["node", os.path.join(NODE_DIR, "headless.js"), "--solver", solver,
                            "--preset", "0", "--frames", "8", "--render", f"{w}x{h}",
                            "--spheres", str(spheres), "--color", str(out)],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert json.loads(r.stdout)["rendered"] == [w, h]
        img = np.fromfile(str(out), dtype=np.float32).reshape(h, w, 4)
        assert np.isfinite(img).all() and img.min() >= 0.0 and img.max() <= 1.0
        fluid = np.any(img[..., :3] != np.float32(0.8), axis=-1)
        assert fluid.mean() > 0.02, fluid.mean()
