'use strict';
// headless.js -- the reference's frame loop without a browser (main.ts:188-283
// minus the UI), for either simulator: preset particle counts and boxes
// (main.ts:128-133), the box-width ratio closing at most 0.007 (MLS-MPM) or
// 0.015 (SPH) per frame toward the slider target (main.ts:249-260),
// changeBoxSize then execute, and with --render the renderer's pass over the
// simulator's buffer every frame (main.ts:264-275, camera as reset by the
// preset buttons, main.ts:229-241).
//
//   node headless.js [--solver mlsmpm|sph] [--preset 0..3] [--frames N]
//                    [--slider 0..100] [--particles FILE --box X,Y,Z]
//                    [--posvel OUT] [--render WxH] [--spheres 0|1] [--color OUT]
// --particles loads n x 80-B (MLS-MPM) or n x 64-B (SPH) records instead of
// initDambreak (--box: the init box, or SPH's init half box); --posvel writes
// the final n x 32-B PosVel records; --color the last frame's RGBA float32
// image (with a synthetic sky cube map).  Prints one JSON line.

const fs = require('fs');
const { MLSMPMSimulator } = require('./mls-mpm');
const { SPHSimulator } = require('./sph');
const { Camera, renderUniformsValues, renderUniformsViews } = require('./camera');
const { FluidRenderer } = require('./fluid-render');

const mlsmpmNumParticleParams = [40000, 70000, 120000, 200000];
const mlsmpmInitBoxSizes = [[35, 25, 55], [40, 30, 60], [45, 40, 80], [50, 50, 80]];
const sphNumParticleParams = [10000, 20000, 30000, 40000];
const sphInitBoxSizes = [[0.7, 2.0, 0.7], [1.0, 2.0, 1.0], [1.2, 2.0, 1.2], [1.4, 2.0, 1.4]];
const mlsmpmInitDistances = [60, 70, 90, 100];   // main.ts:130
const sphInitDistances = [2.6, 3.0, 3.4, 3.8];   // main.ts:133

// A smooth sky over a darker ground, one tint per face (the cube map's PNG
// faces are not decoded here).
function skyCubemap(size) {
    const data = new Float32Array(6 * size * size * 4);
    const tint = [[1.0, 0.9, 0.8], [0.8, 0.9, 1.0], [0.9, 0.95, 1.0], [0.4, 0.35, 0.3],
                  [0.95, 1.0, 0.9], [0.9, 0.9, 0.95]];
    for (let f = 0; f < 6; ++f)
        for (let r = 0; r < size; ++r)
            for (let c = 0; c < size; ++c) {
                const t = (r + 0.5) / size, s = (c + 0.5) / size;
                let sky = f === 2 ? 0.7 + 0.8 * s * (1 - s) : 0.45 + 0.5 * (1 - t);
                if (f === 3) sky = 0.25 + 0.1 * s;
                const o = 4 * ((f * size + r) * size + c);
                for (let k = 0; k < 3; ++k) data[o + k] = sky * tint[f][k];
                data[o + 3] = 1;
            }
    return { size, data };
}

function parseArgs(argv) {
    const a = { solver: 'mlsmpm', preset: 1, frames: 100, slider: 100, particles: null, box: null,
                posvel: null, render: null, spheres: 0, color: null };
    for (let i = 2; i < argv.length; i += 2) {
        const k = argv[i].replace(/^--/, ''), v = argv[i + 1];
        if (k === 'preset' || k === 'frames' || k === 'slider' || k === 'spheres') a[k] = Number(v);
        else if (k === 'render') a.render = v.split('x').map(Number);
        else if (k === 'box') a.box = v.split(',').map(Number);
        else a[k] = v;
    }
    return a;
}

function main() {
    const args = parseArgs(process.argv);
    const sphFl = args.solver === 'sph';
    const sim = sphFl ? new SPHSimulator(null, null, 0.08, {}) : new MLSMPMSimulator(null, null, 1.2, {});
    const recordSize = sphFl ? 64 : 80;
    let initBoxSize;
    if (args.particles) {
        const buf = fs.readFileSync(args.particles);
        const ab = buf.buffer.slice(buf.byteOffset, buf.byteOffset + buf.byteLength);
        initBoxSize = args.box;
        sim.reset(ab.byteLength / recordSize, initBoxSize, ab);
    } else if (sphFl) {
        initBoxSize = sphInitBoxSizes[args.preset];
        sim.reset(sphNumParticleParams[args.preset], initBoxSize);
    } else {
        initBoxSize = mlsmpmInitBoxSizes[args.preset];
        sim.reset(mlsmpmNumParticleParams[args.preset], initBoxSize);
    }
    let renderer = null;
    if (args.render) {
        const [w, h] = args.render;
        const fov = 45 * Math.PI / 180;                       // main.ts:137, :142
        const camera = new Camera(null);
        const target = sphFl ? [0, -initBoxSize[1] + 0.1, 0]
                             : [initBoxSize[0] / 2, initBoxSize[1] / 4, initBoxSize[2] / 2];
        const dist = (sphFl ? sphInitDistances : mlsmpmInitDistances)[args.particles ? 1 : args.preset];
        camera.reset({ width: w, height: h }, dist, target, fov, sphFl ? 0.05 : 1.5);
        renderUniformsViews.texel_size.set([1.0 / w, 1.0 / h]);   // main.ts:106
        renderUniformsViews.sphere_size.set([sim.renderDiameter]);
        renderer = new FluidRenderer(null, { width: w, height: h }, null, sphFl ? 0.04 : 0.6, fov,
                                     sim, renderUniformsValues, skyCubemap(64));
    }
    const minClosingSpeed = sphFl ? -0.015 : -0.007;  // main.ts:253
    const realBoxSize = initBoxSize.slice();
    let boxWidthRatio = 1.0;
    const t0 = process.hrtime.bigint();
    for (let f = 0; f < args.frames; ++f) {
        const curBoxWidthRatio = args.slider / 200 + 0.5;
        const dVal = Math.max(curBoxWidthRatio - boxWidthRatio, minClosingSpeed);
        boxWidthRatio += dVal;
        realBoxSize[2] = initBoxSize[2] * boxWidthRatio;
        sim.changeBoxSize(realBoxSize);
        sim.execute(null);
        if (renderer) renderer.execute(null, null, sim.numParticles, args.spheres !== 0);
    }
    sim.synchronize();
    const ms = Number(process.hrtime.bigint() - t0) / 1e6;
    const pv = new Float32Array(sim.readPositions());
    if (args.posvel) fs.writeFileSync(args.posvel, Buffer.from(pv.buffer));
    if (renderer && args.color) fs.writeFileSync(args.color, Buffer.from(renderer.readColor().buffer));
    let cy = 0;
    for (let i = 0; i < sim.numParticles; ++i) cy += pv[8 * i + 1];
    console.log(JSON.stringify({
        solver: args.solver, numParticles: sim.numParticles, gridCount: sim.gridCount, frames: args.frames,
        ms_per_frame: ms / args.frames, boxWidthRatio: boxWidthRatio,
        mean_y: cy / Math.max(1, sim.numParticles), rendered: args.render,
    }));
    if (renderer) renderer.destroy();
    sim.destroy();
}

main();
