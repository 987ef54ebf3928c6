'use strict';
// headless.js -- the reference's frame loop without a browser (main.ts:188-283
// minus rendering and UI), for either simulator: preset particle counts and
// boxes (main.ts:128-133), the box-width ratio closing at most 0.007 (MLS-MPM)
// or 0.015 (SPH) per frame toward the slider target (main.ts:249-260),
// changeBoxSize then execute.
//
//   node headless.js [--solver mlsmpm|sph] [--preset 0..3] [--frames N]
//                    [--slider 0..100] [--particles FILE --box X,Y,Z]
//                    [--posvel OUT]
// --particles loads n x 80-B (MLS-MPM) or n x 64-B (SPH) records instead of
// initDambreak (--box: the init box, or SPH's init half box); --posvel writes
// the final n x 32-B PosVel records.  Prints one JSON line.

const fs = require('fs');
const { MLSMPMSimulator } = require('./mls-mpm');
const { SPHSimulator } = require('./sph');

const mlsmpmNumParticleParams = [40000, 70000, 120000, 200000];
const mlsmpmInitBoxSizes = [[35, 25, 55], [40, 30, 60], [45, 40, 80], [50, 50, 80]];
const sphNumParticleParams = [10000, 20000, 30000, 40000];
const sphInitBoxSizes = [[0.7, 2.0, 0.7], [1.0, 2.0, 1.0], [1.2, 2.0, 1.2], [1.4, 2.0, 1.4]];

function parseArgs(argv) {
    const a = { solver: 'mlsmpm', preset: 1, frames: 100, slider: 100, particles: null, box: null,
                posvel: null };
    for (let i = 2; i < argv.length; i += 2) {
        const k = argv[i].replace(/^--/, ''), v = argv[i + 1];
        if (k === 'preset' || k === 'frames' || k === 'slider') a[k] = Number(v);
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
    }
    sim.synchronize();
    const ms = Number(process.hrtime.bigint() - t0) / 1e6;
    const pv = new Float32Array(sim.readPositions());
    if (args.posvel) fs.writeFileSync(args.posvel, Buffer.from(pv.buffer));
    let cy = 0;
    for (let i = 0; i < sim.numParticles; ++i) cy += pv[8 * i + 1];
    console.log(JSON.stringify({
        solver: args.solver, numParticles: sim.numParticles, gridCount: sim.gridCount, frames: args.frames,
        ms_per_frame: ms / args.frames, boxWidthRatio: boxWidthRatio,
        mean_y: cy / Math.max(1, sim.numParticles),
    }));
    sim.destroy();
}

main();
