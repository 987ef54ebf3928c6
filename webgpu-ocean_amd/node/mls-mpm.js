'use strict';
// mls-mpm.js -- the reference's MLSMPMSimulator surface (mls-mpm/mls-mpm.ts)
// on Node, backed by libmpm355's gfx950 kernels through the N-API addon.
//
// Same class, method and field names as mls-mpm/mls-mpm.ts:12-276:
//   new MLSMPMSimulator(particleBuffer, posvelBuffer, renderDiameter, device)
//   initDambreak(initBoxSize, numParticles) -> ArrayBuffer
//   reset(numParticles, initBoxSize)
//   execute(commandEncoder)          -- 2 substeps, asynchronous on the GPU
//   changeBoxSize(realBoxSize)
//   numParticles, gridCount
// Buffers are host memory here: `particleBuffer` (ArrayBuffer, >= 80 B per
// particle) receives the initial particles like queue.writeBuffer did, and
// `posvelBuffer` (ArrayBuffer/TypedArray, 32 B per particle) is filled by
// readPositions().  `device` may carry {params: {...}} overriding the
// constants of mls-mpm/mls-mpm.ts:53-59.  Written for Node >= 12 (no ?. / ??).

const path = require('path');
const addon = require(path.join(__dirname, 'mpm.node'));

const mlsmpmParticleStructSize = 80;  // mls-mpm/mls-mpm.ts:10
const posvelStructSize = 32;          // main.ts:117

class MLSMPMSimulator {
    constructor(particleBuffer, posvelBuffer, renderDiameter, device) {
        this.particleBuffer = particleBuffer || null;
        this.posvelBuffer = posvelBuffer || null;
        this.renderDiameter = renderDiameter;
        this.device = device || {};
        this.numParticles = 0;
        this.gridCount = 0;
        this.random = (this.device && this.device.random) || Math.random;
        this.handle = addon.create(this.device.params || null, null, 0, [64, 64, 64]);
        this.gridCount = addon.counts(this.handle).gridCount;
    }

    // mls-mpm/mls-mpm.ts:195-223: lattice y outer, x, z inner, spacing 0.65,
    // one jitter scalar 2*random() added to all three coordinates.
    initDambreak(initBoxSize, numParticles) {
        const particlesBuf = new ArrayBuffer(mlsmpmParticleStructSize * numParticles);
        const spacing = 0.65;
        this.numParticles = 0;
        for (let j = 0; j < initBoxSize[1] * 0.80 && this.numParticles < numParticles; j += spacing) {
            for (let i = 3; i < initBoxSize[0] - 4 && this.numParticles < numParticles; i += spacing) {
                for (let k = 3; k < initBoxSize[2] / 2 && this.numParticles < numParticles; k += spacing) {
                    const offset = mlsmpmParticleStructSize * this.numParticles;
                    const position = new Float32Array(particlesBuf, offset + 0, 3);
                    const jitter = 2.0 * this.random();
                    position.set([i + jitter, j + jitter, k + jitter]);
                    this.numParticles++;
                }
            }
        }
        return particlesBuf.slice(0, mlsmpmParticleStructSize * this.numParticles);
    }

    // mls-mpm/mls-mpm.ts:225-243.  `particles` (optional) replaces the
    // dam-break initialiser with caller data of n x 80 B.
    reset(numParticles, initBoxSize, particles) {
        let data = particles;
        if (data === undefined || data === null) {
            data = this.initDambreak(initBoxSize, numParticles);
        } else {
            this.numParticles = numParticles;
        }
        addon.reset(this.handle, data, this.numParticles, initBoxSize);
        const counts = addon.counts(this.handle);
        this.gridCount = counts.gridCount;
        if (this.particleBuffer) {
            const src = new Uint8Array(data.buffer || data, data.byteOffset || 0,
                                       this.numParticles * mlsmpmParticleStructSize);
            new Uint8Array(this.particleBuffer).set(src);
        }
    }

    // mls-mpm/mls-mpm.ts:245-268: two substeps.  The encoder argument is
    // accepted for signature compatibility; work is queued on the GPU stream.
    execute(commandEncoder) {
        addon.step(this.handle, 2);
    }

    // mls-mpm/mls-mpm.ts:270-275
    changeBoxSize(realBoxSize) {
        addon.setBox(this.handle, realBoxSize);
    }

    // Position readout (the buffer the reference's renderer reads on the GPU).
    readPositions(dst) {
        const out = dst || this.posvelBuffer ||
            new ArrayBuffer(posvelStructSize * this.numParticles);
        addon.readPosvel(this.handle, out);
        return out;
    }

    // Device address of the PosVel buffer (BigInt) for GPU-side consumers.
    posvelDevicePtr() {
        return addon.posvelDevicePtr(this.handle);
    }

    readParticles(dst) {
        const out = dst || new ArrayBuffer(mlsmpmParticleStructSize * this.numParticles);
        addon.readParticles(this.handle, out);
        return out;
    }

    writeParticles(src) {
        addon.writeParticles(this.handle, src);
    }

    synchronize() {
        addon.synchronize(this.handle);
    }

    destroy() {
        addon.destroy(this.handle);
    }
}

module.exports = { MLSMPMSimulator, mlsmpmParticleStructSize, posvelStructSize, addon };
