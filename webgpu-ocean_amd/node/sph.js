'use strict';
// sph.js -- the reference's SPHSimulator surface (sph/sph.ts) on Node, backed
// by libmpm355's gfx950 SPH kernels through the N-API addon.
//
// Same class, method and field names as sph/sph.ts:15-371:
//   new SPHSimulator(particleBuffer, posvelBuffer, renderDiameter, device)
//   initDambreak(initHalfBoxSize, numParticles) -> ArrayBuffer
//   reset(numParticles, initHalfBoxSize)
//   execute(commandEncoder)          -- 2 substeps, asynchronous on the GPU
//   changeBoxSize(realBoxSize)
//   numParticles, gridCount, kernelRadius
// Buffers are host memory here: `particleBuffer` (ArrayBuffer, >= 64 B per
// particle) receives the initial particles like queue.writeBuffer did, and
// `posvelBuffer` (32 B per particle) is filled by readPositions().  `device`
// may carry {params: {...}} overriding the constants of sph/sph.ts:41,72-77
// and {random} replacing Math.random in initDambreak.

const path = require('path');
const addon = require(path.join(__dirname, 'mpm.node'));

const sphParticleStructSize = 64;  // sph/sph.ts:13
const posvelStructSize = 32;       // main.ts:117

class SPHSimulator {
    constructor(particleBuffer, posvelBuffer, renderDiameter, device) {
        this.particleBuffer = particleBuffer || null;
        this.posvelBuffer = posvelBuffer || null;
        this.renderDiameter = renderDiameter;
        this.device = device || {};
        const params = this.device.params || null;
        this.kernelRadius = (params && params.kernelRadius) || 0.07;  // sph/sph.ts:41
        this.numParticles = 0;
        this.random = this.device.random || Math.random;
        this.handle = addon.sphCreate(params, null, 0, [1, 2, 1]);
        this.gridCount = addon.sphCounts(this.handle).gridCount;
    }

    // sph/sph.ts:338-363: y from -0.95*hy upward without bound, x in
    // [-0.95*hx, 0.95*hx), z in [-0.95*hz, 0), spacing 0.5*kernelRadius, one
    // jitter 0.001*random() added to all three coordinates.
    initDambreak(initHalfBoxSize, numParticles) {
        const particlesBuf = new ArrayBuffer(sphParticleStructSize * numParticles);
        this.numParticles = 0;
        const DIST_FACTOR = 0.5;
        for (let y = -initHalfBoxSize[1] * 0.95; this.numParticles < numParticles;
             y += DIST_FACTOR * this.kernelRadius) {
            for (let x = -0.95 * initHalfBoxSize[0];
                 x < 0.95 * initHalfBoxSize[0] && this.numParticles < numParticles;
                 x += DIST_FACTOR * this.kernelRadius) {
                for (let z = -0.95 * initHalfBoxSize[2];
                     z < 0 * initHalfBoxSize[2] && this.numParticles < numParticles;
                     z += DIST_FACTOR * this.kernelRadius) {
                    const jitter = 0.001 * this.random();
                    const offset = sphParticleStructSize * this.numParticles;
                    new Float32Array(particlesBuf, offset, 3).set([x + jitter, y + jitter, z + jitter]);
                    this.numParticles++;
                }
            }
        }
        return particlesBuf;
    }

    // sph/sph.ts:281-299.  `particles` (optional) replaces the dam-break
    // initialiser with caller data of n x 64 B.  The real box becomes the
    // initial half box.
    reset(numParticles, initHalfBoxSize, particles) {
        let data = particles;
        if (data === undefined || data === null) {
            data = this.initDambreak(initHalfBoxSize, numParticles);
        } else {
            this.numParticles = numParticles;
        }
        addon.sphReset(this.handle, data, this.numParticles, initHalfBoxSize);
        this.gridCount = addon.sphCounts(this.handle).gridCount;
        if (this.particleBuffer) {
            const src = new Uint8Array(data.buffer || data, data.byteOffset || 0,
                                       this.numParticles * sphParticleStructSize);
            new Uint8Array(this.particleBuffer).set(src);
        }
    }

    // sph/sph.ts:301-336: two substeps.  The encoder argument is accepted for
    // signature compatibility; work is queued on the GPU stream.
    execute(commandEncoder) {
        addon.sphStep(this.handle, 2);
    }

    // sph/sph.ts:365-370: the half extents the walls push against.
    changeBoxSize(realBoxSize) {
        addon.sphSetBox(this.handle, realBoxSize);
    }

    readPositions(dst) {
        const out = dst || this.posvelBuffer ||
            new ArrayBuffer(posvelStructSize * this.numParticles);
        addon.sphReadPosvel(this.handle, out);
        return out;
    }

    posvelDevicePtr() {
        return addon.sphPosvelDevicePtr(this.handle);
    }

    readParticles(dst) {
        const out = dst || new ArrayBuffer(sphParticleStructSize * this.numParticles);
        addon.sphReadParticles(this.handle, out);
        return out;
    }

    writeParticles(src) {
        addon.sphWriteParticles(this.handle, src);
    }

    synchronize() {
        addon.sphSynchronize(this.handle);
    }

    destroy() {
        addon.sphDestroy(this.handle);
    }
}

module.exports = { SPHSimulator, sphParticleStructSize, posvelStructSize, addon };
