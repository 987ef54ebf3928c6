'use strict';
// fluid-render.js -- the reference's FluidRenderer (render/fluidRender.ts) on
// Node, backed by libmpm355's gfx950 renderer (include/render.h) through the
// N-API addon.
//
//   new FluidRenderer(device, canvas, presentationFormat, radius, fov,
//                     posvelBuffer, renderUniformBuffer, cubemapTextureView)
//                                                        fluidRender.ts:35
//   execute(context, commandEncoder, numParticles, sphereRenderFl)   :332
//
// Arguments as the reference binds them, with host stand-ins for the WebGPU
// objects:
//   device, presentationFormat, context, commandEncoder   accepted, unused
//   canvas               {width, height} of the image
//   posvelBuffer         the PosVel source: an MLSMPMSimulator / SPHSimulator
//                        (its device buffer, read in place), a BigInt device
//                        address, or an ArrayBuffer of host PosVel records
//   renderUniformBuffer  the 272-B uniform ArrayBuffer (camera.js's
//                        renderUniformsValues), read at every execute
//   cubemapTextureView   {size, data}: 6 faces (+X, -X, +Y, -Y, +Z, -Z) of
//                        size x size RGBA float32 texels, rows top to bottom
// After execute(), readColor() / readImage(name) return the frame.

const path = require('path');
const addon = require(path.join(__dirname, 'mpm.node'));

const IMAGES = { depth: 0, filteredDepth: 1, thickness: 2, filteredThickness: 3, color: 4 };

class FluidRenderer {
    constructor(device, canvas, presentationFormat, radius, fov, posvelBuffer,
                renderUniformBuffer, cubemapTextureView) {
        this.width = canvas.width;
        this.height = canvas.height;
        this.radius = radius;
        this.fov = fov;
        this.posvelBuffer = posvelBuffer;
        this.renderUniformBuffer = renderUniformBuffer;
        const cube = cubemapTextureView;
        this.handle = addon.renderCreate(this.width, this.height, radius, fov, cube.data, cube.size);
    }

    posvelSource() {
        const p = this.posvelBuffer;
        if (p && typeof p.posvelDevicePtr === 'function') {
            if (typeof p.synchronize === 'function') p.synchronize();  // PosVel complete
            return p.posvelDevicePtr();
        }
        return p;
    }

    // fluidRender.ts:332-487: depth map, 4 x bilateral, thickness map,
    // 1 x gaussian, fluid shading -- or the sphere pass when sphereRenderFl.
    execute(context, commandEncoder, numParticles, sphereRenderFl) {
        addon.renderExecute(this.handle, this.posvelSource(), numParticles,
                            this.renderUniformBuffer, !!sphereRenderFl);
    }

    // The canvas colour (RGBA float32 per pixel, before 8-bit quantisation),
    // or an intermediate image (one float32 per pixel).
    readImage(name, dst) {
        const n = this.width * this.height * (name === 'color' ? 4 : 1);
        const out = dst || new Float32Array(n);
        addon.renderReadImage(this.handle, IMAGES[name], out);
        return out;
    }

    readColor(dst) {
        return this.readImage('color', dst);
    }

    // Device address (BigInt) of the colour image.
    colorDevicePtr() {
        return addon.renderColorDevicePtr(this.handle);
    }

    // Milliseconds of `frames` frames on the device (events).
    timeFrames(numParticles, sphereRenderFl, frames) {
        return addon.renderTimeFrames(this.handle, this.posvelSource(), numParticles,
                                      this.renderUniformBuffer, !!sphereRenderFl, frames);
    }

    destroy() {
        addon.renderDestroy(this.handle);
    }
}

module.exports = { FluidRenderer, IMAGES, addon };
