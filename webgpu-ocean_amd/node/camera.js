'use strict';
// camera.js -- the reference's orbit camera (camera.ts) and render uniform
// buffer (common.ts) on Node, for the FluidRenderer of fluid-render.js.
//
//   renderUniformsValues / renderUniformsViews   common.ts:1-9 (272 B)
//   new Camera(canvasElement)                     camera.ts:20 (no DOM here:
//       drag(deltaX, deltaY) and wheel(deltaY) stand in for the mouse handlers
//       of camera.ts:21-56)
//   reset(canvasElement, initDistance, target, fov, zoomRate)   camera.ts:61-84
//   recalculateView()                                            camera.ts:86-101
//
// The matrices follow wgpu-matrix 3.3.0's published definitions (the
// version package-lock.json pins): perspective with WebGPU's [0, 1] depth
// range, lookAt, inverse; column-major. They are evaluated in doubles and
// stored as float32, as in render/camera.py.

const renderUniformsValues = new ArrayBuffer(272);
const renderUniformsViews = {
    texel_size: new Float32Array(renderUniformsValues, 0, 2),
    sphere_size: new Float32Array(renderUniformsValues, 8, 2),
    inv_projection_matrix: new Float32Array(renderUniformsValues, 16, 16),
    projection_matrix: new Float32Array(renderUniformsValues, 80, 16),
    view_matrix: new Float32Array(renderUniformsValues, 144, 16),
    inv_view_matrix: new Float32Array(renderUniformsValues, 208, 16),
};

function perspective(fov, aspect, near, far) {
    const f = Math.tan(Math.PI * 0.5 - 0.5 * fov);
    const m = new Array(16).fill(0);
    m[0] = f / aspect;
    m[5] = f;
    m[11] = -1;
    const rangeInv = 1 / (near - far);
    m[10] = far * rangeInv;
    m[14] = far * near * rangeInv;
    return m;
}

function normalize(v) {
    const l = Math.hypot(v[0], v[1], v[2]);
    return [v[0] / l, v[1] / l, v[2] / l];
}

function cross(a, b) {
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]];
}

function dot(a, b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

function lookAt(eye, target, up) {
    const z = normalize([eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]]);
    const x = normalize(cross(up, z));
    const y = normalize(cross(z, x));
    return [x[0], y[0], z[0], 0, x[1], y[1], z[1], 0, x[2], y[2], z[2], 0,
            -dot(x, eye), -dot(y, eye), -dot(z, eye), 1];
}

// Inverse of a column-major 4x4 by Gauss-Jordan elimination with partial
// pivoting (doubles).
function inverse(m) {
    const a = [];
    for (let r = 0; r < 4; ++r) {
        a.push([m[r], m[4 + r], m[8 + r], m[12 + r], r === 0 ? 1 : 0, r === 1 ? 1 : 0,
                r === 2 ? 1 : 0, r === 3 ? 1 : 0]);
    }
    for (let c = 0; c < 4; ++c) {
        let p = c;
        for (let r = c + 1; r < 4; ++r) if (Math.abs(a[r][c]) > Math.abs(a[p][c])) p = r;
        [a[c], a[p]] = [a[p], a[c]];
        const d = a[c][c];
        for (let k = 0; k < 8; ++k) a[c][k] /= d;
        for (let r = 0; r < 4; ++r) {
            if (r === c) continue;
            const f = a[r][c];
            if (f !== 0) for (let k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
        }
    }
    const out = new Array(16);
    for (let r = 0; r < 4; ++r) for (let c = 0; c < 4; ++c) out[c * 4 + r] = a[r][4 + c];
    return out;
}

class Camera {
    constructor(canvasElement) {
        this.canvas = canvasElement || null;
        this.currentXtheta = Math.PI / 4;
        this.currentYtheta = -Math.PI / 12;
        this.currentDistance = 70;
        this.target = [0, 0, 0];
    }

    // camera.ts:61-84.  canvasElement: {clientWidth, clientHeight}
    // (or {width, height}).
    reset(canvasElement, initDistance, target, fov, zoomRate) {
        this.isDragging = false;
        this.currentXtheta = Math.PI / 4 * 1;
        this.currentYtheta = -Math.PI / 12;
        this.maxYTheta = 0;
        this.minYTheta = -0.99 * Math.PI / 2.;
        this.sensitivity = 0.005;
        this.currentDistance = initDistance;
        this.maxDistance = 2. * this.currentDistance;
        this.minDistance = 0.3 * this.currentDistance;
        this.target = target;
        this.fov = fov;
        this.zoomRate = zoomRate;
        const c = canvasElement || this.canvas;
        const w = c.clientWidth !== undefined ? c.clientWidth : c.width;
        const h = c.clientHeight !== undefined ? c.clientHeight : c.height;
        const projection = perspective(fov, w / h, 0.1, 500);
        renderUniformsViews.projection_matrix.set(projection);
        renderUniformsViews.inv_projection_matrix.set(inverse(projection));
        this.recalculateView();
    }

    // The mousemove handler's effect (camera.ts:36-51): deltaX = prevX -
    // currentX, deltaY = prevY - currentY.
    drag(deltaX, deltaY) {
        this.currentXtheta += this.sensitivity * deltaX;
        this.currentYtheta += this.sensitivity * deltaY;
        if (this.currentYtheta > this.maxYTheta) this.currentYtheta = this.maxYTheta;
        if (this.currentYtheta < this.minYTheta) this.currentYtheta = this.minYTheta;
        this.recalculateView();
    }

    // The wheel handler's effect (camera.ts:27-34).
    wheel(deltaY) {
        this.currentDistance += ((deltaY > 0) ? 1 : -1) * this.zoomRate;
        if (this.currentDistance < this.minDistance) this.currentDistance = this.minDistance;
        if (this.currentDistance > this.maxDistance) this.currentDistance = this.maxDistance;
        this.recalculateView();
    }

    // camera.ts:86-101: the eye is target + rotateY(x) rotateX(y) (0, 0, d).
    recalculateView() {
        const cx = Math.cos(this.currentXtheta), sx = Math.sin(this.currentXtheta);
        const cy = Math.cos(this.currentYtheta), sy = Math.sin(this.currentYtheta);
        const d = this.currentDistance;
        const local = [0, -d * sy, d * cy];
        const rotated = [cx * local[0] + sx * local[2], local[1], -sx * local[0] + cx * local[2]];
        const position = [this.target[0] + rotated[0], this.target[1] + rotated[1],
                          this.target[2] + rotated[2]];
        const view = lookAt(position, this.target, [0, 1, 0]);
        renderUniformsViews.view_matrix.set(view);
        renderUniformsViews.inv_view_matrix.set(inverse(view));
    }
}

module.exports = { Camera, renderUniformsValues, renderUniformsViews, perspective, lookAt, inverse };
