"""Camera and render uniforms, as the reference's host code computes them.

Restates camera.ts (Camera.reset / recalculateView) and the matrix functions
it calls from wgpu-matrix 3.3.0 (mat4.perspective with WebGPU's [0, 1] depth
range, mat4.lookAt, mat4.inverse; package-lock.json pins that version), plus
the render uniform buffer of common.ts:1-9.  Matrices are column-major f32
arrays of 16, as in the uniform buffer.  The reference's wgpu-matrix keeps
f32 intermediates; here they are f64, rounded to f32 once, so uniforms agree
with the reference's to f32 rounding.
"""
from __future__ import annotations

import math

import numpy as np


def perspective(fov: float, aspect: float, near: float, far: float) -> np.ndarray:
    """mat4.perspective: depth maps [near, far] to [0, 1]."""
    f = math.tan(math.pi * 0.5 - 0.5 * fov)
    m = np.zeros(16, dtype=np.float64)
    m[0] = f / aspect
    m[5] = f
    m[11] = -1.0
    range_inv = 1.0 / (near - far)
    m[10] = far * range_inv
    m[14] = far * near * range_inv
    return m


def look_at(eye, target, up) -> np.ndarray:
    """mat4.lookAt: the view matrix of a camera at eye looking at target."""
    eye, target, up = (np.asarray(v, dtype=np.float64) for v in (eye, target, up))
    z = eye - target
    z /= np.linalg.norm(z)
    x = np.cross(up, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    y /= np.linalg.norm(y)
    m = np.zeros(16, dtype=np.float64)
    m[0:3] = [x[0], y[0], z[0]]
    m[4:7] = [x[1], y[1], z[1]]
    m[8:11] = [x[2], y[2], z[2]]
    m[12:15] = [-x @ eye, -y @ eye, -z @ eye]
    m[15] = 1.0
    return m


def inverse(m: np.ndarray) -> np.ndarray:
    a = np.asarray(m, dtype=np.float64).reshape(4, 4).T  # column-major -> row-major
    return np.linalg.inv(a).T.reshape(16)


class Camera:
    """camera.ts's orbit camera without the mouse handlers: reset() and
    recalculate_view() as there; orbit(dx, dy) applies a drag, zoom(delta_y)
    a wheel event."""

    def __init__(self):
        self.x_theta = math.pi / 4
        self.y_theta = -math.pi / 12
        self.distance = 70.0
        self.target = [0.0, 0.0, 0.0]
        self.fov = 45 * math.pi / 180
        self.aspect = 1.0
        self.projection = perspective(self.fov, 1.0, 0.1, 500)
        self.view = np.eye(4).reshape(16)

    def reset(self, width: int, height: int, init_distance: float, target, fov: float,
              zoom_rate: float = 1.5):
        """Camera.reset (camera.ts:61-84)."""
        self.x_theta = math.pi / 4 * 1
        self.y_theta = -math.pi / 12
        self.max_y_theta = 0.0
        self.min_y_theta = -0.99 * math.pi / 2.0
        self.sensitivity = 0.005
        self.distance = init_distance
        self.max_distance = 2.0 * init_distance
        self.min_distance = 0.3 * init_distance
        self.target = list(target)
        self.fov = fov
        self.zoom_rate = zoom_rate
        self.aspect = width / height
        self.projection = perspective(fov, self.aspect, 0.1, 500)
        self.recalculate_view()

    def orbit(self, dx: float, dy: float):
        """A mouse drag of (dx, dy) pixels (camera.ts:39-52)."""
        self.x_theta += self.sensitivity * -dx
        self.y_theta = min(max(self.y_theta + self.sensitivity * -dy, self.min_y_theta),
                           self.max_y_theta)
        self.recalculate_view()

    def zoom(self, delta_y: float):
        """A wheel event of deltaY (camera.ts:27-34)."""
        self.distance += (1 if delta_y > 0 else -1) * self.zoom_rate
        self.distance = min(max(self.distance, self.min_distance), self.max_distance)
        self.recalculate_view()

    def recalculate_view(self):
        """Camera.recalculateView (camera.ts:86-101): translate(target),
        rotateY(x_theta), rotateX(y_theta), translate(0, 0, distance)."""
        cx, sx = math.cos(self.x_theta), math.sin(self.x_theta)
        cy, sy = math.cos(self.y_theta), math.sin(self.y_theta)
        d = self.distance
        local = np.array([0.0, -d * sy, d * cy])                       # rotateX applied to (0,0,d)
        rotated = np.array([cx * local[0] + sx * local[2], local[1],  # then rotateY
                            -sx * local[0] + cx * local[2]])
        position = np.asarray(self.target, dtype=np.float64) + rotated
        self.view = look_at(position, self.target, [0.0, 1.0, 0.0])

    def uniforms(self, width: int, height: int, sphere_size: float) -> np.ndarray:
        """The render uniform buffer (common.ts:1-9) as 68 float32:
        texel_size (main.ts:106), sphere_size (the simulator's reset), pad,
        inverse projection, projection, view, inverse view."""
        u = np.zeros(68, dtype=np.float32)
        u[0:2] = [1.0 / width, 1.0 / height]
        u[2] = sphere_size
        u[4:20] = inverse(self.projection)
        u[20:36] = self.projection
        u[36:52] = self.view
        u[52:68] = inverse(self.view)
        return u


def synthetic_cubemap(size: int = 64) -> np.ndarray:
    """A smooth environment (sky gradient over a darker ground, tinted per
    face) as (6, size, size, 4) float32, faces +X, -X, +Y, -Y, +Z, -Z."""
    t = (np.arange(size, dtype=np.float64) + 0.5) / size
    tt, ss = np.meshgrid(t, t, indexing="ij")  # rows (t), columns (s)
    faces = np.zeros((6, size, size, 4), dtype=np.float32)
    tint = [(1.0, 0.9, 0.8), (0.8, 0.9, 1.0), (0.9, 0.95, 1.0), (0.4, 0.35, 0.3),
            (0.95, 1.0, 0.9), (0.9, 0.9, 0.95)]
    for f in range(6):
        sky = 0.45 + 0.5 * (1.0 - tt) if f not in (2, 3) else 0.7 + 0.2 * ss * (1 - ss) * 4
        if f == 3:
            sky = 0.25 + 0.1 * ss
        for c in range(3):
            faces[f, :, :, c] = sky * tint[f][c]
        faces[f, :, :, 3] = 1.0
    return faces
