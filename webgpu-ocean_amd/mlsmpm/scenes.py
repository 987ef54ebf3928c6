"""Particle initialisers: the reference's initDambreak and the seeded
dam-break scenes the benchmarks and parity tests run.

All return ``(n, 20) float32`` arrays in the reference's 80-B particle layout
(position @0, v @16, C column-major @32; mls-mpm/mls-mpm.ts:204-209).
"""
from __future__ import annotations

import math
import random as _random

import numpy as np

PARTICLE_FLOATS = 20


def records(positions: np.ndarray, velocities: np.ndarray | None = None) -> np.ndarray:
    """Packs (n,3) positions (and optional velocities) into 80-B records; C = 0."""
    positions = np.asarray(positions)
    out = np.zeros((positions.shape[0], PARTICLE_FLOATS), dtype=np.float32)
    out[:, 0:3] = positions
    if velocities is not None:
        out[:, 4:7] = velocities
    return out


def init_dambreak(init_box, num_particles: int, rand=None) -> np.ndarray:
    """MLSMPMSimulator.initDambreak (mls-mpm/mls-mpm.ts:195-223).

    Same loop nest, bounds and 0.65 spacing; each particle gets ONE jitter
    scalar 2*rand() added to all three axes (:210-211).  ``rand`` defaults to
    Python's ``random.random`` (the reference uses unseeded Math.random).
    """
    rand = rand or _random.random
    spacing = 0.65
    pos = []
    j = 0.0
    while j < init_box[1] * 0.80 and len(pos) < num_particles:
        i = 3.0
        while i < init_box[0] - 4 and len(pos) < num_particles:
            k = 3.0
            while k < init_box[2] / 2 and len(pos) < num_particles:
                jitter = 2.0 * rand()
                pos.append((i + jitter, j + jitter, k + jitter))
                k += spacing
            i += spacing
        j += spacing
    return records(np.array(pos, dtype=np.float64).reshape(-1, 3))


def lattice(lo, hi, spacing: float, limit: int) -> np.ndarray:
    """Lattice sites lo + k*spacing < hi, filled y outer, x, z inner (the loop
    order of mls-mpm/mls-mpm.ts:201-203), capped at ``limit`` sites."""
    axes = [lo[a] + spacing * np.arange(int(math.ceil((hi[a] - lo[a]) / spacing)) + 1)
            for a in range(3)]
    axes = [ax[ax < hi[a]] for a, ax in enumerate(axes)]
    ys, xs, zs = np.meshgrid(axes[1], axes[0], axes[2], indexing="ij")
    pts = np.stack([xs.ravel(), ys.ravel(), zs.ravel()], axis=1)
    return pts[:limit]


def jittered_block(lo, hi, spacing: float, n: int, rng: np.random.Generator,
                   amp: float = 0.1) -> np.ndarray:
    pts = lattice(lo, hi, spacing, n)
    if pts.shape[0] < n:
        raise ValueError(f"block {lo}-{hi} at spacing {spacing} holds {pts.shape[0]} < {n}")
    return pts + rng.uniform(-amp, amp, size=pts.shape)


def dam_break(n: int, box, spacing: float, seed: int, frac: float = 0.45,
              amp: float = 0.1) -> np.ndarray:
    """Seeded dam-break block in the lower-left corner: x and z span
    [3, 3 + frac*box), y rises from 3 until n particles are placed."""
    rng = np.random.default_rng(seed)
    lo = (3.0, 3.0, 3.0)
    hi = (3.0 + frac * box[0], box[1] - 3.0, 3.0 + frac * box[2])
    return records(jittered_block(lo, hi, spacing, n, rng, amp))


# --- benchmark / parity configurations --------------------------------------

def config0(seed: int = 0):
    """40,000 particles, 64^3 box, block x,z in [3,35), y in [3,45), spacing 0.55."""
    rng = np.random.default_rng(seed)
    pts = jittered_block((3.0, 3.0, 3.0), (35.0, 45.0, 35.0), 0.55, 40_000, rng)
    return records(pts), [64.0, 64.0, 64.0]


def config1(seed: int = 1, n: int = 300_000):
    """300,000 particles on an 80^3 grid; block over the lower-left 45% of x and z."""
    box = [80.0, 80.0, 80.0]
    return dam_break(n, box, 0.55, seed), box


def config2(seed: int = 2, n: int = 2_000_000):
    """2,000,000 particles on 160^3: two spacing-0.5 blocks in opposite corners."""
    rng = np.random.default_rng(seed)
    box = [160.0, 160.0, 160.0]
    half = n // 2
    a = jittered_block((3.0, 3.0, 3.0), (53.0, 157.0, 53.0), 0.5, half, rng)
    b = jittered_block((107.0, 3.0, 107.0), (157.0, 157.0, 157.0), 0.5, n - half, rng)
    return records(np.concatenate([a, b])), box


def small_scene_capacity(box, spacing: float = 0.55) -> int:
    """Largest particle count small_scene() can place in `box`."""
    frac = 0.6
    hi = (3.0 + frac * (box[0] - 6), box[1] - 3.0, 3.0 + frac * (box[2] - 6))
    return lattice((3.0, 3.0, 3.0), hi, spacing, 1 << 62).shape[0]


def small_scene(n: int, box, seed: int, spacing: float = 0.55, velocity: float = 0.0,
                affine: float = 0.0):
    """Small randomized scene for parity tests: jittered block plus optional
    random initial velocities and APIC matrices."""
    rng = np.random.default_rng(seed)
    frac = 0.6
    hi = (3.0 + frac * (box[0] - 6), box[1] - 3.0, 3.0 + frac * (box[2] - 6))
    pts = jittered_block((3.0, 3.0, 3.0), hi, spacing, n, rng, 0.2)
    rec = records(pts)
    if velocity:
        rec[:, 4:7] = rng.uniform(-velocity, velocity, size=(n, 3))
    if affine:
        C = rng.uniform(-affine, affine, size=(n, 3, 3)).astype(np.float32)
        for c in range(3):
            rec[:, 8 + 4 * c:8 + 4 * c + 3] = C[:, c, :]
    return rec


BATCH_EDGES = (32, 48, 64, 96)


def batch_block_capacity(edge: float, spacing: float = 0.55, frac: float = 0.6) -> int:
    """Particles a dam-break block [3, 3+frac*(edge-6))^2 x [3, edge-3) holds."""
    w = int(math.ceil(frac * (edge - 6) / spacing))
    h = int(math.ceil((edge - 6) / spacing))
    return w * w * h


def config3(seed: int = 3, scenes: int = 64, n_lo: int = 5_000, n_hi: int = 200_000):
    """Uneven batch: particle counts log-uniform in [n_lo, n_hi]; each cubic
    box edge drawn from {32, 48, 64, 96} among the edges whose block holds
    the scene's particles; stiffness ~ U[1,5], viscosity ~ U[0.02, 0.2].
    Returns a list of (particles, box, {"stiffness", "dynamic_viscosity"})."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(scenes):
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        edges = [e for e in BATCH_EDGES if batch_block_capacity(e) >= n]
        edge = float(edges[int(rng.integers(len(edges)))])
        box = [edge, edge, edge]
        lo = (3.0, 3.0, 3.0)
        hi = (3.0 + 0.6 * (edge - 6), edge - 3.0, 3.0 + 0.6 * (edge - 6))
        pts = jittered_block(lo, hi, 0.55, n, rng, 0.1)
        phys = {"stiffness": float(rng.uniform(1.0, 5.0)),
                "dynamic_viscosity": float(rng.uniform(0.02, 0.2))}
        out.append((records(pts), box, phys))
    return out


def config4(seed: int = 4, n: int = 100_000, edge: float = 64.0, vel: float = 0.5):
    """Differentiable scene: seeded dam-break block (spacing 0.55, lower-left
    45% of x and z) with initial velocities ~ U(-vel, vel), and a seeded
    target point for the centre-of-mass loss.  The FD-check instance is
    config4(n=1000, edge=16)."""
    rng = np.random.default_rng(seed)
    box = [edge, edge, edge]
    frac = 0.45 if edge >= 32 else 0.6
    lo = (3.0, 3.0, 3.0)
    hi = (3.0 + frac * (edge - 6), edge - 3.0, 3.0 + frac * (edge - 6))
    pts = jittered_block(lo, hi, 0.55, n, rng, 0.1)
    rec = records(pts, rng.uniform(-vel, vel, size=(n, 3)))
    target = rng.uniform(0.3 * edge, 0.7 * edge, size=3).tolist()
    return rec, box, target
