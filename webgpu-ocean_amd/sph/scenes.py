"""Particle initialisers of the SPH path: the reference's initDambreak and a
seeded variant.  Records use the reference's 64-B layout (position @0, v @16,
force @32, density @44, nearDensity @48; sph/sph.ts:348-354) as
``(n, 16) float32`` arrays.
"""
from __future__ import annotations

import random as _random

import numpy as np

PARTICLE_FLOATS = 16

# The reference's scene table (main.ts:131-132): particle counts and initial
# half box sizes of the four SPH settings.
NUM_PARTICLES = (10000, 20000, 30000, 40000)
INIT_HALF_BOXES = ((0.7, 2.0, 0.7), (1.0, 2.0, 1.0), (1.2, 2.0, 1.2), (1.4, 2.0, 1.4))


def records(positions: np.ndarray, velocities: np.ndarray | None = None) -> np.ndarray:
    positions = np.asarray(positions)
    out = np.zeros((positions.shape[0], PARTICLE_FLOATS), dtype=np.float32)
    out[:, 0:3] = positions
    if velocities is not None:
        out[:, 4:7] = velocities
    return out


def init_dambreak(init_half_box, num_particles: int, kernel_radius: float = 0.07,
                  rand=None) -> np.ndarray:
    """SPHSimulator.initDambreak (sph/sph.ts:338-363).

    Same loop nest in double precision: y from -0.95*hy upward, x in
    [-0.95*hx, 0.95*hx), z in [-0.95*hz, 0), spacing 0.5*kernel_radius; each
    particle gets ONE jitter 0.001*rand() added to all three axes (:346,355).
    The y loop has no upper bound: it runs until num_particles are placed.
    ``rand`` defaults to Python's ``random.random`` (the reference uses
    unseeded Math.random).
    """
    rand = rand or _random.random
    step = 0.5 * kernel_radius
    hx, hy, hz = (float(v) for v in init_half_box)
    if num_particles > 0 and not (-0.95 * hx < 0.95 * hx and -0.95 * hz < 0 * hz):
        raise ValueError("the initial box holds no lattice site")
    pos = []
    y = -hy * 0.95
    while len(pos) < num_particles:
        x = -0.95 * hx
        while x < 0.95 * hx and len(pos) < num_particles:
            z = -0.95 * hz
            while z < 0 * hz and len(pos) < num_particles:
                jitter = 0.001 * rand()
                pos.append((x + jitter, y + jitter, z + jitter))
                z += step
            x += step
        y += step
    return records(np.array(pos, dtype=np.float64).reshape(-1, 3))


def dam_break(setting: int = 3, seed: int = 0, num_particles: int | None = None) -> np.ndarray:
    """initDambreak of the reference's SPH setting `setting` (0..3) with a
    seeded jitter; num_particles overrides that setting's count."""
    rng = _random.Random(seed)
    n = NUM_PARTICLES[setting] if num_particles is None else num_particles
    return init_dambreak(INIT_HALF_BOXES[setting], n, rand=rng.random)
