"""Synthetic dam-break scenes.

Fluid blocks follow DemoBase::createFluidBlocks (Demos/Common/DemoBase.cpp:539-600,
mode 0): a cubic lattice of spacing 2r starting at the block corner + 2r.
Walls are single-layer particle lattices on the faces of a box (the
configurations ask for grid walls instead of the reference's Poisson-disk
sampled meshes).  Positions are jittered by U(-j r, j r) per coordinate from a
seeded generator.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def box_surface(lo, hi, spacing, open_top=False):
    """Lattice points on the faces of the box [lo, hi] (no duplicates).

    The spacing is adjusted per axis so the lattice ends exactly on the
    faces.  With open_top the face y = hi[1] is left out."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    counts = np.maximum(np.rint((hi - lo) / spacing).astype(np.int64), 1)
    axes = [lo[a] + (hi[a] - lo[a]) * np.arange(counts[a] + 1) / counts[a] for a in range(3)]
    pts = []
    # iterate y-layers to keep memory bounded for large boxes
    X, Z = np.meshgrid(np.arange(counts[0] + 1), np.arange(counts[2] + 1), indexing="ij")
    X = X.ravel()
    Z = Z.ravel()
    xz_border = (X == 0) | (X == counts[0]) | (Z == 0) | (Z == counts[2])
    for j in range(counts[1] + 1):
        if j == 0 or (j == counts[1] and not open_top):
            sel = np.ones_like(X, dtype=bool)
        else:
            sel = xz_border
        xs, zs = X[sel], Z[sel]
        layer = np.empty((xs.size, 3))
        layer[:, 0] = axes[0][xs]
        layer[:, 1] = axes[1][j]
        layer[:, 2] = axes[2][zs]
        pts.append(layer)
    return np.concatenate(pts, axis=0)


def fluid_block(start, counts, spacing):
    """counts[0] x counts[1] x counts[2] lattice from start (x outer, z inner,
    the loop order of DemoBase::createFluidBlocks)."""
    i, j, k = np.meshgrid(np.arange(counts[0]), np.arange(counts[1]), np.arange(counts[2]), indexing="ij")
    p = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64) * spacing
    return p + np.asarray(start, np.float64)


@dataclass
class Scene:
    name: str
    particle_radius: float
    fluid: np.ndarray            # (n, 3) float32
    boundary: np.ndarray         # (m, 3) float32
    tank_lo: tuple = (0.0, 0.0, 0.0)
    tank_hi: tuple = (1.0, 1.0, 1.0)
    params: dict = field(default_factory=dict)

    @property
    def n_fluid(self):
        return self.fluid.shape[0]

    @property
    def n_boundary(self):
        return self.boundary.shape[0]


def dam_break(block=(20, 40, 20), r=0.025, tank=(4.0, 3.0, 1.5), wall_spacing=None, open_top=False,
              jitter=0.01, seed=12345, obstacle=None, name="dam_break"):
    """Fluid block at the tank corner (start at 2r from the walls, spacing 2r).

    obstacle: optional (lo, hi) box sampled on its surface at the wall spacing."""
    d = 2.0 * r
    ws = d if wall_spacing is None else wall_spacing
    rng = np.random.default_rng(seed)
    fl = fluid_block((d, d, d), block, d)
    if jitter:
        fl = fl + rng.uniform(-jitter * r, jitter * r, size=fl.shape)
    walls = box_surface((0.0, 0.0, 0.0), tank, ws, open_top=open_top)
    if obstacle is not None:
        walls = np.concatenate([walls, box_surface(obstacle[0], obstacle[1], ws)], axis=0)
    return Scene(name=name, particle_radius=r, fluid=fl.astype(np.float32), boundary=walls.astype(np.float32),
                 tank_lo=(0.0, 0.0, 0.0), tank_hi=tuple(tank))


# --- the benchmark configurations -------------------------------------------

def config1(seed=12345):
    """16,000 fluid particles (20x40x20, r = 0.025) in a closed 4 x 3 x 1.5 m tank
    with single-layer walls at spacing 2r (18,002 boundary particles)."""
    return dam_break(block=(20, 40, 20), r=0.025, tank=(4.0, 3.0, 1.5), seed=seed, name="C1_dam_break_16k")


def config2(seed=12345):
    """1,048,576 fluid particles (128 wide x 128 tall x 64 deep, r = 0.01) in a 6 x 3 x 2 m
    tank with an open top and walls at spacing r (~600k boundary particles)."""
    return dam_break(block=(128, 128, 64), r=0.01, tank=(6.0, 3.0, 2.0), wall_spacing=0.01, open_top=True,
                     seed=seed, name="C2_dam_break_1M")


def config4(seed=12345):
    """4,194,304 fluid particles (256 x 128 x 128, r = 0.0075) in a closed 6 x 3 x 2 m tank
    with walls at spacing 2r, plus a 0.5 m cube obstacle on the floor in the flow path."""
    obstacle = ((4.5, 0.0, 0.75), (5.0, 0.5, 1.25))
    return dam_break(block=(256, 128, 128), r=0.0075, tank=(6.0, 3.0, 2.0), seed=seed, obstacle=obstacle,
                     name="C4_dam_break_4M_obstacle")


def small_dam_break(nx=10, ny=16, nz=10, r=0.025, tank=(1.0, 1.2, 0.6), seed=7):
    """A few thousand particles for quick parity tests."""
    return dam_break(block=(nx, ny, nz), r=r, tank=tank, seed=seed, name="small_dam_break")


def seeded_dam_break(k, n_target, r=0.0125):
    """Training scenes of the learned warm start: dam breaks of about n_target
    particles with a random block aspect ratio in [0.5, 2] (seed 1000 + k)."""
    rng = np.random.default_rng(1000 + k)
    ax = rng.uniform(0.5, 2.0)
    az = rng.uniform(0.5, 2.0)
    ny = max(8, int(round((n_target / (ax * az)) ** (1.0 / 3.0))))
    nx = max(4, int(round(ax * ny)))
    nz = max(4, int(round(az * ny)))
    d = 2 * r
    tank = (max(2.5 * nx * d, 1.0), max(1.5 * ny * d, 0.5), max(1.2 * nz * d + 4 * d, 0.5))
    return dam_break(block=(nx, ny, nz), r=r, tank=tank, seed=1000 + k, name="train_%d" % k)
