"""Open-boundary wave tank: a long basin with a sinusoidal piston wave maker at
one end, an absorbing open boundary at the other and a floating box.

The solver is SPH::DFSPHCUDA (libdfsph_host.so) with
  * the piston as a DFSPH_BODY_SINE rigid body (a moving wall; the reference
    has no wave maker),
  * the box as a DFSPH_BODY_DYNAMIC rigid body (vector_dynamic_bodies_data,
    DFSPH_c_arrays_structure.cpp:897-905),
  * the far end as OpenBoundariesSimple (OpenBoundariesSimple.cu:117-621,
    driven as DFSPHCUDA::applyOpenBoundaries does, DFSPH_CUDA.cpp:3889-3924):
    fluid that rises above the rest level past the absorbing line is removed
    (outflow), holes below it are refilled at rest positions (inflow), so the
    far end neither reflects the waves nor drains,
  * every step's compaction of removed particles, and optionally the learned
    warm start.
With window_tank() the simulated region is instead a window around the box
(the reference's DynamicWindow, DynamicWindow.cu:150-3760, driven as
DFSPHCUDA::applyDynamicWindow, DFSPH_CUDA.cpp:3988-4139): the region, walled
at its two x faces by plates that move with it, follows the box, and fluid is
loaded from the basin's rest lattice (the background buffer) along its
leading edge while what it leaves behind is removed.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .scene import box_surface, fluid_block


@dataclass
class WaveTank:
    particle_radius: float
    length: float
    height: float
    width: float
    depth: float                    # nominal water depth
    level: float                    # rest surface: top particle layer + r
    walls: np.ndarray               # (m, 3) float32
    fluid: np.ndarray               # (n, 3) float32
    bodies: list = field(default_factory=list)        # [(DFSPHRigidBody kwargs, body-frame particles)]
    absorb_x: float = 0.0           # start of the absorbing zone
    inflow: np.ndarray = None       # (k, 3) float32 rest positions of the absorbing zone
    piston: dict = None
    box: dict = None

    @property
    def n_fluid(self):
        return len(self.fluid)

    @property
    def n_boundary(self):
        return len(self.walls)


def build(r=0.008, length=10.0, height=2.0, width=1.0, depth=0.9, amplitude=0.1, period=1.5,
          box_side=None, box_density=500.0, box_x=None, absorb=0.6, seed=12345, jitter=0.01) -> WaveTank:
    """The basin [0, length] x [0, height] x [0, width] (y up), water at rest to
    `depth`, the piston plate at x = 0.15 (its rest place) moving along x with
    the given amplitude and period, the box floating half submerged at
    box_x (default: mid-basin), the absorbing zone over the last `absorb` m.
    At r = 0.008 this is about 2M fluid particles (the 10 x 2 x 1 m basin)."""
    d = 2.0 * r
    walls = box_surface((0.0, 0.0, 0.0), (length, height, width), d)
    # piston: a plate two particle layers thick across the basin, above the water
    px = 0.15
    plate_h = min(height - 2 * d, depth + 0.5 * (height - depth))
    plate = box_surface((-0.5 * d, 0.0, 0.0), (0.5 * d, plate_h, width - 4 * d), d)
    plate_c = plate.mean(0)
    piston = dict(type=N.BODY_SINE, position=(px, d + plate_c[1], 2 * d + plate_c[2]), axis=(1.0, 0.0, 0.0),
                  amplitude=amplitude, period=period)
    plate = plate - plate_c
    # jittered water block from one spacing in front of the plate at its rest
    # place to one spacing from the far wall, and to 1-1.5 spacings from the
    # side walls (wider gaps let a deep column collapse into them and throw
    # jets up the walls)
    counts = (int(np.floor((length - d - (px + 1.5 * d)) / d + 1e-6)) + 1, int(round(depth / d)),
              int(np.floor(width / d - 2.0 + 1e-6)) + 1)
    x0 = length - d - (counts[0] - 1) * d
    rng = np.random.default_rng(seed)
    fl = fluid_block((x0, d, d), counts, d)
    if jitter:
        fl = fl + rng.uniform(-jitter * r, jitter * r, size=fl.shape)
    side = box_side if box_side is not None else 0.3 * width
    bx = box_x if box_x is not None else 0.5 * length
    c = np.array([bx, depth, 0.5 * width])
    fl = fl[np.abs(fl - c).max(1) >= 0.5 * side + 1.5 * r]
    cube = box_surface((-0.5 * side,) * 3, (0.5 * side,) * 3, d)
    m = box_density * side ** 3
    inertia = m * 2 * side * side / 12.0
    box = dict(type=N.BODY_DYNAMIC, mass=m, inertia=(inertia,) * 3, position=tuple(c))
    # absorbing zone: rest-lattice positions at least two and a half layers
    # below the surface (refills where the water column has lost particles,
    # not at the free surface, where the density is low anyway)
    level = float(fl[:, 1].max()) + r
    ax = length - absorb
    sel = (fl[:, 0] >= ax) & (fl[:, 1] <= level - 2.5 * d)
    inflow = fl[sel].copy()
    return WaveTank(particle_radius=r, length=length, height=height, width=width, depth=depth, level=level,
                    walls=walls.astype(np.float32), fluid=fl.astype(np.float32),
                    bodies=[(piston, plate.astype(np.float32)), (box, cube.astype(np.float32))],
                    absorb_x=ax, inflow=inflow.astype(np.float32), piston=piston, box=box)


def open_boundary_surfaces(t: WaveTank):
    """(boundary, interior, fluid_surface) of the absorbing end: the interior
    is the basin before the absorbing line; the fluid surface the absorbing
    zone below the rest surface (crests above it are removed); the boundary's
    normal (-x) admits refills where the fluid moves back toward the basin
    (under wave troughs)."""
    h = 2.0 * t.height
    boundary = N.surface("plane", o=(t.absorb_x, 0.0, 0.0), n=(-1.0, 0.0, 0.0))
    interior = N.surface("cuboid", center=(0.5 * t.absorb_x, 0.5 * t.height, 0.5 * t.width),
                         half=(0.5 * t.absorb_x, h, t.width))
    zone = t.length - t.absorb_x
    surf = N.surface("cuboid", center=(t.absorb_x + 0.5 * zone, 0.5 * t.level, 0.5 * t.width),
                     half=(0.5 * zone + t.particle_radius, 0.5 * t.level, t.width))
    return boundary, interior, surf


def make_solver(t: WaveTank, params: N.DFSPHParams, learned: bool = False) -> N.HostSolver:
    """SPH::DFSPHCUDA over the tank: walls, water, the piston and the box as
    rigid bodies, the absorbing open boundary (fluid capacity leaves room for
    the refills); learned: load the committed warm-start model (params must
    select the learned warm start)."""
    hs = N.HostSolver(params, t.walls, t.fluid)
    for desc, pts in t.bodies:
        hs.view.add_rigid_body(N.rigid_body(**desc), pts)
    b, i, s = open_boundary_surfaces(t)
    hs.set_open_boundaries(b, i, s, t.inflow)
    if learned:
        from . import learned as L
        L.attach(hs.view)
    return hs


@dataclass
class WindowTank:
    """the window configuration of a tank: the region (a cuboid about x = cx,
    half length `half`), its x walls as a kinematic body (plates at cx +- plate,
    body frame), the fluid inside, the background buffer with its rest
    surface `level`, and the box (kinematic, moved by the caller)"""
    tank: WaveTank
    cx: float
    half: float
    plate: float
    plates: np.ndarray
    fluid: np.ndarray
    background: np.ndarray
    level: float
    box_side: float


def window_tank(t: WaveTank, window_length: float = 2.0, background: np.ndarray = None) -> WindowTank:
    """The window of `window_length` about the box of tank t.  The background
    buffer is a fluid at rest filling the basin (the reference loads one from
    a file made by letting a fluid settle): `background`, e.g. the tank's
    fluid after it has settled, or by default the tank's water lattice (at
    0.8 rho0 it settles lower than it stands, so windows loading from it gain
    fluid).  The plates stand at the window's ends, the region reaches 2 r
    beyond them (so the fluid interior, 3 r inside the region, reaches 1 r
    inside the plates), the initial fluid is the background more than 1.5 r
    inside the plates, and the rest surface is the background's top."""
    r = t.particle_radius
    d = 2.0 * r
    bg = _basin_lattice(t) if background is None else np.asarray(background, np.float64)
    top = np.sort(bg[:, 1])[-max(1, len(bg) // 1000)]          # the top, less splash
    level = float(top) + r
    cx = float(t.box["position"][0])
    plate = 0.5 * window_length
    ys = np.linspace(-0.5 * t.height, 0.5 * t.height, int(round(t.height / d)) + 1)
    zs = np.linspace(-0.5 * t.width, 0.5 * t.width, int(round(t.width / d)) + 1)
    Y, Z = np.meshgrid(ys, zs, indexing="ij")
    plates = np.concatenate([np.stack([np.full(Y.size, sx * plate), Y.ravel(), Z.ravel()], 1) for sx in (-1, 1)])
    side = float(np.ptp(t.bodies[1][1][:, 0]))      # the box's edge
    c = np.array([cx, level, 0.5 * t.width])
    fl = bg[(np.abs(bg[:, 0] - cx) < plate - 1.5 * r) & (np.abs(bg - c).max(1) >= 0.5 * side + 1.5 * r)]
    return WindowTank(tank=t, cx=cx, half=plate + d, plate=plate, plates=plates.astype(np.float32),
                      fluid=fl.astype(np.float32), background=bg.astype(np.float32), level=level, box_side=side)


def _basin_lattice(t: WaveTank) -> np.ndarray:
    """the water lattice of build() over the whole basin (no jitter, no hole for the box)"""
    r = t.particle_radius
    d = 2.0 * r
    px = 0.15
    counts = (int(np.floor((t.length - d - (px + 1.5 * d)) / d + 1e-6)) + 1, int(round(t.depth / d)),
              int(np.floor(t.width / d - 2.0 + 1e-6)) + 1)
    x0 = t.length - d - (counts[0] - 1) * d
    return fluid_block((x0, d, d), counts, d)


def settle_basin(t: WaveTank, params: N.DFSPHParams, seconds: float = 1.5) -> np.ndarray:
    """a background buffer for window_tank(): the basin's water lattice (no
    piston, no box) let settle for `seconds`; its positions"""
    hs = N.HostSolver(params, t.walls, _basin_lattice(t).astype(np.float32))
    time = 0.0
    while time < seconds:
        time = hs.step().time
    x = hs.view.download(N.FIELD_POS)
    hs.close()
    return x


def make_window_solver(w: WindowTank, params: N.DFSPHParams, learned: bool = False) -> N.HostSolver:
    """SPH::DFSPHCUDA over the window: the basin walls, the window's fluid, the
    box (body 0, kinematic: the caller sets its place each step) and the
    plates (body 1), the dynamic window tracking body 0 with body 1 as its
    walls (no piston, no absorbing end: the window's walls confine the fluid)"""
    t = w.tank
    hs = N.HostSolver(params, t.walls, w.fluid)
    h = 0.5 * w.box_side
    cube = box_surface((-h, -h, -h), (h, h, h), 2.0 * t.particle_radius).astype(np.float32)
    hs.view.add_rigid_body(N.rigid_body(type=N.BODY_KINEMATIC, position=(w.cx, w.level, 0.5 * t.width)), cube)
    hs.view.add_rigid_body(N.rigid_body(type=N.BODY_KINEMATIC, position=(w.cx, 0.5 * t.height, 0.5 * t.width)),
                           w.plates)
    region = N.surface("cuboid", center=(w.cx, 0.5 * t.height, 0.5 * t.width), half=(w.half, t.height, t.width))
    fluid_surf = N.surface("plane", o=(0.0, w.level, 0.0), n=(0.0, -1.0, 0.0))
    hs.set_dynamic_window(region, fluid_surf, w.background, tracked_body=0, wall_body=1)
    if learned:
        from . import learned as L
        L.attach(hs.view)
    return hs
