"""Scene files and text particle files on the CPU (libdfsph_host.so, no GPU call).

Expected values are worked out by hand from the reference's rules:
  * SceneLoader::readScene defaults and keys (Utilities/SceneLoader.cpp:37-276),
    passed to the solver as DemoBase does (Demos/Common/DemoBase.cpp:444-450);
  * createFluidBlocks (Demos/Common/DemoBase.cpp:539-600): steps per axis
    round(diff / shift) - 1, first particle at min + 2r;
  * UnifiedParticleSet::write_to_file / load_from_file
    (SPlisHSPlasH/DFSPH/DFSPH_c_arrays_structure.cpp:424-560).
The scene and mesh files are written here; none is taken from the reference.
"""
import json
import math

import numpy as np
import pytest

from splishsplash_for_cuda_amd import _native
from splishsplash_for_cuda_amd._native import DFSPHError, Scene, read_particles_text, write_particles_text

R = 0.025
D = 2 * R

BOX_OBJ = """# unit cube
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
f 1 2 3
f 1 3 4
f 5 6 7
f 5 7 8
f 1/1 2/2 6/3
f 1 6 5
f 2 3 7
f 2 7 6
f 3 4 8
f 3 8 7
f 4 1 5 8
"""

# unit cube with outward-facing quads (mass properties need a consistent winding)
CUBE_OBJ = "\n".join(["v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0", "v 0 0 1", "v 1 0 1", "v 1 1 1", "v 0 1 1",
                       "f 1 4 3 2", "f 5 6 7 8", "f 1 2 6 5", "f 4 8 7 3", "f 1 5 8 4", "f 2 3 7 6"]) + "\n"

TRIANGLE_OBJ = """v 0 0 0
v 0.5 0 0
v 0 0 0.3
f 1 2 3
"""


def _scene(tmp_path, doc, name="scene.json"):
    (tmp_path / "box.obj").write_text(BOX_OBJ)
    (tmp_path / "tri.obj").write_text(TRIANGLE_OBJ)
    (tmp_path / "cube.obj").write_text(CUBE_OBJ)
    p = tmp_path / name
    p.write_text(json.dumps(doc) if not isinstance(doc, str) else doc)
    return Scene(str(p))


def _block(mode, start=(0, 0, 0), end=(1, 1, 1), **kw):
    b = {"denseMode": mode, "start": list(start), "end": list(end)}
    b.update(kw)
    return b


def test_configuration_defaults(tmp_path):
    p = _scene(tmp_path, {"FluidBlocks": [_block(0)]}).params()
    assert p.particle_radius == 0.025 and p.dt == 0.001
    assert p.use_cfl == 1 and p.cfl_factor == 0.5 and p.dt_max == 0.005
    assert p.max_iter == 100 and p.max_error == 0.01
    assert p.max_iter_v == 100 and p.max_error_v == 0.1
    assert p.viscosity == 0.02 and p.density0 == 1000.0
    assert tuple(p.gravity) == (0.0, -9.81, 0.0) and p.enable_divergence == 1
    d = _native.default_params()
    assert p.min_iter_v == d.min_iter_v and p.max_neighbors == d.max_neighbors   # untouched by scenes


def test_configuration_keys(tmp_path):
    cfg = {"timeStepSize": 0.002, "particleRadius": 0.02, "cflMethod": 0, "cflFactor": 0.7,
           "cflMaxTimeStepSize": 0.004, "maxIterations": 50, "maxError": 0.05, "maxIterationsV": 20,
           "maxErrorV": 0.2, "viscosity": 0.01, "density0": 998.0, "gravitation": [0, 0, -9.8],
           "enableDivergenceSolver": False, "numberOfStepsPerRenderUpdate": 4, "pauseAt": -1}
    p = _scene(tmp_path, {"Configuration": cfg, "FluidBlocks": [_block(0)]}).params()
    assert (p.dt, p.particle_radius, p.use_cfl, p.cfl_factor, p.dt_max) == (0.002, 0.02, 0, 0.7, 0.004)
    assert (p.max_iter, p.max_error, p.max_iter_v, p.max_error_v) == (50, 0.05, 20, 0.2)
    assert (p.viscosity, p.density0, tuple(p.gravity), p.enable_divergence) == (0.01, 998.0, (0.0, 0.0, -9.8), 0)


@pytest.mark.parametrize("mode,steps", [
    (0, (19, 19, 19)),     # 1 / 0.05 = 20 per axis
    (1, (18, 22, 18)),     # x, z: 0.95 / 0.05 = 19; y: 1 / (sqrt(3) r) = 23.09
    (2, (22, 22, 18)),     # x: (1 - 0.04082) / 0.04082 = 23.49; z: 19
])
def test_fluid_block_counts(tmp_path, mode, steps):
    x, v = _scene(tmp_path, {"FluidBlocks": [_block(mode, initialVelocity=[1, 2, 3])]}).fluid()
    assert len(x) == steps[0] * steps[1] * steps[2]
    assert (v == np.float32([1, 2, 3])).all()
    assert np.allclose(x.min(0), 0.05, atol=1e-7)           # first layer at min + 2r


def test_fluid_block_mode0_lattice(tmp_path):
    x, _ = _scene(tmp_path, {"FluidBlocks": [_block(0)]}).fluid()
    g = np.arange(19) * D + D
    want = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)   # x outer, z inner
    assert np.abs(x - want).max() < 1e-6


def test_fluid_block_spacing(tmp_path):
    """Nearest-neighbor distance: 2r for the cubic (0) and layered (1) packings;
    for mode 2 the staggered columns sit sqrt(6) 2r / 3 apart in x and r apart
    in z, i.e. r sqrt(11 / 3)."""
    from scipy.spatial import cKDTree
    for mode, dmin in ((0, D), (1, D), (2, R * math.sqrt(11.0 / 3.0))):
        x, _ = _scene(tmp_path, {"FluidBlocks": [_block(mode)]}, "m%d.json" % mode).fluid()
        d = cKDTree(x.astype(np.float64)).query(x, k=2)[0][:, 1]
        assert abs(d.min() - dmin) < 1e-6 and abs(d.max() - dmin) < 1e-6, (mode, d.min(), d.max())


def test_fluid_block_translation_and_scale(tmp_path):
    b = _block(0, start=(0, 0, 0), end=(1, 0.5, 1), translation=[1, 0, -1], scale=[2, 1, 1])
    x, _ = _scene(tmp_path, {"FluidBlocks": [b]}).fluid()
    # box x in [1, 3], y in [0, 0.5], z in [-1, 0]
    assert len(x) == 39 * 9 * 19
    assert np.allclose(x.min(0), [1.05, 0.05, -0.95], atol=1e-6)
    assert np.allclose(x.max(0), [1.05 + 38 * D, 0.05 + 8 * D, -0.95 + 18 * D], atol=1e-6)


def test_several_blocks_concatenate(tmp_path):
    sc = _scene(tmp_path, {"FluidBlocks": [_block(0), _block(0, start=(2, 0, 0), end=(3, 1, 1))]})
    x, _ = sc.fluid()
    assert len(x) == 2 * 19 ** 3
    assert x[: 19 ** 3, 0].max() < 1 and x[19 ** 3:, 0].min() > 2


def test_box_rigid_body_surface_lattice(tmp_path):
    sc = _scene(tmp_path, {"RigidBodies": [{"geometryFile": "box.obj"}], "FluidBlocks": [_block(0)]})
    b = sc.boundary()
    assert len(b) == 21 ** 3 - 19 ** 3           # lattice points of spacing 2r on the six faces
    on_face = np.isclose(b, 0, atol=1e-6) | np.isclose(b, 1, atol=1e-6)
    assert on_face.any(1).all()
    assert len(np.unique(np.round(b / R).astype(np.int64), axis=0)) == len(b)


def test_rigid_body_scale_rotation_translation(tmp_path):
    # scale x by 2, then 90 degrees about y (x -> -z, z -> x), then translate
    rb = {"geometryFile": "box.obj", "scale": [2, 1, 1], "rotationAxis": [0, 1, 0],
          "rotationAngle": math.pi / 2, "translation": [5, 0, 0]}
    b = _scene(tmp_path, {"RigidBodies": [rb], "FluidBlocks": [_block(0)]}).boundary()
    assert len(b) == 41 * 21 * 21 - 39 * 19 * 19
    assert np.allclose(b.min(0), [5, 0, -2], atol=1e-5) and np.allclose(b.max(0), [6, 1, 0], atol=1e-5)


def test_triangle_mesh_sampling_covers_the_surface(tmp_path):
    b = _scene(tmp_path, {"RigidBodies": [{"geometryFile": "tri.obj"}], "FluidBlocks": [_block(0)]}).boundary()
    assert np.abs(b[:, 1]).max() < 1e-7                    # in the triangle's plane
    assert (b[:, 0] >= -1e-7).all() and (b[:, 2] >= -1e-7).all()
    assert (b[:, 0] / 0.5 + b[:, 2] / 0.3 <= 1 + 1e-6).all()
    rng = np.random.default_rng(0)
    u, w = rng.random((2, 4000))
    keep = u + w <= 1
    pts = np.stack([0.5 * u[keep], np.zeros(keep.sum()), 0.3 * w[keep]], 1)
    d = np.sqrt(((pts[:, None, :] - b[None, :, :]) ** 2).sum(-1)).min(1)
    assert d.max() <= D


def test_scene_errors(tmp_path):
    with pytest.raises(DFSPHError):
        Scene(str(tmp_path / "missing.json"))
    with pytest.raises(DFSPHError, match="JSON"):
        _scene(tmp_path, '{"Configuration": {"particleRadius": 0.025,}')
    with pytest.raises(DFSPHError, match="cannot open"):
        _scene(tmp_path, {"RigidBodies": [{"geometryFile": "nothing.obj"}]})


def test_json_syntax(tmp_path):
    doc = ('{ "Configuration" : { "particleRadius" : 2.5e-2 , "timeStepSize": 1E-3, "name": "a \\"b\\" \\u00e9",'
           ' "list": [true, false, null, [], {}] },\n "FluidBlocks": [ {"denseMode": 0, "start": [0,0,0],'
           ' "end": [1, 1, 1] } ] }')
    sc = _scene(tmp_path, doc)
    assert sc.params().particle_radius == 0.025 and len(sc.fluid()[0]) == 19 ** 3


def test_particle_text_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n = 1000
    x = rng.normal(size=(n, 3)).astype(np.float32)
    v = rng.normal(size=(n, 3)).astype(np.float32)
    m = rng.random(n).astype(np.float32)
    p = str(tmp_path / "fluid_file.txt")
    write_particles_text(p, x, m, v)
    head = open(p).readline().split()
    assert head == [str(n), "1", "1", "0"]
    x2, v2, m2 = read_particles_text(p)
    assert (x2 == x).all() and (v2 == v).all() and (m2 == m).all()     # bit-exact
    write_particles_text(p, x, m, None, False, False)
    x3, v3, _ = read_particles_text(p)
    assert open(p).readline().split() == [str(n), "0", "0", "0"] and (v3 == 0).all() and (x3 == x).all()


def test_particle_text_reference_layout(tmp_path):
    """A file as the reference writes it (default 6-digit stream output), with
    the rigid-body line of a dynamic object (position, velocity, quaternion,
    angular velocity) that the reader skips."""
    p = tmp_path / "solid.txt"
    p.write_text("2 0 0 1\n1 2 3 0 0 0 0 0 0 1 0 0 0 \n0.125 0.1 0.2 0.3 1 2 3 \n"
                 "0.25 -1.5 2.5e-05 3 0 0 0 \n")
    x, v, m = read_particles_text(str(p))
    assert np.array_equal(m, np.float32([0.125, 0.25]))
    assert np.array_equal(x, np.float32([[0.1, 0.2, 0.3], [-1.5, 2.5e-05, 3]]))
    assert np.array_equal(v, np.float32([[1, 2, 3], [0, 0, 0]]))
    p.write_text("3 1 1 0\n1 0 0 0 0 0 0\n")
    with pytest.raises(DFSPHError, match="truncated"):
        read_particles_text(str(p))
    empty = tmp_path / "empty.txt"
    empty.write_text("0 1 1 0\n")
    x, v, m = read_particles_text(str(empty))
    assert x.shape == (0, 3) and m.shape == (0,)


def _quat_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_dynamic_rigid_body_mass_properties(tmp_path):
    """A 0.2 x 0.4 x 0.6 box at 500 kg/m^3 turned 30 degrees about z: mass
    24 kg, principal moments m (b^2 + c^2) / 12 ... = 1.04, 0.80, 0.40 kg m^2,
    center at the turned box center; its particles, mapped back to the world,
    lie on the box surface."""
    ang = math.radians(30.0)
    rb = {"geometryFile": "cube.obj", "isDynamic": True, "density": 500, "scale": [0.2, 0.4, 0.6],
          "rotationAxis": [0, 0, 1], "rotationAngle": ang, "translation": [0.5, 0.5, 0.5], "velocity": [1, 0, 0]}
    tank = {"geometryFile": "box.obj", "scale": [2, 2, 2], "translation": [-0.5, -0.5, -0.5]}
    sc = _scene(tmp_path, {"RigidBodies": [tank, rb], "FluidBlocks": [_block(0)]})
    assert len(sc.boundary()) == 41 ** 3 - 39 ** 3        # the tank only
    bodies = sc.bodies()
    assert len(bodies) == 1
    b, x = bodies[0]
    assert b.type == 0 and b.mass == pytest.approx(24.0, rel=1e-12)
    assert sorted(b.inertia[:]) == pytest.approx([0.40, 0.80, 1.04], rel=1e-10)
    Rz = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0], [0, 0, 1]])
    c = np.array([0.5, 0.5, 0.5]) + Rz @ np.array([0.1, 0.2, 0.3])
    assert np.allclose(b.position[:], c, atol=1e-12)
    assert tuple(b.velocity[:]) == (1.0, 0.0, 0.0)
    R = _quat_mat(b.rotation[:])
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-12) and np.linalg.det(R) == pytest.approx(1.0)
    world = x.astype(np.float64) @ R.T + c
    local = (world - 0.5) @ Rz                              # back into the unturned box frame
    ext = np.array([0.2, 0.4, 0.6])
    assert (local > -1e-5).all() and (local < ext + 1e-5).all()
    on_face = (np.abs(local) < 1e-5) | (np.abs(local - ext) < 1e-5)
    assert on_face.any(1).all()
    # the inertia tensor from the particles' frame is diagonal in the body frame
    # (principal axes): a box's axes are its edges, so each body axis maps to an edge
    for k in range(3):
        axis_local = Rz.T @ R[:, k]
        assert np.isclose(np.abs(axis_local).max(), 1.0, atol=1e-9)
