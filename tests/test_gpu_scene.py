"""Scene files and the reference's text particle files driving the solver on
the GPU (libdfsph_host.so over libdfsph_hip.so).

  * a solver built from a scene file runs bit-identically to one built from
    the same arrays (SceneLoader output -> DFSPHCData constructor);
  * handleSimulationSave / handleSimulationLoad on text files
    (DFSPH_CUDA.cpp:1887-1979): the files hold the device state exactly
    (fluid rows in device order, boundary mass column = psi), a solver made
    from them holds the same particles, psi and time step, and runs within the
    solver tolerances.
"""
import json
import os

import numpy as np
import pytest

import splishsplash_for_cuda_amd as S

pytestmark = pytest.mark.gpu

BOX_OBJ = "\n".join(["v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0", "v 0 0 1", "v 1 0 1", "v 1 1 1", "v 0 1 1",
                     "f 1 2 3 4", "f 5 6 7 8", "f 1 2 6 5", "f 2 3 7 6", "f 3 4 8 7", "f 4 1 5 8"]) + "\n"


def _scene(tmp_path):
    (tmp_path / "tank.obj").write_text(BOX_OBJ)
    doc = {"Configuration": {"particleRadius": 0.025, "timeStepSize": 0.002, "cflMethod": 1,
                             "cflMaxTimeStepSize": 0.004, "maxError": 0.05, "maxErrorV": 0.1},
           "RigidBodies": [{"geometryFile": "tank.obj", "scale": [1.2, 1.0, 0.6]}],
           "FluidBlocks": [{"denseMode": 0, "start": [0, 0, 0], "end": [0.6, 0.8, 0.6]}]}
    p = tmp_path / "dam.json"
    p.write_text(json.dumps(doc))
    return S.Scene(str(p))


def _within_tolerance(st, p):
    """Final errors against the percent targets: the density error is an
    average density deviation (kg/m^3), the divergence error an average
    density rate (kg/m^3/s) taken over the step's h."""
    return (st.err / p.density0 * 100.0 <= p.max_error * 1.0001
            and st.err_v * st.h / p.density0 * 100.0 <= p.max_error_v * 1.0001)


def _lexsorted(x):
    return x[np.lexsort(x.T[::-1])]


def test_scene_solver_equals_array_solver(tmp_path):
    sc = _scene(tmp_path)
    fx, fv = sc.fluid()
    bx = sc.boundary()
    assert len(fx) == 11 * 15 * 11
    print("scene: %d fluid, %d boundary particles" % (len(fx), len(bx)), flush=True)
    a = S.HostSolver.from_scene(sc)
    b = S.HostSolver(sc.params(), bx, fx, fv)
    for _ in range(10):
        sa, sb = a.step(), b.step()
        assert (sa.iters_v, sa.iters, sa.h) == (sb.iters_v, sb.iters, sb.h)
        assert _within_tolerance(sa, sc.params())
        assert sa.n_fluid == len(fx)
    xa, xb = a.view.state_by_id(), b.view.state_by_id()
    for k in ("pos", "vel", "kappa", "kappaV"):
        assert np.array_equal(xa[k], xb[k]), k
    a.close()
    b.close()


def test_text_files_save_and_load(tmp_path):
    sc = _scene(tmp_path)
    p = sc.params()
    a = S.HostSolver.from_scene(sc)
    for _ in range(10):
        a.step()
    out = tmp_path / "fluid_data"
    out.mkdir()
    a.save_text(str(out))
    info = a.view.info()
    h_saved = float(open(out / "general_data.txt").read())
    assert h_saved == info.h

    # the files hold the device state exactly
    x, v, m = S.read_particles_text(str(out / "fluid_file.txt"))
    assert np.array_equal(x, a.view.download(S.FIELD_POS))
    assert np.array_equal(v, a.view.download(S.FIELD_VEL))
    assert np.array_equal(m, a.view.download(S.FIELD_MASS))
    assert open(out / "fluid_file.txt").readline().split()[1:] == ["1", "1", "0"]
    bx, _, bpsi = S.read_particles_text(str(out / "boundaries_file.txt"))
    assert np.array_equal(bx, a.view.download(S.FIELD_BOUNDARY_POS))
    assert np.array_equal(bpsi, a.view.download(S.FIELD_BOUNDARY_PSI))

    # a solver made from the files: same particles, same psi (not recomputed), same h
    b = S.HostSolver.from_text(p, str(out))
    assert b.view.info().h == h_saved
    assert np.array_equal(b.view.state_by_id()["pos"], x)        # ids = file rows
    key_a = np.concatenate([bx, bpsi[:, None]], 1)
    key_b = np.concatenate([b.view.download(S.FIELD_BOUNDARY_POS), b.view.download(S.FIELD_BOUNDARY_PSI)[:, None]], 1)
    assert np.array_equal(_lexsorted(key_a), _lexsorted(key_b))
    for _ in range(5):
        st = b.step()
        assert _within_tolerance(st, p)
    assert np.isfinite(b.view.download(S.FIELD_POS)).all()

    # loading into a running solver: fluid replaced, time kept, h from the file
    for _ in range(3):
        a.step()
    t = a.view.info().time
    a.load_text(str(out), load_velocities=False)
    inf = a.view.info()
    assert inf.time == t and inf.h == inf.h_past == inf.h_future == h_saved
    st = a.view.state_by_id()
    assert np.array_equal(st["pos"], x) and (st["vel"] == 0).all()
    a.step()
    a.close()
    b.close()
    assert sorted(os.listdir(out)) == ["boundaries_file.txt", "fluid_file.txt", "general_data.txt"]


CUBE_OBJ = "\n".join(["v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0", "v 0 0 1", "v 1 0 1", "v 1 1 1", "v 0 1 1",
                      "f 1 4 3 2", "f 5 6 7 8", "f 1 2 6 5", "f 4 8 7 3", "f 1 5 8 4", "f 2 3 7 6"]) + "\n"


def test_scene_with_a_dynamic_body(tmp_path):
    """A 0.2 m cube (isDynamic, 500 kg/m^3) dropped onto the dam's water
    falls, is held up by the fluid and carries its force; the solver comes
    from the scene.  (Bodies meet only the fluid: contacts with walls or
    other bodies are not modelled.)"""
    (tmp_path / "tank.obj").write_text(BOX_OBJ)
    (tmp_path / "cube.obj").write_text(CUBE_OBJ)
    doc = {"Configuration": {"particleRadius": 0.025, "timeStepSize": 0.002, "cflMethod": 1,
                             "cflMaxTimeStepSize": 0.004, "maxError": 0.05, "maxErrorV": 0.1},
           "RigidBodies": [{"geometryFile": "tank.obj", "scale": [1.2, 1.0, 0.6]},
                           {"geometryFile": "cube.obj", "isDynamic": True, "density": 500,
                            "scale": [0.2, 0.2, 0.2], "translation": [0.2, 0.55, 0.2]}],
           "FluidBlocks": [{"denseMode": 0, "start": [0, 0, 0], "end": [0.6, 0.5, 0.6]}]}
    p = tmp_path / "dam_cube.json"
    p.write_text(json.dumps(doc))
    sc = S.Scene(str(p))
    (b, x), = sc.bodies()
    assert len(x) == 5 ** 3 - 3 ** 3 and b.mass == pytest.approx(4.0)
    host = S.HostSolver.from_scene(sc)
    assert host.view.rigid_body_count() == 1
    ys, fy = [], []
    for _ in range(150):
        st = host.step()
        assert _within_tolerance(st, sc.params())
        body, F, T = host.view.rigid_body(0)
        ys.append(body.position[1])
        fy.append(F[1])
    print("cube center y %.3f -> %.3f (min %.3f), last fluid force %s" % (0.65, ys[-1], min(ys), F))
    assert ys[-1] < 0.65 - 0.05                # it fell onto the water
    assert min(ys) > 0.15                      # and floats on it
    assert max(fy) > 10.0                      # the fluid holds it up
    assert np.isfinite(host.view.download(S.FIELD_POS)).all()
    host.close()
