"""CPU checks of the C ABI boundary: the libraries load and export every
function that include/dfsph_hip.h and include/dfsph_host.h declare (no GPU call)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dfsph_\w+)\s*\(", text)))


def test_hip_library_exports_every_declared_function():
    from splishsplash_for_cuda_amd import _native
    lib = _native.load_hip()
    names = _declared("dfsph_hip.h")
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
        assert n in _native.HIP_SIGNATURES, n


def test_host_library_exports_every_declared_function():
    from splishsplash_for_cuda_amd import _native
    lib = _native.load_host()
    for n in _declared("dfsph_host.h"):
        assert hasattr(lib, n), n
        assert n in _native.HOST_SIGNATURES, n


def test_default_params_match_reference_defaults():
    from splishsplash_for_cuda_amd import default_params
    p = default_params()
    assert p.min_iter_v == 3 and p.min_iter == 2          # DFSPH_core_cuda.cu:828,1378
    assert p.max_neighbors == 90                           # DFSPH_define_c.h:23
    assert tuple(p.gravity) == (0.0, -9.81, 0.0)          # DFSPH_c_arrays_structure.cpp:778
    assert p.check_positions_mode == 2                     # DFSPH_CUDA.cpp:456


def test_host_layer_has_no_hip_dependency():
    """The C++ host classes build with g++ against the C ABI only."""
    src = open(os.path.join(ROOT, "splishsplash_for_cuda_amd", "csrc", "DFSPH_CUDA.cpp")).read()
    hdr = open(os.path.join(ROOT, "splishsplash_for_cuda_amd", "csrc", "DFSPH_CUDA.h")).read()
    for t in (src, hdr):
        assert "hip_runtime" not in t and "hipMalloc" not in t


def test_scene_sizes():
    from splishsplash_for_cuda_amd import scene
    s = scene.config1()
    assert s.n_fluid == 16000 and s.n_boundary == 18002
    j = s.fluid - (scene.fluid_block((0.05,) * 3, (20, 40, 20), 0.05))
    assert abs(j).max() <= 0.01 * 0.025 * 1.001      # jitter U(-0.01 r, 0.01 r), float32 positions


def test_ctypes_struct_layouts_match_the_header(tmp_path):
    """sizeof / offsetof of the ABI structs, from the C compiler, equal the
    ctypes mirrors (a field added on one side only fails here)."""
    import shutil
    import subprocess
    from splishsplash_for_cuda_amd import _native
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"dfsph_params": _native.DFSPHParams, "dfsph_info": _native.DFSPHInfo,
               "dfsph_step_stats": _native.DFSPHStepStats, "dfsph_rigid_body": _native.DFSPHRigidBody,
               "dfsph_surface": _native.DFSPHSurface, "dfsph_cuda_stats": _native.DFSPHCudaStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dfsph_host.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (name, name))
        for f in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (name, f[0], name, f[0]))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for name, cls in structs.items():
        assert int(got[name]) == ctypes.sizeof(cls), name
        for f in cls._fields_:
            assert int(got["%s.%s" % (name, f[0])]) == getattr(cls, f[0]).offset, (name, f[0])


C_CONSUMER = r"""
#include <stdio.h>
#include <stdlib.h>
#include "dfsph_host.h"
/* a plain C caller of the host library: scene file -> parameters and
 * particles, then the text particle format (no GPU call) */
int main(int argc, char** argv) {
    void* sc = dfsph_scene_read(argv[1], NULL);
    if (!sc) { printf("error %s\n", dfsph_cuda_last_error()); return 1; }
    dfsph_params p;
    dfsph_scene_params(sc, &p);
    uint32_t nf = dfsph_scene_fluid_count(sc), nb = dfsph_scene_boundary_count(sc);
    float* x = (float*)malloc(sizeof(float) * 3 * nf);
    float* v = (float*)malloc(sizeof(float) * 3 * nf);
    float* m = (float*)malloc(sizeof(float) * nf);
    dfsph_scene_fluid(sc, x, v);
    for (uint32_t i = 0; i < nf; ++i) m[i] = 0.1f;
    if (dfsph_particles_write_text(argv[2], x, v, m, nf, 1, 1) != 0) return 2;
    int64_t back = dfsph_particles_read_text(argv[2], NULL, NULL, NULL, 0);
    printf("%g %u %u %lld %g %g\n", p.particle_radius, nf, nb, (long long)back, x[0], x[3 * nf - 1]);
    dfsph_scene_free(sc);
    free(x); free(v); free(m);
    return 0;
}
"""


def test_c_program_uses_the_host_library(tmp_path):
    """dfsph_host.h from plain C, linked against the built libraries."""
    import shutil
    import subprocess
    from splishsplash_for_cuda_amd import _native
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    _native.load_host()
    pkg = os.path.join(ROOT, "splishsplash_for_cuda_amd")
    (tmp_path / "c.c").write_text(C_CONSUMER)
    exe = tmp_path / "c"
    subprocess.run([cc, "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), str(tmp_path / "c.c"),
                    "-L", pkg, "-ldfsph_host", "-ldfsph_hip", "-Wl,-rpath," + pkg, "-o", str(exe)], check=True)
    (tmp_path / "box.obj").write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0 0 1\nv 1 0 1\nv 1 1 1\nv 0 1 1\n"
                                      "f 1 4 3 2\nf 5 6 7 8\nf 1 2 6 5\nf 4 8 7 3\nf 1 5 8 4\nf 2 3 7 6\n")
    (tmp_path / "s.json").write_text('{"Configuration": {"particleRadius": 0.025}, "RigidBodies": '
                                     '[{"geometryFile": "box.obj"}], "FluidBlocks": [{"denseMode": 0, '
                                     '"start": [0, 0, 0], "end": [0.5, 0.5, 0.5]}]}')
    out = subprocess.run([str(exe), str(tmp_path / "s.json"), str(tmp_path / "f.txt")], check=True,
                         capture_output=True, text=True).stdout.split()
    r, nf, nb, back, x0, xl = float(out[0]), int(out[1]), int(out[2]), int(out[3]), float(out[4]), float(out[5])
    assert r == 0.025 and nf == 9 ** 3 and nb == 21 ** 3 - 19 ** 3 and back == nf
    assert x0 == pytest.approx(0.05) and xl == pytest.approx(0.05 + 8 * 0.05)
