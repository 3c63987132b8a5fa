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
               "dfsph_step_stats": _native.DFSPHStepStats, "dfsph_rigid_body": _native.DFSPHRigidBody}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dfsph_hip.h"', "int main(void) {"]
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
