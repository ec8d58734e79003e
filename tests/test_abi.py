"""The C-ABI library loads without a GPU and exports every entry point include/pbf.h declares."""
import ctypes as C
import os
import re

import numpy as np

from pbf_amd import _lib, scene as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbf_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    names = declared("pbf.h")
    assert len(names) > 40
    for name in names:
        assert hasattr(lib, name), f"{name} declared in include/pbf.h but not exported"
        assert name in _lib.SIGNATURES or name in _lib.OPTIONAL, f"{name} has no ctypes signature"


def test_host_only_entry_points():
    lib = _lib.load()
    # SPH::Wpoly6 of the reference's tensile-instability scale (src/SPH.cpp:142): 1/W(0.2, 2)
    assert np.float32(lib.pbf_wpoly6(0.2, 2.0)) == S.wpoly6(0.2, 2.0)
    assert lib.pbf_wpoly6(1.5, 1.0) == 0.0
    # key bits: 2*ceil(count_sortbits(C-1)/2) for the three grids of the workloads
    assert lib.pbf_key_bits(64, 64, 64) == 18
    assert lib.pbf_key_bits(256, 128, 256) == 24
    assert lib.pbf_key_bits(384, 192, 384) == 26
    assert lib.pbf_key_bits(0, 1, 1) < 0
    p = _lib.PbfParams()
    lib.pbf_params_default(C.byref(p), 128, 64, 128)
    assert (p.one_over_rho_0, p.epsilon, p.gravity, p.iterations, p.h) == (1.0, 5.0, 10.0, 5, 2.0)
    assert list(p.wall_lo) == [16.0, 0.0, 16.0] and list(p.wall_hi) == [112.0, 64.0, 112.0]
    assert list(p.gravity_dir) == [0.0, -1.0, 0.0]


def test_cpp_classes_compile_and_link(tmp_path):
    """include/pbf.hpp compiles with g++ alone (no HIP headers) and the demo
    program built over the library starts without a GPU."""
    import subprocess
    from pbf_amd.build import CPP_DEMO, build_cpp_demo
    src = tmp_path / "use.cpp"
    src.write_text('#include "pbf.hpp"\nint main() { return pbf::SPH::Wpoly6(0.5f, 1.0f) > 0.0f ? 0 : 1; }\n')
    exe = str(tmp_path / "use")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", exe,
                    "-L", os.path.dirname(_lib.LIB_PATH), "-l:libpbf_amd.so",
                    "-Wl,-rpath," + os.path.dirname(_lib.LIB_PATH)], check=True)
    assert subprocess.run([exe]).returncode == 0
    if not os.path.exists(CPP_DEMO):
        build_cpp_demo()
    res = subprocess.run([CPP_DEMO], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert res.returncode == 2 and "usage" in res.stdout


def test_noalloc_stage_rejects_sizes_beyond_its_scratch_without_a_gpu():
    """The size check of pbf_stage_vorticity_noalloc comes before any HIP call."""
    lib = _lib.load()
    p = _lib.PbfParams()
    lib.pbf_params_default(C.byref(p), 128, 64, 128)
    g3 = (C.c_int32 * 3)(128, 64, 128)
    dummy = C.c_void_p(256)  # never dereferenced
    assert lib.pbf_stage_vorticity_noalloc(131072 + 256, g3, C.byref(p), dummy, dummy, dummy, dummy, None) == -1
    assert b"PBF_NOALLOC_MAX_N" in lib.pbf_last_error()
