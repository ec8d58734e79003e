"""Scene files (include/pbf_scene.hpp, pbf_amd/scene_file.py): the C++ reader
behind pbf_run and the Python reader give the same particles and parameter
schedules bit for bit, and the shipped scene files restate the workloads of
pbf_amd.scene exactly. No GPU: pbf_run's --dump-initial and --print-params
modes stop before any device call."""
import os
import subprocess

import numpy as np
import pytest

from pbf_amd import scene as S
from pbf_amd import scene_file as F
from pbf_amd import sph as P

SCENES = sorted(f for f in os.listdir(F.SCENES) if f.endswith(".scene"))


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


def cli(*args):
    assert os.path.exists(F.CLI), "pbf_run not built (python -m pbf_amd.build)"
    r = subprocess.run([F.CLI, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def params_line(step, p):
    h = lambda x: float(np.float32(x)).hex()
    return (f"step {step} one_over_rho_0 {h(p.one_over_rho_0)} epsilon {h(p.epsilon)} gravity {h(p.gravity)} "
            f"timestep {h(p.timestep)} tensile_k {h(p.tensile_instability_k)} "
            f"tensile_scale {h(p.tensile_instability_scale)} xsph {h(p.xsph_viscosity_c)} "
            f"vorticity_epsilon {h(p.vorticity_epsilon)} iterations {p.iterations} vorticity {p.vorticity} "
            f"extforce {p.extforce} h {h(p.h)}")


def c_hex(line):
    """C's %a for floats to Python's float.hex() spelling (same value)."""
    out = []
    for w in line.split():
        if w.startswith("0x") or w.startswith("-0x"):
            w = float.fromhex(w).hex()
        out.append(w)
    return " ".join(out)


@pytest.mark.parametrize("name", SCENES)
def test_cpp_and_python_readers_agree(name, tmp_path):
    path = os.path.join(F.SCENES, name)
    sf = F.load(path)
    dump = tmp_path / "init.f32"
    cli(path, "--dump-initial", str(dump))
    pos, vel = F.read_state(str(dump), sf.n)
    assert np.array_equal(bits(pos), bits(sf.pos)), f"{name}: positions differ"
    if name == "adjoint_16k.scene":
        # normal velocities go through log, cos and sin of each library's
        # libm; equal up to their last bits
        np.testing.assert_allclose(vel, sf.vel, rtol=2e-7, atol=1e-12)
    else:
        assert np.array_equal(bits(vel), bits(sf.vel)), f"{name}: velocities differ"
    # parameter block and every scheduled change
    got = [c_hex(l) for l in cli(path, "--print-params").strip().splitlines()]
    p = F.copy_params(sf.params)
    want = [params_line(0, p)]
    for a in sf.actions:
        F.apply(p, a)
        want.append(params_line(a.step, p))
    assert got == [c_hex(w) for w in want]


@pytest.mark.parametrize("name,make", [("dam_32k.scene", S.cfg_dam_32k), ("block_1m.scene", S.cfg_block_1m),
                                       ("adjoint_16k.scene", S.cfg_adjoint_16k)])
def test_scene_files_restate_the_workloads(name, make):
    sf = F.load(os.path.join(F.SCENES, name))
    sc = make()
    if name == "adjoint_16k.scene":
        sc.params.epsilon = 5.0
    assert sf.grid == tuple(sc.grid) and sf.n == sc.n and sf.steps == sc.steps
    assert np.array_equal(bits(sf.pos), bits(sc.pos))
    assert np.array_equal(bits(sf.vel), bits(sc.vel))
    want = P.scene_params(sc)
    for f, _ in type(want)._fields_:
        a, b = getattr(sf.params, f), getattr(want, f)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), f
        else:
            assert a == b, f


def test_reference_scene_pattern():
    """reference_default.scene: two 32^3 blocks from opposite corners at spacing
    0.94 (src/Simulation.cpp:214-246) with the reference's defaults."""
    sf = F.load(os.path.join(F.SCENES, "reference_default.scene"))
    assert sf.n == 65536 and sf.grid == (128, 64, 128)
    a, b = sf.pos[:32768, :3].astype(np.float64), sf.pos[32768:, :3].astype(np.float64)
    assert np.all(np.abs(a[0] - (32.5, 0.5, 32.5)) <= 0.005 + 1e-6)
    assert np.all(np.abs(b[0] - (95.5, 0.5, 95.5)) <= 0.005 + 1e-6)
    # id order x outer, z, y fastest; the second block mirrored in x and z
    assert abs((a[1] - a[0])[1] - 0.94) < 0.011 and abs((b[32] - b[0])[2] + 0.94) < 0.011
    assert (sf.params.h, sf.params.iterations, sf.params.vorticity, sf.params.epsilon) == (2.0, 5, 0, 5.0)


def test_bad_scene_lines_are_rejected(tmp_path):
    for text, msg in [("grid 64 64\n", "grid"), ("blok 1 2 3\n", "unknown directive"),
                      ("at 3 adjust nothing 1\n", "unknown parameter"), ("gravity x\n", "not a number")]:
        with pytest.raises(ValueError, match=msg):
            F.parse(text)
        p = tmp_path / "bad.scene"
        p.write_text(text)
        r = subprocess.run([F.CLI, str(p), "--print-params"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True)
        assert r.returncode == 1 and msg in r.stderr
