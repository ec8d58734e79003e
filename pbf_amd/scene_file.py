"""Scene files: the Python reader and runner of the format of include/pbf_scene.hpp.

A scene file states what the reference's application sets up by hand: the
particles of ``Simulation::ResetParticleBuffer`` (src/Simulation.cpp:206-273,
here seeded), the parameter block of ``SPH`` (src/SPH.cpp:137-144 defaults,
src/SPH.h setters) and the keyboard changes of ``Simulation::OnKeyDown`` /
``OnKeyUp`` (src/Simulation.cpp:276-300, 363-423) scheduled at given steps.
The format and its arithmetic are those of the C++ reader used by
``pbf_amd/pbf_run``; both give the same particles and parameters bit for bit.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import scene as S

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = os.path.join(HERE, "scenes")
CLI = os.path.join(HERE, "pbf_run")

SETTERS = ("rest_density", "cfm_epsilon", "gravity", "timestep", "solver_iterations", "tensile_instability_k",
           "tensile_instability_scale", "xsph_viscosity", "vorticity_epsilon")

f32 = np.float32


@dataclass
class Action:
    step: int
    kind: str            # "adjust", "set", "external_force", "vorticity", "toggle_vorticity"
    param: str = ""
    value: float = 0.0


@dataclass
class SceneFile:
    name: str
    grid: tuple
    params: _lib.PbfParams
    pos: np.ndarray      # (n, 4) float32, id order
    vel: np.ndarray
    steps: int
    actions: list = field(default_factory=list)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


def _fail(line, msg):
    raise ValueError(f"scene line {line}: {msg}")


def _num(s, line):
    try:
        return float(s)
    except ValueError:
        _fail(line, f"not a number: '{s}'")


def _int(s, line):
    try:
        return int(s, 10)
    except ValueError:
        _fail(line, f"not an integer: '{s}'")


def _flag(s, line):
    if s in ("on", "true", "1"):
        return True
    if s in ("off", "false", "0"):
        return False
    _fail(line, f"expected on or off, got '{s}'")


def _after(t, key, count, line):
    if key not in t or len(t) - t.index(key) < count + 1:
        _fail(line, f"missing '{key}' and its values")
    k = t.index(key)
    return t[k + 1:k + 1 + count]


def set_param(p: _lib.PbfParams, name: str, v: float, line: int = 0):
    """The named setter of ``SPH`` (src/SPH.h; SetRestDensity stores 1/rho, src/SPH.cpp:166-169)."""
    if name == "rest_density":
        p.one_over_rho_0 = float(f32(1.0) / f32(v))
    elif name == "cfm_epsilon":
        p.epsilon = v
    elif name == "gravity":
        p.gravity = v
    elif name == "timestep":
        p.timestep = v
    elif name == "solver_iterations":
        p.iterations = int(v)
    elif name == "tensile_instability_k":
        p.tensile_instability_k = v
    elif name == "tensile_instability_scale":
        p.tensile_instability_scale = v
    elif name == "xsph_viscosity":
        p.xsph_viscosity_c = v
    elif name == "vorticity_epsilon":
        p.vorticity_epsilon = v
    else:
        _fail(line, f"unknown parameter '{name}'")


def adjust(p: _lib.PbfParams, name: str, factor: float):
    """One +/- key press on the selected parameter (src/Simulation.cpp:382-421):
    the reference's clamps, float or double arithmetic as there."""
    fa = f32(factor)
    if name == "rest_density":
        rho = f32(1.0) / f32(p.one_over_rho_0)
        v = f32(max(float(rho) + float(fa) * 0.01, 0.01))
        p.one_over_rho_0 = float(f32(1.0) / v)
    elif name == "cfm_epsilon":
        p.epsilon = float(max(f32(p.epsilon) + fa, f32(0.01)))
    elif name == "gravity":
        p.gravity = float(f32(p.gravity) + fa)
    elif name == "timestep":
        p.timestep = float(f32(max(float(f32(p.timestep)) + 0.001 * float(fa), 0.001)))
    elif name == "solver_iterations":
        iters = int(f32(p.iterations) + fa)
        p.iterations = max(iters, 1)
    elif name == "tensile_instability_k":
        p.tensile_instability_k = float(f32(max(float(f32(p.tensile_instability_k)) + float(fa) * 0.1, 0.1)))
    elif name == "tensile_instability_scale":
        p.tensile_instability_scale = float(
            f32(max(float(f32(p.tensile_instability_scale)) + float(fa) * 0.1, 0.1)))
    elif name == "xsph_viscosity":
        p.xsph_viscosity_c = float(f32(max(float(f32(p.xsph_viscosity_c)) + float(fa) * 0.01, 0.01)))
    elif name == "vorticity_epsilon":
        p.vorticity_epsilon = float(f32(max(float(f32(p.vorticity_epsilon)) + float(fa) * 0.1, 0.1)))
    else:
        raise ValueError(f"unknown parameter '{name}'")


def apply(p: _lib.PbfParams, a: Action):
    if a.kind == "adjust":
        adjust(p, a.param, a.value)
    elif a.kind == "set":
        set_param(p, a.param, a.value)
    elif a.kind == "external_force":
        p.extforce = int(a.value != 0.0)
    elif a.kind == "vorticity":
        p.vorticity = int(a.value != 0.0)
    elif a.kind == "toggle_vorticity":
        p.vorticity = int(not p.vorticity)


def copy_params(p: _lib.PbfParams) -> _lib.PbfParams:
    q = _lib.PbfParams()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(p))
    return q


def parse(text: str, name: str = "scene") -> SceneFile:
    lines = []
    for no, raw in enumerate(text.splitlines(), 1):
        t = raw.split("#", 1)[0].split()
        if t:
            lines.append((no, t))
    grid = (128, 64, 128)
    h = None
    for no, t in lines:
        if t[0] == "grid":
            if len(t) != 4:
                _fail(no, "grid takes three integers")
            grid = tuple(_int(x, no) for x in t[1:4])
        elif t[0] == "h":
            if len(t) != 2:
                _fail(no, "h takes one value")
            h = float(f32(_num(t[1], no)))
    p = _lib.PbfParams()
    _lib.load().pbf_params_default(C.byref(p), *grid)
    if h is not None and h > 0.0:
        p.h = h
    h = float(p.h)
    pos = []
    velocity = ("zero", 0.0, 0)
    steps = 1
    actions = []
    for no, t in lines:
        k = t[0]
        if k in ("grid", "h"):
            continue
        if k == "rest_density" and len(t) == 3 and t[1] == "lattice":
            rho = S.lattice_rest_density(_num(t[2], no) * h, h)
            p.one_over_rho_0 = float(f32(1.0 / rho))
        elif k == "tensile_instability_scale" and len(t) == 3 and t[1] == "wpoly6":
            q = _num(t[2], no) * h
            p.tensile_instability_scale = float(f32(1.0) / f32(S.wpoly6(q, h)))
        elif k in SETTERS:
            if len(t) != 2:
                _fail(no, f"{k} takes one value")
            set_param(p, k, _num(t[1], no), no)
        elif k == "vorticity":
            if len(t) != 2:
                _fail(no, "vorticity takes on or off")
            p.vorticity = int(_flag(t[1], no))
        elif k == "external_force":
            if len(t) != 2:
                _fail(no, "external_force takes on or off")
            p.extforce = int(_flag(t[1], no))
        elif k == "walls":
            if len(t) != 7:
                _fail(no, "walls takes three lower bounds and three margins")
            for a in range(3):
                p.wall_lo[a] = _num(t[1 + a], no)
                p.wall_hi[a] = float(f32(grid[a]) - f32(_num(t[4 + a], no)))
        elif k == "gravity_direction":
            if len(t) != 4:
                _fail(no, "gravity_direction takes three values")
            for a in range(3):
                p.gravity_dir[a] = _num(t[1 + a], no)
        elif k == "block":
            dims = tuple(_int(x, no) for x in t[1:4])
            if any(d <= 0 for d in dims):
                _fail(no, "block dimensions must be positive")
            spacing = _num(_after(t, "spacing", 1, no)[0], no)
            corner = tuple(_num(x, no) for x in _after(t, "corner", 3, no))
            jitter = _num(_after(t, "jitter", 1, no)[0], no) if "jitter" in t else 0.0
            seed = _int(_after(t, "seed", 1, no)[0], no) if "seed" in t else 0
            pos.append(S.lattice_block(dims, spacing, corner, jitter, seed, mirror="mirror" in t))
        elif k == "uniform":
            m = _int(t[1], no)
            if m <= 0:
                _fail(no, "uniform count must be positive")
            lo = np.array([_num(x, no) for x in _after(t, "lo", 3, no)])
            hi = np.array([_num(x, no) for x in _after(t, "hi", 3, no)])
            seed = _int(_after(t, "seed", 1, no)[0], no) if "seed" in t else 0
            u = S.uniform(seed, 3 * m).reshape(m, 3)
            b = np.zeros((m, 4), np.float32)
            b[:, :3] = (lo + u * (hi - lo)).astype(np.float32)
            pos.append(b)
        elif k == "velocity":
            if len(t) >= 2 and t[1] == "zero":
                velocity = ("zero", 0.0, 0)
            elif len(t) >= 3 and t[1] == "normal":
                seed = _int(_after(t, "seed", 1, no)[0], no) if "seed" in t else 0
                velocity = ("normal", _num(t[2], no), seed)
            else:
                _fail(no, "velocity zero | velocity normal SIGMA seed S")
        elif k == "steps":
            steps = _int(t[1], no)
        elif k == "at":
            if len(t) < 3:
                _fail(no, "at STEP ACTION ...")
            step = _int(t[1], no)
            what = t[2]
            if what == "adjust" and len(t) == 5:
                adjust(copy_params(p), t[3], 1.0)
                actions.append(Action(step, "adjust", t[3], _num(t[4], no)))
            elif what == "set" and len(t) == 5:
                set_param(copy_params(p), t[3], _num(t[4], no), no)
                actions.append(Action(step, "set", t[3], _num(t[4], no)))
            elif what in ("external_force", "vorticity") and len(t) == 4:
                actions.append(Action(step, what, "", 1.0 if _flag(t[3], no) else 0.0))
            elif what == "toggle" and len(t) == 4 and t[3] == "vorticity":
                actions.append(Action(step, "toggle_vorticity"))
            else:
                _fail(no, "unknown action")
        else:
            _fail(no, f"unknown directive '{k}'")
    pos = np.concatenate(pos) if pos else np.zeros((0, 4), np.float32)
    n = pos.shape[0]
    vel = np.zeros_like(pos)
    if velocity[0] == "normal":
        vel[:, :3] = (velocity[1] * S.normal(velocity[2], 3 * n).reshape(n, 3)).astype(np.float32)
    actions.sort(key=lambda a: a.step)  # stable, as the C++ reader's
    return SceneFile(name, grid, p, pos, vel, steps, actions)


def load(path: str) -> SceneFile:
    with open(path) as f:
        return parse(f.read(), os.path.basename(path))


def run(sf: SceneFile, steps: int | None = None, graph: bool = False):
    """The scene through pbf_amd.sph.SPH, as pbf_run drives pbf::SPH: parameter
    changes applied before their step. Returns the final (pos, vel)."""
    from . import sph as P
    steps = sf.steps if steps is None else steps
    h = P.SPH(sf.n, sf.grid)
    try:
        p = copy_params(sf.params)
        h.set_params(p)
        h.upload(sf.pos, sf.vel)
        h.set_graph(graph)
        nxt = 0
        for s in range(steps):
            changed = False
            while nxt < len(sf.actions) and sf.actions[nxt].step <= s:
                apply(p, sf.actions[nxt])
                nxt += 1
                changed = True
            if changed:
                h.set_params(p)
            h.Run()
        return h.download()
    finally:
        h.close()


def read_state(path: str, n: int):
    """A state file of pbf_run: n float4 positions, then n float4 velocities."""
    a = np.fromfile(path, dtype=np.float32)
    assert a.size == 8 * n, f"{path}: {a.size} floats, expected {8 * n}"
    return a[:4 * n].reshape(n, 4), a[4 * n:].reshape(n, 4)
