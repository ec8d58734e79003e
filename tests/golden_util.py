"""Loading of the llvmpipe golden fixtures in tests/golden (see make_golden.py)."""
import os

import numpy as np

from pbf_amd import scene as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["dam4k", "pairs", "collide", "extforce", "reference_default"]


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def params_of(g):
    u = g["ubo"]
    p = S.Params(one_over_rho_0=float(u[0]), epsilon=float(u[1]), gravity=float(u[2]), timestep=float(u[3]),
                 tensile_instability_k=float(u[4]), tensile_instability_scale=float(u[5]),
                 xsph_viscosity_c=float(u[6]), vorticity_epsilon=float(u[7]), iterations=int(g["iterations"]),
                 vorticity=bool(g["vorticity"]), extforce=bool(g["extforce"]), h=float(g["h"]))
    return p


def cells(g):
    gx, gy, gz = [int(x) for x in g["grid"]]
    return gx * gy * gz


def dense(g, prefix, name, fill):
    t = np.full(cells(g), fill, np.int32)
    t[g[f"{prefix}{name}_idx"]] = g[f"{prefix}{name}_val"]
    return t


def step_inputs(g, s):
    """(pos, vel, end, vort) before step s."""
    n = g["pos0"].shape[0]
    if s == 0:
        return g["pos0"], g["vel0"], np.zeros(cells(g), np.int32), np.zeros(n, np.float32)
    p = f"s{s}_"
    vort = g[p + "in_vort"] if p + "in_vort" in g else np.zeros(n, np.float32)
    return g[p + "in_pos"], g[p + "in_vel"], dense(g, p, "in_end", 0), vort


def sorted_records(g, s):
    pre = g[f"s{s}_predict"]
    return pre[g[f"s{s}_perm"]]


def solved_records(g, s, it):
    """Records after solver iteration `it` (xyz from the fixture, ids of the sorted order)."""
    r = sorted_records(g, s).copy()
    r[:, :3] = g[f"s{s}_solved"][it]
    return r


def highlight_case(g, name):
    """(scene params, grid, pos0, vel0, flags before the first step, flags after each step) of one
    case of highlight.npz."""
    u = g[f"{name}_ubo"]
    p = S.config_params()
    p.one_over_rho_0, p.epsilon, p.gravity, p.timestep = (float(x) for x in u[:4])
    p.tensile_instability_k, p.tensile_instability_scale, p.xsph_viscosity_c, p.vorticity_epsilon = (
        float(x) for x in u[4:])
    p.iterations = int(g[f"{name}_iterations"])
    p.vorticity = bool(g[f"{name}_vorticity"])
    p.h = float(g[f"{name}_h"])
    grid = tuple(int(x) for x in g[f"{name}_grid"])
    return p, grid, g[f"{name}_pos0"], g[f"{name}_vel0"], g[f"{name}_hl0"], g[f"{name}_hl"]
