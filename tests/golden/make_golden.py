"""Writes the golden fixtures of tests/golden/*.npz from the reference's own
compute shaders, run headless under Mesa llvmpipe with one thread (workgroups
then execute in dispatch order, so the reference's in-place passes are
deterministic).

Run in the development container only (it needs the reference checkout and
the runner built by `make -C oracle ref`):

    python tests/golden/make_golden.py [--only highlight]

Each fixture holds, for a few steps, the state before the step (id-order
pos/vel, the persistent end table and vorticity buffer) and every stage output
of that step, plus the state after a multi-step rollout. Large, mostly
constant tables are stored sparsely (indices and values that differ from the
fill value).
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

import llvmpipe  # noqa: E402
from pbf_amd import scene as S  # noqa: E402


def sparse(a, fill):
    idx = np.nonzero(a != fill)[0].astype(np.int64)
    return idx, a[idx].astype(np.int32)


def pairs_scene():
    """512 isolated pairs at random separations below h: every lambda and
    position correction depends on one pair, which pins the per-pair arithmetic."""
    sites = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)
    sites = sites * 4.0 + np.array([17.3, 1.3, 17.3])
    d = S.normal(11, 3 * 512).reshape(512, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= (0.05 + 0.93 * S.uniform(12, 512)).reshape(512, 1)
    p = np.zeros((1024, 4), np.float32)
    p[0::2, :3] = sites
    p[1::2, :3] = sites + d
    sc = S.dam_break((16, 16, 16), (64, 32, 64), steps=1, name="pairs")
    sc.pos = p
    sc.vel = np.zeros_like(p)
    return sc


def collide_scene():
    """Keys at or above 2^24 on the 384x192x384 grid, where the reference's FP32
    dot product rounds and x-adjacent cells can share a key: 2048 particles
    spread over the whole domain (some outside it) plus a dense cluster of 2048
    at y in [150, 160)."""
    grid = (384, 192, 384)
    sc = S.dam_break((16, 16, 16), grid, steps=1, name="collide")
    sc.params.iterations = 1
    u = S.uniform(21, 3 * 2048).reshape(2048, 3)
    a = (u * np.array([grid[0] + 8.0, grid[1] + 8.0, grid[2] + 8.0]) - 4.0)
    v = S.uniform(22, 3 * 2048).reshape(2048, 3)
    b = np.array([100.0, 150.0, 100.0]) + v * np.array([8.0, 10.0, 8.0])
    p = np.zeros((4096, 4), np.float32)
    p[:2048, :3] = a
    p[2048:, :3] = b
    sc.pos = p
    sc.vel = np.zeros_like(p)
    return sc


def reference_default_scene():
    """The reference's own parameters (h 2, grid 128x64x128, epsilon 5, gravity 10,
    dt 0.016, scale 1/W(0.2, 2), vorticity epsilon 5, 5 iterations, src/SPH.cpp:
    25-26, 58, 137-144) and its two mirrored blocks at spacing 0.94 with corners
    (32.5, 0.5, 32.5) and (95.5, 0.5, 95.5) (src/Simulation.cpp:212-246), at 16^3
    per block instead of 32^3; vorticity confinement switched on."""
    a = S.lattice_block((16, 16, 16), 0.94, (32.5, 0.5, 32.5), 0.005, 31)
    b = S.lattice_block((16, 16, 16), 0.94, (95.5, 0.5, 95.5), 0.005, 31, offset=3 * 4096, mirror=True)
    pos = np.concatenate([a, b])
    p = S.Params()
    p.vorticity = True
    return S.Scene("reference_default", (128, 64, 128), p, pos, np.zeros_like(pos), 3)


def dump_step(run, s, prefix, out, cells_fill=True):
    """Stage outputs of step s of a runner output folder."""
    n = run.n
    out[f"{prefix}predict"] = run.records(f"s{s}_predict")
    srt = run.records(f"s{s}_sorted")
    out[f"{prefix}perm"] = srt[:, 3].view(np.int32).copy()
    si, sv = sparse(run.table(f"s{s}_start"), -1)
    out[f"{prefix}start_idx"], out[f"{prefix}start_val"] = si, sv
    ei, ev = sparse(run.table(f"s{s}_end"), 0)
    out[f"{prefix}end_idx"], out[f"{prefix}end_val"] = ei, ev
    out[f"{prefix}nbr"] = run.nbr(f"s{s}_nbr")
    it = 0
    lam, pos = [], []
    while run.has(f"s{s}_lambda{it}"):
        lam.append(run.floats(f"s{s}_lambda{it}"))
        pos.append(run.records(f"s{s}_pos{it}")[:, :3].copy())
        it += 1
    out[f"{prefix}lambda"] = np.stack(lam)
    out[f"{prefix}solved"] = np.stack(pos)
    out[f"{prefix}upd_pos"] = run.records(f"s{s}_upd_pos")
    out[f"{prefix}upd_vel"] = run.records(f"s{s}_upd_vel")
    if run.has(f"s{s}_vort"):
        out[f"{prefix}vort"] = run.floats(f"s{s}_vort")
        out[f"{prefix}vel"] = run.records(f"s{s}_vel")
    return n


def state_after(run, s, prefix, out):
    """Persistent state after step s (the input of step s+1)."""
    out[f"{prefix}in_pos"] = run.records(f"s{s}_upd_pos")
    out[f"{prefix}in_vel"] = run.records(f"s{s}_vel") if run.has(f"s{s}_vel") else run.records(f"s{s}_upd_vel")
    ei, ev = sparse(run.table(f"s{s}_end"), 0)
    out[f"{prefix}in_end_idx"], out[f"{prefix}in_end_val"] = ei, ev
    if run.has(f"s{s}_vort"):
        out[f"{prefix}in_vort"] = run.floats(f"s{s}_vort")


def meta(sc, nsteps, out, dumped):
    p = sc.params
    out["grid"] = np.array(sc.grid, np.int32)
    out["h"] = np.float32(p.h)
    out["ubo"] = p.ubo()
    out["iterations"] = np.int32(p.iterations)
    out["vorticity"] = np.int32(p.vorticity)
    out["extforce"] = np.int32(p.extforce)
    out["nsteps"] = np.int32(nsteps)
    out["steps_dumped"] = np.array(dumped, np.int32)
    out["pos0"] = sc.pos
    out["vel0"] = sc.vel


def make(name, sc, nsteps, dumped, final=True):
    work = tempfile.mkdtemp(prefix=f"golden_{name}_")
    lo = min(dumped) - 1 if min(dumped) > 0 else 0
    run = llvmpipe.run(sc, nsteps, lo, max(dumped) + 1, threads=1, workdir=work)
    out = {}
    meta(sc, nsteps, out, dumped)
    for s in dumped:
        dump_step(run, s, f"s{s}_", out)
        if s > 0:
            state_after(run, s - 1, f"s{s}_", out)
    if final:
        out["final_pos"] = np.fromfile(os.path.join(work, "final_pos.bin"), np.float32).reshape(-1, 4)
        out["final_vel"] = np.fromfile(os.path.join(work, "final_vel.bin"), np.float32).reshape(-1, 4)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    shutil.rmtree(work)
    print(f"{name}: {os.path.getsize(path) / 1024:.0f} KiB")


def highlight_flags(n, seed):
    """Initial highlight words: bit 0 (the host's selection, src/Simulation.cpp:
    172-190) on a few particles, some with other bits that the clear pass must
    drop, and stray high bits on particles without bit 0."""
    f = np.zeros(n, np.uint32)
    pick = (S.uniform(seed, 48) * n).astype(np.int64)
    f[pick[:24]] = 1
    f[pick[24:32]] = 3
    f[pick[32:40]] = 0xFFFFFFF2
    f[pick[40:44]] = 0xFFFFFFFF
    f[pick[44:48]] = 2
    return f


def make_highlight():
    """Highlight passes (src/SPH.cpp:288-296, clearhighlight.glsl, highlight.glsl):
    the flag words after each step of a 3-step dam break, and after one step of
    the key-collision scene (keys >= 2^24, particles outside the grid)."""
    out = {}
    for name, sc, nsteps in (("dam", S.dam_break((16, 16, 16), (64, 32, 64), steps=3, name="hl_dam"), 3),
                             ("collide", collide_scene(), 1)):
        work = tempfile.mkdtemp(prefix=f"golden_hl_{name}_")
        hl0 = highlight_flags(sc.n, 41 if name == "dam" else 42)
        run = llvmpipe.run(sc, nsteps, 0, nsteps, threads=1, workdir=work, highlight_init=hl0)
        p = sc.params
        out[f"{name}_grid"] = np.array(sc.grid, np.int32)
        out[f"{name}_h"] = np.float32(p.h)
        out[f"{name}_ubo"] = p.ubo()
        out[f"{name}_iterations"] = np.int32(p.iterations)
        out[f"{name}_vorticity"] = np.int32(p.vorticity)
        out[f"{name}_pos0"] = sc.pos
        out[f"{name}_vel0"] = sc.vel
        out[f"{name}_hl0"] = hl0
        out[f"{name}_hl"] = np.stack([run.words(f"s{k}_highlight") for k in range(nsteps)])
        shutil.rmtree(work)
    path = os.path.join(HERE, "highlight.npz")
    np.savez_compressed(path, **out)
    print(f"highlight: {os.path.getsize(path) / 1024:.0f} KiB")


def main():
    if not llvmpipe.available():
        raise SystemExit("llvmpipe runner not built or reference shaders missing (make -C oracle ref)")
    if "--only" in sys.argv:
        only = sys.argv[sys.argv.index("--only") + 1]
        if only != "highlight":
            raise SystemExit("--only takes: highlight")
        make_highlight()
        return
    # Dam break, 4096 particles, steps 0-1 and 38-39 of a 40-step rollout.
    make("dam4k", S.dam_break((16, 16, 16), (64, 32, 64), steps=40, name="dam4k"), 40, [0, 1, 39])
    make("pairs", pairs_scene(), 1, [0], final=False)
    make("collide", collide_scene(), 1, [0], final=False)
    ext = S.dam_break((16, 16, 16), (64, 32, 64), steps=3, name="extforce")
    ext.params.extforce = True
    ext.pos[:, 2] += 8.0  # half of the block above GRID_SIZE.z / 2
    make("extforce", ext, 3, [0])
    make("reference_default", reference_default_scene(), 3, [2])
    make_highlight()


if __name__ == "__main__":
    main()
