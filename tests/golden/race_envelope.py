"""Measures the reference's own race envelope and writes race_envelope.json.

The reference's position update (updatepos.glsl:53-55, 104) and vorticity pass
(vorticity.glsl:60, 79-85) update particles in place while other workgroups
read them, so its result depends on the order in which the driver runs
workgroups. With LP_NUM_THREADS=1 llvmpipe runs them in dispatch order (the
golden fixtures); with 8 threads the order varies from run to run. For each
scene this script runs the reference's shaders (oracle/_ref/glsl_runner, this
container only) once with 1 thread and RUNS times with 8 threads, and the
oracle's race-free order once, all from the same initial state, and records
max and mean |difference| of the positions and velocities against the
1-thread run after 1 and 10 steps (1 step for the 1M block). The race-free
order is what the HIP path computes (bit for bit); the envelope says how far
the reference itself moves between runs. It also records the distance of the
FP32 race-free state from the oracle's FP64 replay of the same steps.

    python tests/golden/race_envelope.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from pbf_amd import scene as S  # noqa: E402
from tests.golden import llvmpipe as L  # noqa: E402
from tests.golden.make_golden import reference_default_scene  # noqa: E402

RUNS = 3


def diff(a, b):
    d = np.abs(a[:, :3].astype(np.float64) - b[:, :3].astype(np.float64))
    return {"max": float(d.max()), "mean": float(d.mean())}


def measure(sc, steps):
    one = L.run(sc, steps, threads=1)
    p1, v1 = one.records("final_pos"), one.records("final_vel")
    env = []
    for _ in range(RUNS):
        r = L.run(sc, steps, threads=8)
        env.append({"pos": diff(r.records("final_pos"), p1), "vel": diff(r.records("final_vel"), v1)})
    cfg = O.cfg_for(sc)
    pos, vel, end, vort = sc.pos, sc.vel, None, None
    for _ in range(steps):
        pos, vel, end, vort = O.step(cfg, pos, vel, end, vort, O.ORDER_JACOBI)
    pd, vd, endd, vortd = sc.pos, sc.vel, None, None
    for _ in range(steps):
        pd, vd, endd, vortd = O.step_f64(cfg, pd, vd, endd, vortd)
    return {"scene": sc.name, "n": sc.n, "h": sc.params.h, "steps": steps,
            "reference_8_threads": env, "race_free": {"pos": diff(pos, p1), "vel": diff(vel, v1)},
            "race_free_vs_fp64_replay": {"pos": diff(pos, pd), "vel": diff(vel, vd)}}


def main():
    scenes = [S.dam_break((16, 16, 16), (64, 32, 64), steps=40, name="dam4k"), S.cfg_dam_32k(),
              reference_default_scene()]
    out = []
    for sc in scenes:
        for steps in (1, 10):
            m = measure(sc, steps)
            print(json.dumps(m), flush=True)
            out.append(m)
    m = measure(S.cfg_block_1m(), 1)
    print(json.dumps(m), flush=True)
    out.append(m)
    with open(os.path.join(HERE, "race_envelope.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
