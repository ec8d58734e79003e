"""Times the reference's own compute shaders under Mesa llvmpipe (8 threads) on
two of the workloads, as a CPU baseline for the reference itself. Runs in the
development container only (the shaders are read from the reference checkout);
writes llvmpipe_timing.json beside this file.

    python tests/golden/llvmpipe_timing.py

Per-phase seconds are wall-clock intervals between glFinish calls at the five
phase boundaries of SPH::Run (src/SPH.cpp:246-334): predict, sort, cell
tables, solver iterations, update + vorticity."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pbf_amd import scene as S  # noqa: E402
from tests.golden import llvmpipe as L  # noqa: E402

THREADS = 8
WARMUP = 2


def mean_row(timing_text):
    for line in timing_text.splitlines():
        if line.startswith("# mean"):
            head, vals = line.split(":", 1)
            toks = vals.split()
            return {toks[k]: float(toks[k + 1]) for k in range(0, len(toks), 2)}, head
    raise RuntimeError("runner printed no mean row")


def main():
    if not L.available():
        raise SystemExit("llvmpipe runner or reference shaders not available")
    out = {"threads": THREADS, "warmup_steps": WARMUP, "driver": "Mesa llvmpipe", "workloads": {}}
    for name, sc, steps in [("dam_32k", S.cfg_dam_32k(), 200), ("block_1m", S.cfg_block_1m(), 5)]:
        t0 = time.time()
        r = L.run(sc, steps, threads=THREADS, warmup=WARMUP)
        wall = time.time() - t0
        phases, head = mean_row(r.timing())
        out["workloads"][name] = {
            "n": sc.n, "steps": steps, "timed_steps": steps - WARMUP, "wall_s": round(wall, 2),
            "mean_phase_s": phases,
            "particle_steps_per_s": sc.n / phases["total"],
        }
        print(name, json.dumps(out["workloads"][name]), flush=True)
    with open(os.path.join(HERE, "llvmpipe_timing.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
