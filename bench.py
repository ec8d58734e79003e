#!/usr/bin/env python3
"""Benchmark of the pbf simulation step on one MI355X.

A step is one SPH::Run (src/SPH.cpp:246-334) over the whole scene: predict,
sort, cell table, I x (lambda, position update), update, XSPH/vorticity.
Default workload: 1,048,576 particles as a 128x64x128 lattice block at 0.5h
(seeded +-0.01h jitter), grid 256x128x256, 4 solver iterations, dt 0.008,
vorticity confinement and XSPH on; inputs are resident in HBM before timing.

Prints one JSON line on stdout (everything else goes to stderr):
  value        particle-steps/s over the timed steps (whole job, 1 GPU)
  roofline     the dominant kernel: algorithmic bytes per launch / its mean
               launch time (HIP events on the library's stream), against 8 TB/s
  cpu_baseline the oracle (C restatement, OpenMP) timed on this host for a
               bounded sample of the same workload
  cpu_reference the reference's own CPU path (its GLSL shaders on Mesa
               llvmpipe), as recorded in tests/golden/llvmpipe_timing.json on
               the machine that made the fixtures (not timed by this run)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s @1M (4 iters), % HBM roofline; batched & adjoint step time"

# Algorithmic HBM bytes per particle of each kernel launch: every stream read and
# written once, neighbour gathers counted as cache reuse (FP32 float4 records,
# u32 keys and indices). lambda: record 16 in + lambda 4 out; updatepos: record
# 16 + lambda 4 in, record 16 out; vorticity passes: 56 each (112 for both).
ALG_BYTES = {
    "lambda": 20.0,
    "updatepos": 36.0,
    "vorticity_a": 56.0,
    "vorticity_b": 56.0,
    "predict": 40.0,
    "update": 48.0,
    "neighbourcells": 12.0,
}

# HBM bytes per particle of one launch, measured for the block_1m workload with
# rocprofv3 FETCH_SIZE and WRITE_SIZE counters over steps 1-2 of the rollout,
# scaled by the counters of a plain 256 MiB copy (profiles/hbm_traffic_block_1m.md).
# lambda writes its pair lists (2 B per in-range neighbour) and updatepos reads
# them back; both also read the 9-row neighbour table (36 B per particle). The
# last updatepos launch of each step also performs the closing update (pos and
# vel by particle id), so updatepos is the mean over launches with and without.
MEASURED_TRAFFIC_BLOCK_1M = {
    "lambda": 183.8,
    "updatepos": 237.1,
    "vorticity_a": 257.8,
    "vorticity_b": 245.3,
    "neighbourcells": 71.0,
    "predict": 60.0,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from pbf_amd import scene as S
    return {"block_1m": S.cfg_block_1m, "dam_32k": S.cfg_dam_32k, "stress_4m": S.cfg_stress_4m}[name]()


def step_alg_bytes(scene) -> float:
    """B_alg = 296 + 20 P8 + 56 I + 4 C/N bytes per particle-step."""
    from pbf_amd.scene import key_bits
    p8 = (key_bits(scene.grid) + 7) // 8
    cells = scene.grid[0] * scene.grid[1] * scene.grid[2]
    return 296.0 + 20.0 * p8 + 56.0 * scene.params.iterations + 4.0 * cells / scene.n


def cpu_baseline(scene, steps: int):
    import oracle as O
    cfg = O.cfg_for(scene)
    pos, vel, end, vort = scene.pos, scene.vel, None, None
    t0 = time.perf_counter()
    for _ in range(steps):
        pos, vel, end, vort = O.step(cfg, pos, vel, end, vort, O.ORDER_JACOBI)
    dt = time.perf_counter() - t0
    return {"value": scene.n * steps / dt, "unit": "particle-steps/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"the first {steps} step(s) of the {scene.name} rollout ({scene.n} particles, "
                      f"{scene.params.iterations} iterations), oracle/pbf_oracle.c FP32 race-free order, "
                      f"{O.num_threads()} OpenMP threads, {dt:.2f} s"}


def cpu_reference(scene):
    """The reference's shaders (calclambda.glsl and the rest of SPH::Run) run
    by Mesa's llvmpipe CPU driver through oracle/glsl_ref, timed when the golden
    fixtures were made (tests/golden/llvmpipe_timing.py); a recorded figure, for
    comparison beside cpu_baseline."""
    path = os.path.join(ROOT, "tests", "golden", "llvmpipe_timing.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    w = rec["workloads"].get(scene.name)
    if not w:
        return None
    return {"value": w["particle_steps_per_s"], "unit": "particle-steps/s", "cores": rec["threads"],
            "kind": "reference",
            "sample": f"{w['timed_steps']} timed steps of the {scene.name} rollout after {rec['warmup_steps']} "
                      f"warm-up steps, the reference's GLSL shaders on {rec['driver']}, {rec['threads']} threads, "
                      f"recorded in tests/golden/llvmpipe_timing.json (the fixture machine, not this host)"}


def bench_batched(steps: int, tune_steps: int):
    """The 256-scene batched mix (pbf_amd.scene.batched_scenes) from its initial
    state: one autotuned, graph-captured batch against the plain solver looped
    over the scenes (one handle and stream per scene, each step issued scene by
    scene), same GPU, same steps."""
    import numpy as np
    import torch
    from pbf_amd import _lib
    from pbf_amd import batch as Bt
    from pbf_amd import scene as S
    from pbf_amd import sph as P
    lib = _lib.load()
    scenes = S.batched_scenes()
    total = sum(s.n for s in scenes)
    cells = sum(int(np.prod(s.grid)) for s in scenes)
    b = Bt.Batch(scenes)
    zero_end = np.zeros(cells, np.int32)

    def reset_batch():
        b.upload()
        b.set_end_tables(zero_end)

    reset_batch()
    b.step(2)
    best, trials = b.autotune(tune_steps)
    log(f"[bench] batched autotune: best {best} of {len(trials)} shapes")
    b.set_graph(True)
    reset_batch()
    b.step(1)  # capture
    reset_batch()
    _lib.check(lib.pbf_synchronize(b.stream()))
    t0 = time.perf_counter()
    b.step(steps)
    _lib.check(lib.pbf_synchronize(b.stream()))
    tb = time.perf_counter() - t0
    b.close()
    handles = []
    for sc in scenes:
        h = P.SPH(sc.n, sc.grid)
        h.set_params(P.scene_params(sc))
        handles.append(h)

    def reset_loop():
        for h, sc in zip(handles, scenes):
            h.upload(sc.pos, sc.vel)
            h.set_end_table(np.zeros(int(np.prod(sc.grid)), np.int32))

    reset_loop()
    for h in handles:
        h.Run()
    reset_loop()
    torch.cuda.synchronize()
    for h in handles:
        _lib.check(lib.pbf_synchronize(h.stream()))
    t0 = time.perf_counter()
    for _ in range(steps):
        for h in handles:
            h.Run()
    for h in handles:
        _lib.check(lib.pbf_synchronize(h.stream()))
    tl = time.perf_counter() - t0
    for h in handles:
        h.close()
    vb, vl = total * steps / tb, total * steps / tl
    p8 = (max(S.key_bits(s.grid) for s in scenes) + 8 + 7) // 8
    b_alg = 296.0 + 20.0 * p8 + 56.0 * scenes[0].params.iterations + 4.0 * cells / total
    log(f"[bench] batched: {1e3 * tb / steps:.3f} ms/step, {vb:.4e} particle-steps/s; looped plain: "
        f"{1e3 * tl / steps:.3f} ms/step, {vl:.4e}; ratio {vb / vl:.2f}")
    return {"workload": f"{len(scenes)} scenes, n in [{min(s.n for s in scenes)}, {max(s.n for s in scenes)}], "
                        f"total {total}, 3 iterations, dt 0.016, {steps} steps from the initial state",
            "value": vb, "unit": "particle-steps/s", "ms_per_step": 1e3 * tb / steps,
            "looped_plain_value": vl, "looped_plain_ms_per_step": 1e3 * tl / steps, "speedup_vs_looped": vb / vl,
            "tuning": {k: best[k] for k in ("neighbour_block", "stream_block", "stream_ppt", "list_depth", "persistent")},
            "step_roofline": {"alg_bytes_per_particle_step": b_alg, "frac": vb * b_alg / (HBM_PEAK_GBS * 1e9)}}


def bench_adjoint(reps: int):
    """The differentiable workload: 16,384 particles (32x16x32 lattice, grid
    64^3), v0 ~ N(0, 0.3), 32 steps x 3 iterations, checkpoints every 8 steps,
    relaxation epsilon 5. Times the FP32 forward rollout and the reverse pass
    (segment recomputation included) and reports per-step times."""
    import numpy as np
    from pbf_amd import adjoint as A
    from pbf_amd import scene as S
    sc = S.cfg_adjoint_16k()
    sc.params.epsilon = 5.0
    tgt = S.cfg_adjoint_16k(seed=40)
    tgt.params.epsilon = 5.0
    r = A.Rollout.for_scene(sc, steps=32, checkpoint_every=8)
    t = r.forward(sc.pos, tgt.vel)
    x = r.forward(sc.pos, sc.vel)
    g = np.zeros_like(x)
    g[:, :3] = 2.0 * (x[:, :3] - t[:, :3]) / (3 * sc.n)
    r.backward(g)
    tf = tb = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        r.forward(sc.pos, sc.vel)
        t1 = time.perf_counter()
        r.backward(g)
        t2 = time.perf_counter()
        tf += t1 - t0
        tb += t2 - t1
    r.close()
    # the plain step (SPH::Run) on the same scene, for comparison
    from pbf_amd import _lib
    from pbf_amd import sph as P
    lib = _lib.load()
    h = P.SPH(sc.n, sc.grid)
    h.set_params(P.scene_params(sc))
    h.upload(sc.pos, sc.vel)
    for _ in range(4):
        h.Run()
    _lib.check(lib.pbf_synchronize(h.stream()))
    h.upload(sc.pos, sc.vel)
    t0 = time.perf_counter()
    for _ in range(reps * 32):
        h.Run()
    _lib.check(lib.pbf_synchronize(h.stream()))
    plain = 1e3 * (time.perf_counter() - t0) / (reps * 32)
    h.close()
    fwd = 1e3 * tf / (reps * 32)
    bwd = 1e3 * tb / (reps * 32)
    log(f"[bench] adjoint workload: forward {fwd:.3f} ms/step, reverse {bwd:.3f} ms/step, ratio {bwd / fwd:.2f}; "
        f"plain step {plain:.3f} ms")
    return {"workload": "16384 particles, 32 steps x 3 iterations, checkpoint every 8, FP32, epsilon 5",
            "forward_ms_per_step": fwd, "adjoint_ms_per_step": bwd, "adjoint_over_forward": bwd / fwd,
            "plain_step_ms": plain, "adjoint_over_plain_step": bwd / plain}


def reset(sph, scene):
    """Initial state of the workload: positions, velocities and the persistent
    end table of the cell finder (never cleared by a step) back to a fresh handle's."""
    import numpy as np
    sph.upload(scene.pos, scene.vel)
    gx, gy, gz = scene.grid
    sph.set_end_table(np.zeros(gx * gy * gz, np.int32))


def timed_run(sph, scene, steps, window):
    """Runs `steps` steps from the initial state; returns (seconds, per-window ms/step)."""
    from pbf_amd import _lib
    lib = _lib.load()
    reset(sph, scene)
    _lib.check(lib.pbf_synchronize(sph.stream()))
    import torch
    torch.cuda.synchronize()
    windows = []
    t0 = time.perf_counter()
    tw = t0
    for k in range(steps):
        sph.Run()
        if window and (k + 1) % window == 0:
            _lib.check(lib.pbf_synchronize(sph.stream()))
            now = time.perf_counter()
            windows.append(1e3 * (now - tw) / window)
            tw = now
    _lib.check(lib.pbf_synchronize(sph.stream()))
    torch.cuda.synchronize()
    return time.perf_counter() - t0, windows


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's step count)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="block_1m", choices=["block_1m", "dam_32k", "stress_4m"])
    ap.add_argument("--window", type=int, default=100, help="steps per reported window of the timed run")
    ap.add_argument("--no-kernel-timing", action="store_true", help="skip the per-kernel HIP-event run")
    ap.add_argument("--eager", action="store_true",
                    help="launch every kernel of a step from the host instead of replaying the step's hipGraph")
    ap.add_argument("--cpu-steps", type=int, default=20, help="oracle steps of the CPU baseline sample (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=100, help="steps of the batched mix (0: skip)")
    ap.add_argument("--tune-steps", type=int, default=2, help="steps per autotuner trial of the batched mix")
    ap.add_argument("--adjoint-reps", type=int, default=3, help="timed forward+reverse rollouts of the adjoint workload (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, launch a plain FP32 streaming copy (y = 1 * x) of MIB MiB three times: "
                         "the known-size kernel that scales FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one GPU only: the library has no multi-GPU path (a scene would shard "
            "by slabs of the sorted cell order with halo exchange, a batch by scene). Run with --gpus 1.")
        sys.exit(2)

    import torch
    from pbf_amd import sph as P

    scene = workload(args.workload)
    n = scene.n
    steps = args.steps if args.steps is not None else scene.steps
    log(f"[bench] workload {scene.name}: n={n} grid={scene.grid} iterations={scene.params.iterations} "
        f"dt={scene.params.timestep}, {steps} timed steps from the initial state")
    sph = P.SPH(n, scene.grid)
    sph.set_params(P.scene_params(scene))
    # each step one replay of a hipGraph of its launches (same kernels and
    # work; the per-kernel timing run below launches eagerly)
    sph.set_graph(not args.eager)
    torch.cuda.init()
    reset(sph, scene)
    for _ in range(args.warmup):
        sph.Run()
    elapsed, windows = timed_run(sph, scene, steps, args.window)
    value = n * steps / elapsed
    ms_per_step = 1e3 * elapsed / steps
    log(f"[bench] {steps} steps in {elapsed:.4f} s: {ms_per_step:.4f} ms/step, {value:.4e} particle-steps/s")
    if windows:
        log("[bench] ms/step per window of " + str(args.window) + ": " + " ".join(f"{w:.3f}" for w in windows))

    roofline = None
    if not args.no_kernel_timing:
        # the same run again with every kernel launch bracketed by HIP events
        sph.set_kernel_timing(True)
        elapsed_k, _ = timed_run(sph, scene, steps, 0)
        kt = sph.kernel_times()
        sph.set_kernel_timing(False)
        log(f"[bench] with per-kernel events: {1e3 * elapsed_k / steps:.4f} ms/step")
        for name, (ms, cnt) in sorted(kt.items(), key=lambda kv: -kv[1][0]):
            if cnt:
                log(f"[bench] kernel {name:16s} {ms / cnt * 1e3:10.2f} us/launch x {cnt / steps:.0f} per step, "
                    f"{ms / steps:.4f} ms/step")
        dom = max((k for k in kt if k in ALG_BYTES and kt[k][1]), key=lambda k: kt[k][0])
        dom_ms = kt[dom][0] / kt[dom][1]
        alg = ALG_BYTES[dom] * n
        achieved = alg / (dom_ms * 1e-3) / 1e9
        traffic = None
        if scene.name == "block_1m" and dom in MEASURED_TRAFFIC_BLOCK_1M:
            traffic = int(MEASURED_TRAFFIC_BLOCK_1M[dom] * n)
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": dom,
                    "launch_us": round(dom_ms * 1e3, 2), "alg_bytes_per_launch": int(alg),
                    "kernel_share_of_step": round(kt[dom][0] / (1e3 * elapsed_k), 4)}
    b_alg = step_alg_bytes(scene)
    step_frac = b_alg * value / (HBM_PEAK_GBS * 1e9)

    batched = None
    if args.batched_steps > 0:
        sph.close()
        batched = bench_batched(args.batched_steps, args.tune_steps)

    adjoint = None
    if args.adjoint_reps > 0:
        adjoint = bench_adjoint(args.adjoint_reps)

    cpu = None
    if args.cpu_steps > 0:
        log(f"[bench] CPU baseline: {args.cpu_steps} oracle step(s) ...")
        cpu = cpu_baseline(scene, args.cpu_steps)
        log(f"[bench] CPU baseline {cpu['value']:.4e} particle-steps/s on {cpu['cores']} threads")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{scene.name}: {n} particles, 128x64x128 lattice block at 0.5h, grid "
                               f"{scene.grid[0]}x{scene.grid[1]}x{scene.grid[2]}, {scene.params.iterations} "
                               f"iterations, dt {scene.params.timestep}, XSPH + vorticity on, {steps}-step "
                               f"rollout from the initial state"
                   if scene.name == "block_1m" else scene.name,
                   "particles": n, "grid": list(scene.grid), "iterations": scene.params.iterations,
                   "parallelism": "single GPU", "launch": "eager" if args.eager else "graph replay per step",
                   "ms_per_step_windows": [round(w, 3) for w in windows]},
        "step_roofline": {"alg_bytes_per_particle_step": b_alg, "frac": round(step_frac, 5)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_reference": cpu_reference(scene),
        "batched": batched,
        "adjoint": adjoint,
    }
    print(json.dumps(line), flush=True)
    sph.close()
    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: 3 x {args.copy_calibration} MiB read + {args.copy_calibration} MiB written")


if __name__ == "__main__":
    main()
