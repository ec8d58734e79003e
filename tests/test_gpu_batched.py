"""Batched form on the GPU, through the C ABI.

Bar: every scene of a batch equals the plain SPH handle run on that scene
alone, bit for bit (positions, velocities, end tables), over several steps;
a hipGraph-captured step equals the eagerly launched one on every replay, with
eager work and synchronisations between replays; the autotuner leaves the
state where it found it and its choice changes no bit of the results."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from pbf_amd import batch as B
from pbf_amd import scene as S
from pbf_amd import sph as P
from pbf_amd._lib import PbfError

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.int32), np.ascontiguousarray(b).view(np.int32))


def small_batch(nscenes=6, seed=5):
    return S.batched_scenes(nscenes=nscenes, seed=seed, n_lo=2048, n_hi=8192)


def plain_run(scene, steps):
    sph = P.SPH(scene.n, scene.grid)
    sph.set_params(P.scene_params(scene))
    sph.upload(scene.pos, scene.vel)
    for _ in range(steps):
        sph.Run()
    pos, vel = sph.download()
    end = sph.end_table()
    sph.close()
    return pos, vel, end


def test_batched_equals_plain_per_scene(gpu_lib):
    scenes = small_batch()
    steps = 6
    say(f"[batched vs plain] scenes={[s.n for s in scenes]} steps={steps}")
    b = B.Batch(scenes)
    b.upload()
    b.step(steps)
    pos, vel = b.download()
    ends = b.end_tables()
    b.close()
    cell_off = np.concatenate([[0], np.cumsum([int(np.prod(s.grid)) for s in scenes])])
    offs = np.concatenate([[0], np.cumsum([s.n for s in scenes])])
    for k, sc in enumerate(scenes):
        wp, wv, we = plain_run(sc, steps)
        assert same(pos[offs[k]:offs[k + 1]], wp), f"scene {k} positions"
        assert same(vel[offs[k]:offs[k + 1]], wv), f"scene {k} velocities"
        assert np.array_equal(ends[cell_off[k]:cell_off[k + 1]], we), f"scene {k} end table"
        moved = np.abs(wp[:, :3] - sc.pos[:, :3]).max()
        assert moved > 1e-3


def test_full_256_scene_mix_against_oracle(gpu_lib):
    """The whole 256-scene workload (4,277,760 particles, n from 2,048 to 65,536; composite keys of
    8 scene bits above up to 18 cell-key bits) for 10 graph-replayed steps at the tuned launch shape:
    every scene's positions, velocities and end table bit-exact against the oracle's race-free step
    with that scene's gravity direction, walls and XSPH coefficient (the per-scene forms of
    predictpos.glsl:32, updatepos.glsl:98-100 and vorticity.glsl:57; the step src/SPH.cpp:246-334)."""
    scenes = S.batched_scenes()
    steps = 10
    total = sum(s.n for s in scenes)
    say(f"[full mix] scenes={len(scenes)} particles={total} max n={max(s.n for s in scenes)} steps={steps}")
    assert len(scenes) == 256 and total == 4277760
    assert len({s.params.gravity_dir for s in scenes}) == 256
    assert len({s.params.xsph_viscosity_c for s in scenes}) > 200
    b = B.Batch(scenes)
    b.set_tuning(256, 128, 3)
    b.set_graph(True)
    b.upload()
    b.step(steps)
    pos, vel = b.download()
    ends = b.end_tables()
    b.close()
    cell_off = np.concatenate([[0], np.cumsum([int(np.prod(s.grid)) for s in scenes])])
    moved = 0
    for k, sc in enumerate(scenes):
        cfg = O.make_cfg(sc.n, sc.grid, sc.params)
        p, v, e, vo = sc.pos, sc.vel, None, None
        for _ in range(steps):
            p, v, e, vo = O.step(cfg, p, v, e, vo, O.ORDER_JACOBI)
        o = slice(b.offsets[k], b.offsets[k + 1])
        assert same(pos[o], p), f"scene {k} (n={sc.n}) positions"
        assert same(vel[o], v), f"scene {k} (n={sc.n}) velocities"
        assert np.array_equal(ends[cell_off[k]:cell_off[k + 1]], e), f"scene {k} end table"
        moved += int(np.abs(p[:, :3] - sc.pos[:, :3]).max() > 1e-3)
    say(f"[full mix] all {len(scenes)} scenes bit-exact; {moved} moved by more than 1e-3 h")
    assert moved == len(scenes)


@pytest.mark.parametrize("iterations,vorticity,extforce", [(0, True, False), (2, False, False), (3, True, True)])
def test_iteration_counts_against_oracle(gpu_lib, iterations, vorticity, extforce):
    """A batch whose steps run no solver iteration (src/SPH.cpp:303-311 with the count 0: the
    closing update reads the sorted predicted positions and the first vorticity pass alone writes
    the neighbour records the second reads), and one with vorticity off (the reference's default,
    src/SPH.cpp:25: the last position update does the closing update and keeps no sorted
    velocities), and one with the external force on (predictpos.glsl:26-27, against each scene's
    own grid height: about half of every scene's particles). Four scenes for 4 steps, each
    bit-exact against the oracle's step with that scene's gravity, walls and XSPH coefficient."""
    scenes = small_batch(nscenes=4, seed=9)
    for sc in scenes:
        sc.params = dataclasses.replace(sc.params, iterations=iterations, vorticity=vorticity, extforce=extforce)
    steps = 4
    say(f"[batched {iterations} iterations, vorticity {vorticity}, extforce {extforce}] "
        f"scenes={[s.n for s in scenes]} steps={steps}")
    b = B.Batch(scenes)
    b.upload()
    b.step(steps)
    pos, vel = b.download()
    ends = b.end_tables()
    b.close()
    cell_off = np.concatenate([[0], np.cumsum([int(np.prod(s.grid)) for s in scenes])])
    for k, sc in enumerate(scenes):
        cfg = O.make_cfg(sc.n, sc.grid, sc.params)
        assert cfg.iterations == iterations and cfg.vorticity == int(vorticity) and cfg.extforce == int(extforce)
        p, v, e, vo = sc.pos, sc.vel, None, None
        for _ in range(steps):
            p, v, e, vo = O.step(cfg, p, v, e, vo, O.ORDER_JACOBI)
        o = slice(b.offsets[k], b.offsets[k + 1])
        assert same(pos[o], p), f"scene {k} (n={sc.n}) positions"
        assert same(vel[o], v), f"scene {k} (n={sc.n}) velocities"
        assert np.array_equal(ends[cell_off[k]:cell_off[k + 1]], e), f"scene {k} end table"


def test_graph_replays_equal_eager_steps(gpu_lib):
    scenes = small_batch(4, seed=6)
    say(f"[graph] scenes={[s.n for s in scenes]}")
    eager, graph = B.Batch(scenes), B.Batch(scenes)
    eager.upload()
    graph.upload()
    graph.set_graph(True)
    for k in range(5):
        eager.step(1)
        graph.step(1)
        pe, ve = eager.download()
        pg, vg = graph.download()
        assert same(pe, pg) and same(ve, vg), f"replay {k}"
        assert np.array_equal(eager.end_tables(), graph.end_tables()), f"replay {k} end tables"
    graph.step(3)  # several replays queued back to back
    eager.step(3)
    pe, ve = eager.download()
    pg, vg = graph.download()
    assert same(pe, pg) and same(ve, vg)
    eager.close()
    graph.close()


def test_autotune_restores_state_and_keeps_results(gpu_lib):
    scenes = small_batch(3, seed=7)
    say(f"[autotune] scenes={[s.n for s in scenes]}")
    ref, tuned = B.Batch(scenes), B.Batch(scenes)
    ref.upload()
    tuned.upload()
    ref.step(2)
    tuned.step(2)
    p0, v0 = tuned.download()
    e0 = tuned.end_tables()
    best, trials = tuned.autotune(1)
    say(f"[autotune] best {best}, {len(trials)} trials")
    # 16 neighbour shapes (width x list depth x persistent), then the 19 other streaming shapes
    assert len(trials) == 35 and all(t["step_ms"] > 0 for t in trials)
    assert {(t["neighbour_block"], t["list_depth"], t["persistent"]) for t in trials[:16]} == {
        (nb, lc, p) for nb in (64, 128, 256, 512) for lc in (32, 56) for p in (0, 1)}
    p1, v1 = tuned.download()
    assert same(p0, p1) and same(v0, v1) and np.array_equal(e0, tuned.end_tables())
    assert tuned.tuning() == {k: best[k] for k in ("neighbour_block", "stream_block", "stream_ppt", "list_depth",
                                                    "persistent")}
    # shapes change no bit: widths, streaming shapes, list depths, persistent launches
    for shape in [(128, 64, 4), (512, 1024, 2, 32), (256, 256, 1, 56, 1), (64, 128, 3, 32, 1)]:
        tuned.set_tuning(*shape)
        tuned.step(1)
        ref.step(1)
        pr, vr = ref.download()
        pt, vt = tuned.download()
        assert same(pr, pt) and same(vr, vt), f"shape {shape}"
    ref.close()
    tuned.close()


def test_batch_rejects_bad_input(gpu_lib):
    sc = small_batch(1)[0]
    bad = S.Scene("bad", sc.grid, sc.params, sc.pos[:1000], sc.vel[:1000])
    with pytest.raises(PbfError):
        B.Batch([bad])
    b = B.Batch([sc])
    with pytest.raises(Exception):
        b.set_tuning(100, 256, 1)
    with pytest.raises(Exception):
        b.set_tuning(256, 256, 1, 40)
    with pytest.raises(Exception):
        b.set_tuning(256, 256, 1, 56, 2)
    b.close()
