"""The dense-particle workgroups on the GPU: rollouts in which they carry many
particles (a 4,096-particle dam break: every particle with more than 24
candidates) and rollouts that split the particles between both paths (the
32,768-particle dam break) equal the oracle bit for bit, and the split really
happens (pbf_sph_dense_info)."""
import numpy as np
import pytest

import oracle as O
from pbf_amd import scene as S
from pbf_amd import sph as P
from tests import golden_util as G
from tests.test_gpu_parity import assert_bitexact, oracle_rollout

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rollout_with_split(sc, steps):
    sph = P.SPH(sc.n, sc.grid)
    sph.set_params(P.scene_params(sc))
    sph.upload(sc.pos, sc.vel)
    split = []
    for _ in range(steps):
        sph.Run()
        split.append(sph.dense_info())
    pos, vel = sph.download()
    end = sph.end_table()
    sph.close()
    return pos, vel, end, split


def test_dense_path_carries_small_scene(gpu_lib):
    g = G.load("dam4k")
    sc = S.Scene("dam4k", tuple(int(x) for x in g["grid"]), G.params_of(g), g["pos0"], g["vel0"])
    steps = 40
    say(f"[dense dam4k] n={sc.n} steps={steps}")
    pos, vel, end, split = rollout_with_split(sc, steps)
    counts = [c for _, c in split]
    say(f"[dense dam4k] thresholds {sorted(set(t for t, _ in split))}, dense counts min {min(counts)} max {max(counts)}")
    # a scene smaller than the budget takes the lowest level in every step; on
    # the initial lattice (216 candidates each) that is every particle, and
    # the spreading dam keeps some above it throughout
    assert all(t == 24 for t, _ in split)
    assert counts[0] == sc.n
    assert min(counts) > 0
    wpos, wvel, wend = oracle_rollout(sc, steps)
    assert_bitexact(pos, wpos, "dam4k positions")
    assert_bitexact(vel, wvel, "dam4k velocities")
    assert_bitexact(end, wend, "dam4k end table")


def test_dense_and_light_paths_split_a_scene(gpu_lib):
    sc = S.cfg_dam_32k()
    steps = 25
    say(f"[dense dam_32k] n={sc.n} steps={steps}")
    pos, vel, _, split = rollout_with_split(sc, steps)
    say(f"[dense dam_32k] (threshold, count) per step {split}")
    assert any(0 < c < sc.n for _, c in split), "some step splits the particles between both paths"
    wpos, wvel, _ = oracle_rollout(sc, steps)
    assert_bitexact(pos, wpos, "dam_32k positions")
    assert_bitexact(vel, wvel, "dam_32k velocities")


def test_dense_info_before_a_step(gpu_lib):
    sph = P.SPH(4096, (64, 64, 64))
    assert sph.dense_info() == (0, 0)
    sph.close()
