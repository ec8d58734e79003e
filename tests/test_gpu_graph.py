"""Graph replay of the single-scene step (pbf_sph_set_graph) on the GPU:
replays equal eager steps bit for bit, replay after replay, with eager work,
downloads and synchronizations between them; a parameter change recaptures;
phase timing refuses graph steps and per-kernel timing steps eagerly."""
import numpy as np
import pytest

from pbf_amd import _lib
from pbf_amd import scene as S
from pbf_amd import sph as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def handle(sc):
    h = P.SPH(sc.n, sc.grid)
    h.set_params(P.scene_params(sc))
    h.upload(sc.pos, sc.vel)
    return h


def same(a, b):
    return np.array_equal(a.view(np.int32), b.view(np.int32))


def test_graph_replays_equal_eager_steps(gpu_lib):
    sc = S.cfg_dam_32k(seed=5)
    say(f"[graph sph] n={sc.n}")
    eager, graph = handle(sc), handle(sc)
    graph.set_graph(True)
    for k in range(6):
        eager.Run()
        graph.Run()
        pe, ve = eager.download()
        pg, vg = graph.download()
        assert same(pe, pg) and same(ve, vg), f"replay {k}"
        assert np.array_equal(eager.end_table(), graph.end_table()), f"replay {k} end table"
    eager.step(4)  # replays queued back to back
    graph.step(4)
    pe, ve = eager.download()
    pg, vg = graph.download()
    assert same(pe, pg) and same(ve, vg)
    # a parameter change recaptures the step
    for h in (eager, graph):
        h.SetVorticityEpsilon(0.002)
        h.SetNumSolverIterations(2)
    eager.step(3)
    graph.step(3)
    pe, ve = eager.download()
    pg, vg = graph.download()
    assert same(pe, pg) and same(ve, vg), "after a parameter change"
    with pytest.raises(_lib.PbfError):
        graph.timing()
    graph.set_kernel_timing(True)  # per-kernel timing steps eagerly
    graph.Run()
    eager.Run()
    assert graph.kernel_times()["lambda"][1] == 2
    assert graph.timing()["solver_ms"] > 0.0
    pe, ve = eager.download()
    pg, vg = graph.download()
    assert same(pe, pg) and same(ve, vg)
    eager.close()
    graph.close()
