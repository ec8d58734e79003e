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


@pytest.mark.parametrize("graph_on", [False, True])
def test_stream_switch_keeps_step_order(gpu_lib, graph_on):
    """Steps queued on one stream, then more on another (torch's streams, then the handle's own
    again), equal a rollout on one stream bit for bit: set_stream waits for the old stream's work
    (and a graph replay captured on it) before the new stream's steps start."""
    import torch
    sc = S.cfg_dam_32k(seed=6)
    say(f"[stream switch graph={graph_on}] n={sc.n}")
    ref, sw = handle(sc), handle(sc)
    ref.set_graph(graph_on)
    sw.set_graph(graph_on)
    ref.step(9)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    sw.step(3)                         # own stream, not waited for
    sw.set_stream(s1.cuda_stream)
    sw.step(3)
    sw.set_stream(s2.cuda_stream)
    sw.step(2)
    sw.set_stream(None)                # back to the handle's own stream
    sw.step(1)
    assert sw.stream() != s2.cuda_stream
    pr, vr = ref.download()
    ps, vs = sw.download()
    assert same(pr, ps) and same(vr, vs)
    assert np.array_equal(ref.end_table(), sw.end_table())
    ref.close()
    sw.close()
