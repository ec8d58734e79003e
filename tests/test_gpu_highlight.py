"""The highlight passes of SPH::Run (src/SPH.cpp:288-296: clearhighlight.glsl:5-8,
highlight.glsl:16-30) and the highlight buffer of the boundary
(SPH::GetHighlightBuffer, src/SPH.h:205-207), through the C ABI.

The first step's flag words depend only on the predicted positions and the
neighbour table, which equal the reference's bit for bit, so they are compared
with the llvmpipe fixture; later steps follow the race-free positions and are
compared with the oracle's race-free step with the same passes."""
import numpy as np
import pytest

import oracle as O
from pbf_amd import scene as S
from pbf_amd import sph as P
from tests import golden_util as G
from tests import gpu_util as U

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def words_equal(got, want, what):
    bad = np.nonzero(np.asarray(got, np.uint32) != np.asarray(want, np.uint32))[0]
    assert bad.size == 0, f"{what}: {bad.size} of {got.size} words differ, first at {bad[:8]}"


@pytest.mark.parametrize("case", ["dam", "collide"])
@pytest.mark.parametrize("graph", [False, True])
def test_highlight_against_fixture_and_oracle(gpu_lib, case, graph):
    g = G.load("highlight")
    p, grid, pos0, vel0, hl0, want = G.highlight_case(g, case)
    n = pos0.shape[0]
    say(f"[highlight {case} graph={graph}] n={n} grid={grid} steps={want.shape[0]}")
    sc = S.Scene("hl", grid, p, pos0, vel0)
    sph = P.SPH(n, grid)
    sph.set_params(P.scene_params(sc))
    sph.set_graph(graph)
    sph.upload(pos0, vel0)
    sph.set_highlight(hl0)
    cfg = O.make_cfg(n, grid, p)
    pos, vel = pos0.copy(), vel0.copy()
    end = np.zeros(int(np.prod(grid)), np.int32)
    vort = np.zeros(n, np.float32)
    hl = hl0.copy()
    for s in range(want.shape[0]):
        sph.Run()
        got = sph.highlight()
        pos, vel, end, vort = O.step(cfg, pos, vel, end, vort, O.ORDER_JACOBI, highlight=hl)
        words_equal(got, hl, f"{case} step {s} against the oracle")
        if s == 0:
            words_equal(got, want[0], f"{case} step 0 against the reference's shaders")
    # the physics does not see the flags: positions equal the oracle's
    gpos, _ = sph.download()
    assert (gpos.view(np.int32) == pos.view(np.int32)).all()
    sph.close()


def test_highlight_buffer_boundary(gpu_lib):
    """GetHighlightBuffer: uint32 by id, stride 4, zero after creation; the toggle and the clear of
    src/Simulation.cpp:163-193 and :270-272; bit 0 survives a step, other host bits do not."""
    sc = S.dam_break((16, 16, 8), (64, 32, 64), steps=1, name="hlb")
    sph = P.SPH(sc.n, sc.grid)
    sph.set_params(P.scene_params(sc))
    sph.upload(sc.pos, sc.vel)
    import ctypes as C
    ptr, stride = C.c_void_p(), C.c_size_t()
    assert sph._lib.pbf_sph_highlight_device(sph._h, C.byref(ptr), C.byref(stride)) == 0
    assert stride.value == 4 and ptr.value == sph.GetHighlightBuffer()
    words = U.copy_from_device(ptr.value, (sc.n,), np.uint32, sph.stream())
    assert (words == 0).all()
    assert sph.toggle_highlight(7) == 1
    assert sph.toggle_highlight(7) == 0
    assert sph.toggle_highlight(7) == 1
    w = sph.highlight()
    w[9] = 0xF0
    sph.set_highlight(w)
    sph.Run()
    after = U.copy_from_device(ptr.value, (sc.n,), np.uint32, sph.stream())
    assert after[7] & 1 == 1 and (after & ~np.uint32(3)).sum() == 0
    assert (after & 2).sum() > 0 and after[9] & 1 == 0
    sph.set_highlight(None)
    assert (sph.highlight() == 0).all()
    with pytest.raises(Exception):
        sph.toggle_highlight(sc.n)
    sph.close()
