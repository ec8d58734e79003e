"""The C restatement (oracle/) against the reference's own shaders run on llvmpipe.

Every stage of every dumped step is compared bit-exactly, each from the
fixture's own input of that stage, so one stage's mismatch cannot hide
behind another's; the multi-step rollouts then compare the whole step in the
reference's dispatch order."""
import numpy as np
import pytest

import oracle as O
from tests import golden_util as G

FIXTURES = G.NAMES


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.int32) if a.dtype == np.float32 else a


def assert_bitexact(got, want, what):
    got, want = bits(got), bits(want)
    assert got.shape == want.shape, what
    bad = np.nonzero(got.ravel() != want.ravel())[0]
    assert bad.size == 0, f"{what}: {bad.size} of {got.size} differ, first at {bad[:5]}"


def cfg_of(g):
    return O.make_cfg(g["pos0"].shape[0], tuple(int(x) for x in g["grid"]), G.params_of(g))


@pytest.mark.parametrize("name", FIXTURES)
def test_stages_bitexact(name):
    g = G.load(name)
    cfg = cfg_of(g)
    for s in [int(x) for x in g["steps_dumped"]]:
        p = f"s{s}_"
        pos, vel, end0, vort0 = G.step_inputs(g, s)
        assert_bitexact(O.predict(cfg, pos, vel), g[p + "predict"], f"{name} step {s} predict")
        srt, perm = O.sort(cfg, g[p + "predict"])
        assert_bitexact(perm, g[p + "perm"], f"{name} step {s} sort permutation")
        sorted_rec = G.sorted_records(g, s)
        assert_bitexact(srt, sorted_rec, f"{name} step {s} sorted records")
        start, end = O.findcells(cfg, sorted_rec, end0)
        assert_bitexact(start, G.dense(g, p, "start", -1), f"{name} step {s} start table")
        assert_bitexact(end, G.dense(g, p, "end", 0), f"{name} step {s} end table")
        nbr = O.neighbourcells(cfg, sorted_rec, start, end)
        assert_bitexact(nbr, g[p + "nbr"], f"{name} step {s} neighbour table")
        prev = sorted_rec
        for it in range(int(g["iterations"])):
            lam = O.lambdas(cfg, prev, g[p + "nbr"])
            assert_bitexact(lam, g[p + "lambda"][it], f"{name} step {s} lambda {it}")
            up = O.updatepos(cfg, prev, g[p + "lambda"][it], g[p + "nbr"], O.ORDER_REFERENCE)
            assert_bitexact(up[:, :3], g[p + "solved"][it], f"{name} step {s} updatepos {it}")
            prev = G.solved_records(g, s, it)
        upos, uvel = O.update(cfg, prev, pos, vel)
        assert_bitexact(upos, g[p + "upd_pos"], f"{name} step {s} update pos")
        assert_bitexact(uvel, g[p + "upd_vel"], f"{name} step {s} update vel")
        if p + "vort" in g:
            v2, vo = O.vorticity(cfg, prev, g[p + "nbr"], g[p + "upd_vel"], vort0, O.ORDER_REFERENCE)
            assert_bitexact(vo, g[p + "vort"], f"{name} step {s} |omega|")
            assert_bitexact(v2, g[p + "vel"], f"{name} step {s} vorticity velocity")


@pytest.mark.parametrize("name", ["dam4k", "extforce", "reference_default"])
def test_rollout_reference_order_bitexact(name):
    """Whole steps in the reference's dispatch order reproduce the llvmpipe rollout."""
    g = G.load(name)
    cfg = cfg_of(g)
    pos, vel = g["pos0"], g["vel0"]
    end = np.zeros(G.cells(g), np.int32)
    vort = np.zeros(pos.shape[0], np.float32)
    for _ in range(int(g["nsteps"])):
        pos, vel, end, vort = O.step(cfg, pos, vel, end, vort, O.ORDER_REFERENCE)
    assert_bitexact(pos, g["final_pos"], f"{name} final positions")
    assert_bitexact(vel, g["final_vel"], f"{name} final velocities")


def test_collide_fixture_has_rounded_keys():
    """The collision fixture exercises keys >= 2^24, where FP32 rounding merges cells."""
    g = G.load("collide")
    cfg = cfg_of(g)
    rec = g["s0_predict"]
    k = O.keys(cfg, rec)
    ix = np.clip(rec[:, :3], 0, np.array(g["grid"], np.float32)).astype(np.int64)
    gx, gy, gz = [int(x) for x in g["grid"]]
    exact = ix[:, 0] + ix[:, 1] * gx * gz + ix[:, 2] * gx
    assert (exact >= 1 << 24).sum() > 1000
    assert (k.astype(np.int64) != exact).sum() > 100


@pytest.mark.parametrize("case", ["dam", "collide"])
def test_highlight_reference_order_bitexact(case):
    """The highlight passes (clearhighlight.glsl, highlight.glsl; src/SPH.cpp:288-296), run inside
    whole steps in the reference's order, reproduce the flag words llvmpipe wrote after each step."""
    g = G.load("highlight")
    p, grid, pos, vel, hl0, want = G.highlight_case(g, case)
    cfg = O.make_cfg(pos.shape[0], grid, p)
    end = np.zeros(int(np.prod(grid)), np.int32)
    vort = np.zeros(pos.shape[0], np.float32)
    hl = hl0.copy()
    for s in range(want.shape[0]):
        pos, vel, end, vort = O.step(cfg, pos, vel, end, vort, O.ORDER_REFERENCE, highlight=hl)
        assert_bitexact(hl, want[s], f"highlight {case} step {s}")
    # the fixture marks neighbours and keeps only bit 0 of the host's words
    assert (want[0] & 2).sum() > 0 and (want[0] & ~np.uint32(3)).sum() == 0


def test_highlight_closed_form():
    """Flags after the passes = (bit 0 of the old word) | 2 for every id that is a candidate (not
    the particle's own slot) in a row of a flagged particle, counted by brute force over the rows."""
    g = G.load("highlight")
    p, grid, pos, vel, hl0, _ = G.highlight_case(g, "collide")
    cfg = O.make_cfg(pos.shape[0], grid, p)
    n = pos.shape[0]
    srt, _ = O.sort(cfg, O.predict(cfg, pos, vel))
    start, end = O.findcells(cfg, srt, np.zeros(int(np.prod(grid)), np.int32))
    nbr = O.neighbourcells(cfg, srt, start, end)
    ids = srt[:, 3].view(np.int32)
    want = hl0 & np.uint32(1)
    words = nbr[:, :, :3].reshape(n, 9)
    for gsl in np.nonzero((hl0[ids] & 1) == 1)[0]:
        for d in words[gsl]:
            e, s0 = int(d) >> 24, int(d) & 0xFFFFFF
            for j in range(s0, s0 + e):
                if j != gsl:
                    want[ids[j] if j < n else 0] |= 2
    assert_bitexact(O.highlight(cfg, srt, nbr, hl0), want, "highlight closed form")
