"""HIP path against the golden fixtures (the reference's shaders on llvmpipe) and
the oracle, through the C ABI.

Bar: bit-exact everywhere. Keys, permutation and tables are integers; the FP32
stages follow the reference's operation order without fused multiply-add, so
the race-free stages equal the fixtures bit for bit and the race-free
replacements of the reference's in-place passes (updatepos, vorticity) equal
the oracle's race-free order bit for bit."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from pbf_amd import scene as S
from pbf_amd import sph as P
from tests import golden_util as G
from tests import gpu_util as U

pytestmark = pytest.mark.gpu

FIXTURES = G.NAMES


def say(*a):
    print(*a, flush=True)


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.int32) if a.dtype == np.float32 else a


def assert_bitexact(got, want, what):
    got, want = bits(got), bits(want)
    assert got.shape == want.shape, f"{what}: shape {got.shape} vs {want.shape}"
    bad = np.nonzero(got.ravel() != want.ravel())[0]
    assert bad.size == 0, f"{what}: {bad.size} of {got.size} differ, first at {bad[:8]}"


def fixture_params(g):
    p = G.params_of(g)
    sc = S.Scene("fx", tuple(int(x) for x in g["grid"]), p, g["pos0"], g["vel0"])
    return P.scene_params(sc), O.make_cfg(g["pos0"].shape[0], sc.grid, p)


@pytest.mark.parametrize("name", FIXTURES)
def test_stages_against_fixtures(gpu_lib, name):
    g = G.load(name)
    pp, cfg = fixture_params(g)
    grid = tuple(int(x) for x in g["grid"])
    n = g["pos0"].shape[0]
    st = U.stream()
    for s in [int(x) for x in g["steps_dumped"]]:
        p = f"s{s}_"
        say(f"[{name} step {s}] n={n} grid={grid}")
        pos, vel, end0, vort0 = G.step_inputs(g, s)
        # predict
        dpos, dvel = U.dev(pos), U.dev(vel)
        rec = U.empty_like_host((n, 4), np.float32)
        P.stage_predict(n, grid, pp, U.ptr(dpos), U.ptr(dvel), U.ptr(rec), st)
        assert_bitexact(U.host(rec), g[p + "predict"], f"{name} {s} predict")
        # sort (RadixSort class)
        rs = P.RadixSort(512, n // 512, grid)
        U.copy_to_device(rs.GetBuffer(), g[p + "predict"], rs.stream())
        rs.Run()
        sorted_rec = U.copy_from_device(rs.GetBuffer(), (n, 4), np.float32, rs.stream())
        assert_bitexact(sorted_rec, G.sorted_records(g, s), f"{name} {s} sorted records")
        perm = U.copy_from_device(rs.permutation_ptr(), (n,), np.int32, rs.stream())
        assert_bitexact(perm, g[p + "perm"], f"{name} {s} permutation")
        rs.close()
        # cells (NeighbourCellFinder class)
        f = P.NeighbourCellFinder(n, grid)
        f.set_end_table(end0)
        dsorted = U.dev(G.sorted_records(g, s))
        f.FindNeighbourCells(U.ptr(dsorted))
        start, end = f.tables()
        assert_bitexact(start, G.dense(g, p, "start", -1), f"{name} {s} start")
        assert_bitexact(end, G.dense(g, p, "end", 0), f"{name} {s} end")
        nbr = U.copy_from_device(f.GetResult(), (n, 3, 4), np.int32, f.stream())
        assert_bitexact(nbr, g[p + "nbr"], f"{name} {s} neighbour table")
        f.close()
        # solver stages from the fixture's own inputs
        dnbr = U.dev(g[p + "nbr"])
        prev = G.sorted_records(g, s)
        lam = U.empty_like_host((n,), np.float32)
        out = U.empty_like_host((n, 4), np.float32)
        for it in range(int(g["iterations"])):
            dprev = U.dev(prev)
            P.stage_lambda(n, grid, pp, U.ptr(dprev), U.ptr(dnbr), U.ptr(lam), st)
            assert_bitexact(U.host(lam), g[p + "lambda"][it], f"{name} {s} lambda {it}")
            dlam = U.dev(g[p + "lambda"][it])
            P.stage_updatepos(n, grid, pp, U.ptr(dprev), U.ptr(dlam), U.ptr(dnbr), U.ptr(out), st)
            want = O.updatepos(cfg, prev, g[p + "lambda"][it], g[p + "nbr"], O.ORDER_JACOBI)
            assert_bitexact(U.host(out), want, f"{name} {s} updatepos {it} (race-free)")
            prev = G.solved_records(g, s, it)
        dprev = U.dev(prev)
        dpos, dvel = U.dev(pos), U.dev(vel)
        P.stage_update(n, pp, U.ptr(dprev), U.ptr(dpos), U.ptr(dvel), st)
        assert_bitexact(U.host(dpos), g[p + "upd_pos"], f"{name} {s} update pos")
        assert_bitexact(U.host(dvel), g[p + "upd_vel"], f"{name} {s} update vel")
        if p + "vort" in g:
            dvel = U.dev(g[p + "upd_vel"])
            vort = U.empty_like_host((n,), np.float32)
            P.stage_vorticity(n, grid, pp, U.ptr(dprev), U.ptr(dnbr), U.ptr(dvel), U.ptr(vort), st)
            wv, wo = O.vorticity(cfg, prev, g[p + "nbr"], g[p + "upd_vel"], vort0, O.ORDER_JACOBI)
            assert_bitexact(U.host(vort), wo, f"{name} {s} |omega| (race-free)")
            assert_bitexact(U.host(dvel), wv, f"{name} {s} vorticity velocity (race-free)")


def run_sph(scene, steps, pparams=None):
    sph = P.SPH(scene.n, scene.grid)
    sph.set_params(pparams or P.scene_params(scene))
    sph.upload(scene.pos, scene.vel)
    for _ in range(steps):
        sph.Run()
    pos, vel = sph.download()
    return sph, pos, vel


def oracle_rollout(scene, steps):
    cfg = O.cfg_for(scene)
    pos, vel, end, vort = scene.pos, scene.vel, None, None
    for _ in range(steps):
        pos, vel, end, vort = O.step(cfg, pos, vel, end, vort, O.ORDER_JACOBI)
    return pos, vel, end


@pytest.mark.parametrize("name,steps", [("dam4k", 40), ("extforce", 3), ("reference_default", 3)])
def test_rollout_matches_oracle(gpu_lib, name, steps):
    """Whole SPH::Run steps on the GPU equal the oracle's race-free rollout bit for bit."""
    g = G.load(name)
    p = G.params_of(g)
    sc = S.Scene(name, tuple(int(x) for x in g["grid"]), p, g["pos0"], g["vel0"])
    say(f"[rollout {name}] n={sc.n} steps={steps}")
    sph, pos, vel = run_sph(sc, steps)
    wpos, wvel, wend = oracle_rollout(sc, steps)
    assert_bitexact(pos, wpos, f"{name} positions after {steps} steps")
    assert_bitexact(vel, wvel, f"{name} velocities after {steps} steps")
    assert_bitexact(sph.end_table(), wend, f"{name} end table after {steps} steps")


@pytest.mark.parametrize("iterations,vorticity", [(0, True), (0, False), (1, True), (2, False), (3, False)])
def test_iteration_counts_rollout(gpu_lib, iterations, vorticity):
    """Zero solver iterations (the loop of src/SPH.cpp:303-311 runs no lambda or position update:
    the closing update of src/SPH.cpp:315-318 reads the sorted predicted positions, and the first
    vorticity pass alone writes the neighbour records the second reads), a single one (the
    closing update done by the only position update), and vorticity off with even and odd
    counts (the reference's default, src/SPH.cpp:25: no vorticity pass, so the closing update
    keeps no sorted velocities), on a 4,096-particle dam break for 8 steps: positions,
    velocities and end table bit-exact against the oracle."""
    sc = S.dam_break((16, 16, 16), (64, 32, 64), steps=8, name="iters")
    sc.params = dataclasses.replace(sc.params, iterations=iterations, vorticity=vorticity)
    say(f"[iterations {iterations} vorticity {vorticity}] n={sc.n}")
    sph, pos, vel = run_sph(sc, 8)
    wpos, wvel, wend = oracle_rollout(sc, 8)
    assert_bitexact(pos, wpos, f"positions ({iterations} iterations)")
    assert_bitexact(vel, wvel, f"velocities ({iterations} iterations)")
    assert_bitexact(sph.end_table(), wend, f"end table ({iterations} iterations)")
    assert np.abs(pos[:, :3] - sc.pos[:, :3]).max() > 1e-2


def test_dam_32k_rollout(gpu_lib):
    """The 32,768-particle dam break (grid 64^3, 3 iterations): 25 steps bit-exact."""
    sc = S.cfg_dam_32k()
    say(f"[dam_32k] n={sc.n}")
    _, pos, vel = run_sph(sc, 25)
    wpos, wvel, _ = oracle_rollout(sc, 25)
    assert_bitexact(pos, wpos, "dam_32k positions")
    assert_bitexact(vel, wvel, "dam_32k velocities")


def test_block_1m_one_step(gpu_lib):
    """The 1,048,576-particle block (grid 256x128x256, 4 iterations): one step
    bit-exact against the oracle, plus size-independent checks of the sort."""
    sc = S.cfg_block_1m()
    say(f"[block_1m] n={sc.n}")
    sph, pos, vel = run_sph(sc, 1)
    wpos, wvel, _ = oracle_rollout(sc, 1)
    assert_bitexact(pos, wpos, "block_1m positions")
    assert_bitexact(vel, wvel, "block_1m velocities")
    rec = sph.sorted_records()
    ids = rec[:, 3].view(np.int32)
    assert np.array_equal(np.sort(ids), np.arange(sc.n)), "sorted ids are a permutation"


def test_block_1m_graph_rollout(gpu_lib):
    """The timed configuration itself as hipGraph replays, bit-exact against the oracle with the
    never-cleared end table: after the first 20 steps (the window that bench.py --steps 20
    times: the lattice phase, whose positions are tame) and after 100 (through the violent
    collapse, where particles travel up to about 220 h and the dense-particle and list-overflow
    paths are busiest). The oracle takes about a minute for the 100 steps on 16 threads."""
    sc = S.cfg_block_1m()
    cfg = O.cfg_for(sc)
    say(f"[block_1m graph] n={sc.n} steps=100, checked at 20 and 100")
    sph = P.SPH(sc.n, sc.grid)
    sph.set_params(P.scene_params(sc))
    sph.upload(sc.pos, sc.vel)
    sph.set_graph(True)
    got, done = {}, 0
    for steps in (20, 100):
        sph.step(steps - done)
        done = steps
        got[steps] = (*sph.download(), sph.end_table().copy())
    sph.close()
    pos, vel, end, vort = sc.pos, sc.vel, None, None
    for s in range(1, 101):
        pos, vel, end, vort = O.step(cfg, pos, vel, end, vort, O.ORDER_JACOBI)
        if s in got:
            say(f"[block_1m graph] oracle at step {s}")
            gp, gv, ge = got[s]
            assert_bitexact(gp, pos, f"block_1m positions after {s} graph steps")
            assert_bitexact(gv, vel, f"block_1m velocities after {s} graph steps")
            assert_bitexact(ge, end, f"block_1m end table after {s} graph steps")


def test_stress_4m_one_step(gpu_lib):
    """The 4,194,304-particle stress scene (grid 384x192x384: 26-bit keys, four
    sort passes, keys above 2^24 that the reference's FP32 hash rounds; a dense
    block under uniform-random particles, 0 to 216 candidates per particle): one
    step bit-exact against the oracle, with its sorted records and neighbour table."""
    sc = S.cfg_stress_4m()
    say(f"[stress_4m] n={sc.n}")
    sph, pos, vel = run_sph(sc, 1)
    cfg = O.cfg_for(sc)
    wpos, wvel, _, _, st = O.step(cfg, sc.pos, sc.vel, None, None, O.ORDER_JACOBI, stages=True)
    # records after the solver iterations, in sorted order: the sort
    # permutation (ids in w) and the last iteration's positions
    assert_bitexact(sph.sorted_records(), st["solved"][-1], "stress_4m solved records in sorted order")
    assert_bitexact(sph.sorted_records()[:, 3], st["sorted"][:, 3], "stress_4m sort permutation")
    assert_bitexact(sph.neighbour_table()[:, :, :3], st["nbr"][:, :, :3], "stress_4m neighbour table")
    assert_bitexact(pos, wpos, "stress_4m positions")
    assert_bitexact(vel, wvel, "stress_4m velocities")
    th, cnt = sph.dense_info()
    say(f"[stress_4m] dense threshold {th}, {cnt} dense particles")
    sph.close()


def test_stress_4m_graph_rollout(gpu_lib):
    """The 4,194,304-particle stress scene (a dense block under uniform-random particles, whose
    densest particles take the dense-particle path) for 3 hipGraph-replayed steps, bit-exact
    against the oracle with the never-cleared end table."""
    sc = S.cfg_stress_4m()
    steps = 3
    say(f"[stress_4m graph] n={sc.n} steps={steps}")
    sph = P.SPH(sc.n, sc.grid)
    sph.set_params(P.scene_params(sc))
    sph.upload(sc.pos, sc.vel)
    sph.set_graph(True)
    sph.step(steps)
    pos, vel = sph.download()
    wpos, wvel, wend = oracle_rollout(sc, steps)
    assert_bitexact(pos, wpos, f"stress_4m positions after {steps} graph steps")
    assert_bitexact(vel, wvel, f"stress_4m velocities after {steps} graph steps")
    assert_bitexact(sph.end_table(), wend, f"stress_4m end table after {steps} graph steps")
    th, cnt = sph.dense_info()
    say(f"[stress_4m graph] dense threshold {th}, {cnt} dense particles")
    sph.close()


def test_error_paths(gpu_lib):
    """C-ABI errors: bad sizes and odd sort blocks fail with the reference's error kinds."""
    with pytest.raises(P._lib.PbfError):
        P.SPH(1000, (64, 64, 64))
    with pytest.raises(P._lib.PbfLogicError):
        P.RadixSort(511, 2, (64, 64, 64))
    with pytest.raises(P._lib.PbfError):
        P.SPH(512, (0, 64, 64))


def test_violent_rollout(gpu_lib):
    """Large random velocities (30 h/s per axis) scatter a dam-break block within a few steps:
    cells empty and refill, rows change length every step and particles leave through the walls'
    clamp; the rollout must still equal the oracle bit for bit."""
    sc = S.dam_break((16, 16, 16), (64, 32, 64), steps=12, name="violent")
    n = sc.n
    sc.vel = np.zeros_like(sc.pos)
    sc.vel[:, :3] = (30.0 * S.normal(77, 3 * n).reshape(n, 3)).astype(np.float32)
    say(f"[violent] n={n}")
    _, pos, vel = run_sph(sc, 12)
    wpos, wvel, _ = oracle_rollout(sc, 12)
    assert_bitexact(pos, wpos, "violent positions")
    assert_bitexact(vel, wvel, "violent velocities")


def isolated_pairs(seed=91):
    """3,072 particles in 1,536 pairs whose members are about h apart, each pair more than 2h
    from every other. Pair midpoints lie on a 4h lattice that includes the multiples of 16h in
    every axis, so many pairs straddle one; separations h (1 - 10^U(-7, -2)) for three pairs in
    four, h (1 + 10^U(-7, -2)) for the rest; random orientations and velocities."""
    rng = np.random.default_rng(seed)
    mx = np.arange(20.0, 77.0, 4.0)
    my = np.arange(4.0, 29.0, 4.0)
    mid = np.stack(np.meshgrid(mx, my, mx, indexing="ij"), axis=-1).reshape(-1, 3)[:1536]
    u = rng.normal(size=(1536, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    d = 10.0 ** rng.uniform(-7, -2, 1536)
    r = np.where(np.arange(1536) % 4 == 3, 1.0 + d, 1.0 - d)
    pos = np.zeros((3072, 4), np.float32)
    pos[0::2, :3] = mid - 0.5 * r[:, None] * u
    pos[1::2, :3] = mid + 0.5 * r[:, None] * u
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(size=(3072, 3))
    return S.Scene("pairs", (96, 32, 96), S.config_params(1.0, 0.008, 4), pos, vel, 1)


def test_candidate_filter_keeps_pairs_at_h(gpu_lib):
    """The candidate scans test distances on quantized positions before the exact test. A pair
    it dropped would add nothing, so the scene is isolated pairs at separations next to h,
    straddling the 16h period of the quantized coordinates: each particle's |omega| is nonzero
    exactly when its partner is closer than h (fl(sqrt(sq)) < h, counted here in float32 the
    reference's way), and lambda and the vorticity stage equal the oracle bit for bit."""
    sc = isolated_pairs()
    n = sc.n
    cfg = O.cfg_for(sc)
    pp = P.scene_params(sc)
    xs = O.predict(cfg, sc.pos, sc.vel)
    rec, _ = O.sort(cfg, xs)
    start, end = O.findcells(cfg, rec, np.zeros(O.cells(cfg), np.int32))
    nbr = O.neighbourcells(cfg, rec, start, end)
    say(f"[pairs at h] n={n}")
    st = U.stream()
    drec, dnbr = U.dev(rec), U.dev(nbr)
    lam = U.empty_like_host((n,), np.float32)
    P.stage_lambda(n, sc.grid, pp, U.ptr(drec), U.ptr(dnbr), U.ptr(lam), st)
    assert_bitexact(U.host(lam), O.lambdas(cfg, rec, nbr), "lambda of the pairs")
    dvel = U.dev(sc.vel)
    vort = U.empty_like_host((n,), np.float32)
    P.stage_vorticity(n, sc.grid, pp, U.ptr(drec), U.ptr(dnbr), U.ptr(dvel), U.ptr(vort), st)
    wv, wo = O.vorticity(cfg, rec, nbr, sc.vel, np.zeros(n, np.float32), O.ORDER_JACOBI)
    got = U.host(vort)
    assert_bitexact(U.host(dvel), wv, "velocities of the pairs")
    assert_bitexact(got, wo, "|omega| of the pairs")
    # partner of each sorted slot, and the reference's range test in float32
    ids = rec[:, 3].view(np.int32)
    slot = np.empty(n, np.int64)
    slot[ids] = np.arange(n)
    partner = slot[ids ^ 1]
    d = rec[:, :3] - rec[partner, :3]
    sq = (d[:, 2] * d[:, 2] + d[:, 1] * d[:, 1]) + d[:, 0] * d[:, 0]
    # a partner at fl(sqrt(sq)) == h is in range but adds exact zeros (h - r = 0)
    within = np.sqrt(sq) < np.float32(1.0)
    # the partner among the particle's candidates: inside one of its nine row ranges (the
    # reference's first and last cells are not: findcells.glsl writes no start for the cell of
    # sorted slot 0 and no end for the last one)
    rows = nbr.reshape(n, 12)[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]]
    rs, rc = rows & 0xFFFFFF, rows >> 24
    listed = ((rs <= partner[:, None]) & (partner[:, None] < rs + rc)).any(axis=1)
    assert (~listed).sum() <= 8, f"{(~listed).sum()} partners outside the rows"
    assert np.array_equal(got != 0.0, within & listed), "a pair within h was dropped or one beyond h was kept"
    assert within.mean() > 0.4 and (~within).mean() > 0.2, f"{within.mean():.2f} of the particles have a partner"


def test_tiny_coordinates_rollout(gpu_lib):
    """The pair passes skip the division's range tests only when no position of the set they read
    has a coordinate in (0, 2^-25) in magnitude. Here, without gravity or initial velocity (x* = x),
    a layer of the block sits at y = 1e-9 .. 2.9e-8 above the floor and a column at z = 1e-12 in a
    grid whose walls start at 0, so the first iterations read tiny coordinates; 6 steps must equal
    the oracle bit for bit."""
    sc = S.dam_break((16, 8, 16), (64, 32, 64), steps=6, name="tiny")
    sc.params.gravity = 0.0
    sc.params.wall_lo = (0.0, 0.0, 0.0)
    sc.params.wall_hi_margin = (0.0, 0.0, 0.0)
    sc.pos = sc.pos.copy()
    low = sc.pos[:, 1] < 0.5
    sc.pos[low, 1] = np.linspace(1e-9, 2.9e-8, int(low.sum())).astype(np.float32)
    col = np.arange(sc.n) % 97 == 0
    sc.pos[col, 2] = np.float32(1e-12)
    assert np.any((np.abs(sc.pos[:, :3]) > 0) & (np.abs(sc.pos[:, :3]) < 2.0 ** -25))
    say(f"[tiny coordinates] n={sc.n}")
    _, pos, vel = run_sph(sc, 6)
    wpos, wvel, _ = oracle_rollout(sc, 6)
    assert_bitexact(pos, wpos, "tiny-coordinate positions")
    assert_bitexact(vel, wvel, "tiny-coordinate velocities")


@pytest.mark.parametrize("index", [0, 97, 200])
def test_rotated_gravity_one_cell_walls_rollout(gpu_lib, index):
    """The plain handle on scenes of the batched workload: gravity turned off the y axis, walls one
    cell in from the x and z faces (so the wall clamp and the neighbour rows reach the grid's
    edge cells), a per-scene XSPH coefficient; 12 steps, each bit-exact against the oracle."""
    sc = S.batched_scenes()[index]
    steps = 12
    say(f"[rotated gravity] scene {index} n={sc.n} grid={sc.grid} g={sc.params.gravity_dir} "
        f"c={sc.params.xsph_viscosity_c:.4f}")
    assert sc.params.gravity_dir[0] != 0.0 and sc.params.wall_lo == (1.0, 0.0, 1.0)
    sph = P.SPH(sc.n, sc.grid)
    sph.set_params(P.scene_params(sc))
    sph.upload(sc.pos, sc.vel)
    cfg = O.make_cfg(sc.n, sc.grid, sc.params)
    pos, vel, end, vort = sc.pos, sc.vel, None, None
    for k in range(steps):
        sph.Run()
        pos, vel, end, vort = O.step(cfg, pos, vel, end, vort, O.ORDER_JACOBI)
        gp, gv = sph.download()
        assert_bitexact(gp, pos, f"scene {index} step {k} positions")
        assert_bitexact(gv, vel, f"scene {index} step {k} velocities")
    assert_bitexact(sph.end_table(), end, f"scene {index} end table")
    # the walls were reached: some particle sits on a one-cell wall
    gp, _ = sph.download()
    assert ((gp[:, 0] == 1.0) | (gp[:, 2] == 1.0) | (gp[:, 0] == sc.grid[0] - 1.0) |
            (gp[:, 2] == sc.grid[2] - 1.0)).any()
    sph.close()


def test_position_export_render_feed(gpu_lib):
    """SPH::GetPositionBuffer / GetVelocityBuffer (src/SPH.h:49-51, 181-183) read through the
    device pointer: float4 in id order, stride 16 bytes, w = 0, equal to DownloadState and to the
    oracle's id-order state; the render feed (pbf_sph_export_positions) writes the same xyz packed at
    12 bytes and as float4 at 16 bytes and at a padded 32-byte stride."""
    sc = S.dam_break((16, 16, 16), (64, 32, 64), steps=3, name="export")
    sph = P.SPH(sc.n, sc.grid)
    sph.set_params(P.scene_params(sc))
    sph.upload(sc.pos, sc.vel)
    cfg = O.cfg_for(sc)
    pos, vel, end, vort = sc.pos, sc.vel, None, None
    for _ in range(3):
        sph.Run()
        pos, vel, end, vort = O.step(cfg, pos, vel, end, vort, O.ORDER_JACOBI)
    import ctypes as C
    for fn, want in ((sph._lib.pbf_sph_positions_device, pos), (sph._lib.pbf_sph_velocities_device, vel)):
        ptr, stride = C.c_void_p(), C.c_size_t()
        assert fn(sph._h, C.byref(ptr), C.byref(stride)) == 0 and stride.value == 16
        raw = U.copy_from_device(ptr.value, (sc.n, 4), np.float32, sph.stream())
        assert_bitexact(raw, want, "device buffer against the oracle")
        assert (raw[:, 3] == 0.0).all()
    assert sph.GetPositionBuffer() == C.c_void_p(sph.GetPositionBuffer()).value
    dpos, dvel = sph.download()
    assert_bitexact(dpos, pos, "DownloadState positions")
    assert_bitexact(dvel, vel, "DownloadState velocities")
    assert_bitexact(sph.export_positions(3), np.ascontiguousarray(pos[:, :3]), "packed xyz feed")
    assert_bitexact(sph.export_positions(4), pos, "float4 feed")
    wide = sph.export_positions(3, 32)
    assert_bitexact(np.ascontiguousarray(wide[:, :3]), np.ascontiguousarray(pos[:, :3]), "32-byte stride feed")
    assert (wide[:, 3:] == 0.0).all()
    with pytest.raises(Exception):
        sph.export_positions(3, 8)
    sph.close()


def test_pair_list_overflow_on_the_one_thread_path(gpu_lib):
    """Two compact clumps of 176 particles, each inside a ball of radius 0.45h around a cell corner
    (22 per cell, 44 per row, so no row count wraps), beside a 48x32x48 lattice whose 216-candidate
    interior puts the step's dense threshold at 256 candidates. A clump particle then has 176
    candidates (one-thread path) of which 175 lie within h: more than the 128 entries its pair list
    holds, so the position update rescans its candidates. Bit-exact against the oracle."""
    bg = S.lattice_block((48, 32, 48), 0.5, (16.25, 0.25, 16.25), 0.01, 91)[:73728 - 352]
    u = S.uniform(92, 2 * 176 * 3).reshape(2 * 176, 3)
    d = S.normal(93, 2 * 176 * 3).reshape(2 * 176, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = 0.45 * np.cbrt(u[:, 0:1])
    centre = np.repeat(np.array([[70.0, 10.0, 70.0], [70.0, 10.0, 60.0]]), 176, axis=0)
    clump = np.zeros((352, 4), np.float32)
    clump[:, :3] = (centre + r * d).astype(np.float32)
    pos = np.concatenate([bg, clump])
    sc = S.Scene("clumps", (96, 48, 96), S.config_params(1.0, 0.008, 4), pos, np.zeros_like(pos), 3)
    assert sc.n % 512 == 0
    say(f"[list overflow] n={sc.n}")
    cfg = O.cfg_for(sc)
    # the first step's structure: clump particles have <= 256 candidates, > 128 within h
    srt, _ = O.sort(cfg, O.predict(cfg, sc.pos, sc.vel))
    start, end = O.findcells(cfg, srt, np.zeros(96 * 48 * 96, np.int32))
    nbr = O.neighbourcells(cfg, srt, start, end)[:, :, :3].reshape(sc.n, 9)
    cand = np.maximum(nbr >> 24, 0).sum(1)
    assert (nbr >> 24).max() < 128, "no wrapped row counts"
    ids = srt[:, 3].view(np.int32)
    cl = np.nonzero(ids >= 73728 - 352)[0]
    inr = np.array([(np.linalg.norm(srt[cl, :3] - srt[g, :3], axis=1) <= 1.0).sum() - 1 for g in cl])
    assert (cand[cl] <= 256).all() and (inr > 128).sum() >= 300
    sph = P.SPH(sc.n, sc.grid)
    sph.set_params(P.scene_params(sc))
    sph.upload(sc.pos, sc.vel)
    pos_o, vel_o, end_o, vort_o = sc.pos, sc.vel, None, None
    for k in range(3):
        sph.Run()
        th, count = sph.dense_info()
        if k == 0:
            assert th == 256, f"dense threshold {th}"
        pos_o, vel_o, end_o, vort_o = O.step(cfg, pos_o, vel_o, end_o, vort_o, O.ORDER_JACOBI)
        gp, gv = sph.download()
        assert_bitexact(gp, pos_o, f"clumps step {k} positions")
        assert_bitexact(gv, vel_o, f"clumps step {k} velocities")
    sph.close()
