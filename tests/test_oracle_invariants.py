"""The oracle against what the mathematics fixes, independently of the golden
fixtures: closed forms (stable sort by the cell key, the cell tables of a
sorted array, the perfect-lattice density and lambda = 0), brute force (the
neighbour set of the packed table equals the particles of the 27 surrounding
cells), invariants (momentum of the race-free position update, a uniform
velocity field passes the XSPH/vorticity pass unchanged), a physical sign
(a compressed pair separates) and the FP64 replay against the FP32 step.

Each check is chosen so that a plausible mistake fails it: a self term in the
density, a dropped lambda_j, a transposed key weight, an off-by-one in the cell
ranges, a flipped gradient sign, a missing term of the FP64 restatement."""
import numpy as np

import oracle as O
from pbf_amd import scene as S

H = 1.0


def records(xyz):
    n = xyz.shape[0]
    rec = np.zeros((n, 4), np.float32)
    rec[:, :3] = xyz
    rec[:, 3] = np.arange(n, dtype=np.int32).view(np.float32)
    return rec


def params(rho0, eps=0.01, k=0.1):
    p = S.config_params(h=H)
    p.one_over_rho_0 = float(np.float32(1.0 / rho0))
    p.epsilon = eps
    p.tensile_instability_k = k
    return p


def free_cfg(n, grid, p):
    """Walls at the grid faces (nothing inside is clamped)."""
    return O.make_cfg(n, grid, p, wall_lo=(-1e6, -1e6, -1e6), wall_hi=(1e6, 1e6, 1e6))


def lattice(side, spacing, corner, jitter=0.0, seed=0):
    a = np.arange(side, dtype=np.float64) * spacing
    x, y, z = np.meshgrid(a, a, a, indexing="ij")
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], 1) + np.asarray(corner, np.float64)
    if jitter:
        xyz += (S.uniform(seed, xyz.size).reshape(xyz.shape) * 2 - 1) * jitter
    return xyz.astype(np.float32)


def with_sentinels(xyz, grid):
    """Prepends a lone particle in cell (0, 0, 0), the lowest key, whose cell the
    reference's findcells leaves invisible (gid 0 writes start[(0,0,0)] = 0 and
    nothing else), and appends one in the highest cell, whose end entry is never
    written by the step; both lie far from the others."""
    lo = np.array([[0.5, 0.5, 0.5]], np.float32)
    hi = np.array([[grid[0] - 0.5, grid[1] - 0.5, grid[2] - 0.5]], np.float32)
    return np.concatenate([lo, xyz, hi])


def cell_of(xyz, grid):
    g = np.asarray(grid, np.float32)
    return np.clip(xyz, 0, g).astype(np.int64)  # clamp inclusive of the grid size, then truncate


# ------------------------------------------------------------------ sort + keys
def test_sort_is_stable_sort_by_exact_cell_key():
    """Keys below 2^24 are exact integers x + y gx gz + z gx of the clamped cell
    (counting.glsl:53-57); the sort is the stable sort of the records by that
    key (src/RadixSort.cpp:124-133 over all key bits)."""
    grid = (24, 10, 18)
    rng = np.random.default_rng(7)
    n = 3000
    xyz = rng.uniform(-2.0, [26.0, 12.0, 20.0], size=(n, 3)).astype(np.float32)
    xyz[::7] = xyz[3]  # many equal cells
    rec = records(xyz)
    cfg = O.make_cfg(n, grid, params(1.0))
    c = cell_of(xyz, grid)
    want_keys = (c[:, 0] + c[:, 1] * grid[0] * grid[2] + c[:, 2] * grid[0]) & ((1 << cfg.key_bits) - 1)
    assert np.array_equal(O.keys(cfg, rec).astype(np.int64), want_keys)
    out, perm = O.sort(cfg, rec)
    want = np.argsort(want_keys, kind="stable")
    assert np.array_equal(perm, want)
    assert np.array_equal(out.view(np.int32), rec[want].view(np.int32))


# ------------------------------------------------------------------ cell tables
def test_findcells_closed_form_and_persistent_end():
    """start[c] = first sorted index of cell c, end[c] = one past its last,
    except the quirks findcells.glsl:34-53 has: the lowest cell's start is
    written only by gid 0, at (0,0,0); the last cell's end keeps its old value;
    cells outside the grid are dropped. Everything not written keeps -1 (start,
    cleared every call) or its previous value (end, never cleared)."""
    grid = (9, 7, 8)
    rng = np.random.default_rng(11)
    n = 700
    xyz = rng.uniform(0.0, [9.0, 7.0, 8.0], size=(n, 3)).astype(np.float32)
    cfg = O.make_cfg(n, grid, params(1.0))
    srt, _ = O.sort(cfg, records(xyz))
    c = cell_of(srt[:, :3], grid)
    lin = c[:, 0] + grid[0] * (c[:, 1] + grid[1] * c[:, 2])
    old_end = rng.integers(-5, 5, grid[0] * grid[1] * grid[2]).astype(np.int32)
    start, end = O.findcells(cfg, srt, old_end)
    want_start = np.full_like(start, -1)
    want_end = old_end.copy()
    want_start[0] = 0
    for g in range(1, n):
        if lin[g] != lin[g - 1]:
            want_start[lin[g]] = g
            want_end[lin[g - 1]] = g
    assert np.array_equal(start, want_start)
    assert np.array_equal(end, want_end)


def test_neighbour_table_visits_the_27_cells_brute_force():
    """Iterating the packed table as FOR_EACH_NEIGHBOUR does
    (foreachneighbour.glsl:1-10: count = d >> 24, start = d & 0xFFFFFF, skip
    self) visits exactly the particles whose cells differ from the particle's
    by at most one in every axis, found by brute force over all pairs."""
    grid = (12, 12, 12)
    body = lattice(8, 0.55, (3.1, 3.1, 3.1), jitter=0.2, seed=3)
    xyz = with_sentinels(body, grid)
    n = xyz.shape[0]
    cfg = O.make_cfg(n, grid, params(1.0))
    srt, _ = O.sort(cfg, records(xyz))
    start, end = O.findcells(cfg, srt, np.zeros(grid[0] * grid[1] * grid[2], np.int32))
    nbr = O.neighbourcells(cfg, srt, start, end)
    c = cell_of(srt[:, :3], grid)
    ids = srt[:, 3].view(np.int32)
    for i in range(n):
        if ids[i] in (0, n - 1):
            continue  # the two sentinels
        got = []
        for d in nbr[i, :, :3].ravel():
            cnt, s0 = int(d) >> 24, int(d) & 0xFFFFFF
            got += [j for j in range(s0, s0 + cnt) if j != i]
        want = np.nonzero(np.all(np.abs(c - c[i]) <= 1, axis=1))[0]
        want = want[want != i]
        assert sorted(got) == list(want), f"particle {i}"
        assert got == sorted(got), "rows are visited in ascending sorted order"


# ------------------------------------------------------------------ lambda / delta p
def test_lattice_rest_density_closed_form():
    """Sum of W_poly6 over a perfect 0.5h lattice without the self term."""
    assert abs(S.lattice_rest_density(0.5, 1.0) - 6.511519864097) < 1e-9


def lattice_setup(jitter=0.0):
    grid = (24, 24, 24)
    body = lattice(14, 0.5, (5.25, 5.25, 5.25), jitter=jitter, seed=5)
    xyz = with_sentinels(body, grid)
    n = xyz.shape[0]
    cfg = free_cfg(n, grid, params(S.lattice_rest_density(0.5, 1.0)))
    srt, _ = O.sort(cfg, records(xyz))
    start, end = O.findcells(cfg, srt, np.zeros(grid[0] * grid[1] * grid[2], np.int32))
    nbr = O.neighbourcells(cfg, srt, start, end)
    ids = srt[:, 3].view(np.int32)
    return cfg, srt, nbr, ids, n


def test_perfect_lattice_interior_lambda_and_delta_are_zero():
    """Interior of a perfect lattice at rest density: rho = rho0 so lambda = 0
    (calclambda.glsl:98-102), and the position update of particles whose
    neighbours are interior too is a sum of opposite pairs, so it vanishes
    (updatepos.glsl:57-66). A self term in rho or a one-sided gradient sum fails."""
    cfg, srt, nbr, ids, n = lattice_setup()
    lam = O.lambdas(cfg, srt, nbr)
    q = ids - 1
    ijk = np.stack([q // 196, (q // 14) % 14, q % 14], 1)  # lattice index (x, y, z) of body particles
    body = (ids > 0) & (ids < n - 1)
    deep2 = body & np.all((ijk >= 2) & (ijk <= 11), axis=1)
    deep4 = body & np.all((ijk >= 4) & (ijk <= 9), axis=1)
    assert deep4.sum() == 216
    assert np.abs(lam[deep2]).max() < 1e-5
    assert np.abs(lam[body & ~deep2]).max() > 1e-2  # the surface is below rest density
    out = O.updatepos(cfg, srt, lam, nbr, O.ORDER_JACOBI)
    dp = (out[:, :3].astype(np.float64) - srt[:, :3].astype(np.float64))
    assert np.abs(dp[deep4]).max() < 1e-5
    assert np.abs(dp[body]).max() > 1e-3


def test_position_update_conserves_momentum():
    """lambda_i + lambda_j + s_corr(r_ij) is symmetric and grad W antisymmetric,
    so the race-free updates sum to zero when neighbour sets are symmetric and
    nothing is clamped. A dropped lambda_j or a wrong gradient index fails."""
    cfg, srt, nbr, ids, n = lattice_setup(jitter=0.05)
    lam = O.lambdas(cfg, srt, nbr)
    out = O.updatepos(cfg, srt, lam, nbr, O.ORDER_JACOBI)
    dp = out[:, :3].astype(np.float64) - srt[:, :3].astype(np.float64)
    body = (ids > 0) & (ids < n - 1)
    total = np.abs(dp[body]).sum()
    assert total > 1.0
    assert np.abs(dp[body].sum(0)).max() < 1e-5 * total + 1e-4
    assert np.all(dp[~body] == 0)  # the sentinels have no neighbours


def test_compressed_pair_separates():
    """Two particles closer than the rest spacing (rho > rho0): lambda < 0 and
    the update pushes them apart along their axis (calclambda.glsl:98-102,
    updatepos.glsl:57-66 with the spiky gradient of calclambda.glsl:57-64)."""
    grid = (16, 16, 16)
    xyz = with_sentinels(np.array([[8.0, 8.0, 8.0], [8.3, 8.0, 8.0]], np.float32), grid)
    w = float(S.wpoly6(0.3, 1.0))
    cfg = free_cfg(4, grid, params(0.5 * w, eps=0.01, k=0.0))
    srt, _ = O.sort(cfg, records(xyz))
    start, end = O.findcells(cfg, srt, np.zeros(16 ** 3, np.int32))
    nbr = O.neighbourcells(cfg, srt, start, end)
    lam = O.lambdas(cfg, srt, nbr)
    ids = srt[:, 3].view(np.int32)
    a, b = int(np.nonzero(ids == 1)[0][0]), int(np.nonzero(ids == 2)[0][0])
    assert lam[a] < 0 and lam[b] < 0 and lam[a] == lam[b]
    out = O.updatepos(cfg, srt, lam, nbr, O.ORDER_JACOBI)
    dxa, dxb = out[a, 0] - srt[a, 0], out[b, 0] - srt[b, 0]
    assert dxa < 0 < dxb and abs(dxa + dxb) < 1e-6
    assert np.all(out[[a, b], 1:3] == srt[[a, b], 1:3])


# ------------------------------------------------------------------ XSPH / vorticity
def test_uniform_velocity_field_is_unchanged():
    """v_ij = 0 for every pair, so the XSPH sum, omega and eta vanish and
    N = eta/|eta| takes its zero guard (vorticity.glsl:34-86)."""
    cfg, srt, nbr, ids, n = lattice_setup(jitter=0.05)
    cfg.vorticity = 1
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = (0.7, -1.3, 0.25)
    v2, om = O.vorticity(cfg, srt, nbr, vel, np.zeros(n, np.float32), O.ORDER_JACOBI)
    assert np.array_equal(v2.view(np.int32), vel.view(np.int32))
    assert np.all(om == 0)


# ------------------------------------------------------------------ FP64 replay
def test_fp64_replay_tracks_the_fp32_step():
    """The FP64 restatement of a whole step stays within FP32 rounding of the
    FP32 step on a stable scene (the reference's own parameters); replaying
    with the FP32 step's sort order and neighbour table gives the same FP64
    result as the FP64 step's own (its structure agrees)."""
    sc = S.lattice_block((12, 12, 12), 0.94, (40.5, 0.5, 40.5), 0.005, 31)
    p = S.Params()
    p.vorticity = True
    scene = S.Scene("ref12", (128, 64, 128), p, sc, np.zeros_like(sc), 1)
    cfg = O.cfg_for(scene)
    pos, vel, _, _, st = O.step(cfg, scene.pos, scene.vel, stages=True)
    pd, vd, _, _ = O.step_f64(cfg, scene.pos, scene.vel)
    assert np.abs(pos[:, :3] - pd[:, :3]).max() < 1e-4 * p.h
    assert np.abs(vel[:, :3] - vd[:, :3]).max() < 1e-2
    moved = np.abs(pd[:, :3] - scene.pos[:, :3]).max()
    assert moved > 1e-2  # the step did something
    perm = st["sorted"][:, 3].view(np.int32).copy()
    pf, vf = O.step_fixed_f64(cfg, scene.pos, scene.vel, perm, st["nbr"])
    assert np.abs(pf - pd).max() < 1e-9
    assert np.abs(vf - vd).max() < 1e-7
