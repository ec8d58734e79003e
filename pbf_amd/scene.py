"""Synthetic scenes and reference-compatible parameters.

Restates the scene pattern of the reference's ``Simulation::ResetParticleBuffer``
(src/Simulation.cpp:206-273: jittered lattice blocks, zero velocity, id order
x-major / z / y-fastest) with a seeded splitmix64 generator instead of
``srand(time)``, and the parameter block of ``SPH::SPH`` (src/SPH.cpp:137-144,
``sphparams_t`` src/SPH.h:252-285).

Every workload used by the tests, the golden-vector generator and the benchmark
comes from this module, so one seed names one input everywhere.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

POLY6 = 1.56668147106    # 315/(64 pi), the literal of calclambda.glsl:46
SPIKY = 4.774648292756860  # 15/pi, the literal of calclambda.glsl:54

# Reference wall of updatepos.glsl:98-100: x and z are clamped to [16, GRID-16], y to [0, GRID].
REFERENCE_WALL = (16.0, 0.0, 16.0)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n outputs of the splitmix64 sequence of ``seed`` starting at ``offset``."""
    i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Uniform float64 in [0, 1) with 53 random bits."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Standard normals by Box-Muller over pairs of uniforms."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, offset)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


def wpoly6(r: float, h: float) -> float:
    """``SPH::Wpoly6`` (src/SPH.cpp:159-164) in float32 arithmetic."""
    r = np.float32(r)
    h = np.float32(h)
    if r > h:
        return np.float32(0.0)
    tmp = np.float32(h * h - r * r)
    den = np.float32(h * h * h * h * h * h * h * h * h)
    return np.float32(np.float32(np.float32(np.float32(POLY6) * tmp) * tmp) * tmp / den)


def lattice_rest_density(spacing: float, h: float) -> float:
    """Density of a perfect infinite lattice without the self term (calclambda.glsl:75-82)."""
    k = int(math.ceil(h / spacing)) + 1
    rho = 0.0
    for a in range(-k, k + 1):
        for b in range(-k, k + 1):
            for c in range(-k, k + 1):
                if a == 0 and b == 0 and c == 0:
                    continue
                r = spacing * math.sqrt(a * a + b * b + c * c)
                if r <= h:
                    rho += POLY6 * (h * h - r * r) ** 3 / h ** 9
    return rho


@dataclass
class Params:
    """The eight floats of the reference's parameter block (src/SPH.h:252-285) plus host state."""
    one_over_rho_0: float = 1.0
    epsilon: float = 5.0
    gravity: float = 10.0
    timestep: float = 0.016
    tensile_instability_k: float = 0.1
    tensile_instability_scale: float = float(1.0 / wpoly6(0.2, 2.0))
    xsph_viscosity_c: float = 0.01
    vorticity_epsilon: float = 5.0
    iterations: int = 5                # src/SPH.cpp:26
    vorticity: bool = False            # src/SPH.cpp:25
    extforce: bool = False             # predictpos.glsl:13, SPH::SetExternalForce
    h: float = 2.0                     # src/SPH.cpp:58
    wall_lo: tuple = REFERENCE_WALL
    wall_hi_margin: tuple = REFERENCE_WALL
    gravity_dir: tuple = (0.0, -1.0, 0.0)

    def ubo(self) -> np.ndarray:
        return np.array([self.one_over_rho_0, self.epsilon, self.gravity, self.timestep,
                         self.tensile_instability_k, self.tensile_instability_scale,
                         self.xsph_viscosity_c, self.vorticity_epsilon], dtype=np.float32)


@dataclass
class Scene:
    name: str
    grid: tuple
    params: Params
    pos: np.ndarray          # (N, 4) float32, id order, w = 0
    vel: np.ndarray          # (N, 4) float32, id order, w = 0
    steps: int = 1
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


def config_params(h: float = 1.0, dt: float = 0.016, iterations: int = 3, spacing: float = 0.5) -> Params:
    """Kernel constants shared by the dam-break workloads: g 9.81, eps 0.01, s_corr k 0.1 n 4
    dq 0.2h, XSPH c 0.01, vorticity eps 0.0005, rest density of the 0.5h lattice."""
    rho0 = lattice_rest_density(spacing * h, h)
    return Params(one_over_rho_0=float(np.float32(1.0 / rho0)), epsilon=0.01, gravity=9.81, timestep=dt,
                  tensile_instability_k=0.1, tensile_instability_scale=float(np.float32(1.0 / wpoly6(0.2 * h, h))),
                  xsph_viscosity_c=0.01, vorticity_epsilon=0.0005, iterations=iterations, vorticity=True,
                  extforce=False, h=h)


def lattice_block(dims, spacing: float, corner, jitter: float, seed: int, offset: int = 0,
                  mirror: bool = False) -> np.ndarray:
    """Jittered lattice in the reference's id order (x outer, then z, y fastest;
    src/Simulation.cpp:216-229). Jitter is U(-jitter, +jitter) per axis."""
    nx, ny, nz = dims
    x, z, y = np.meshgrid(np.arange(nx), np.arange(nz), np.arange(ny), indexing="ij")
    idx = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.float64)
    if mirror:
        idx[:, 0] *= -1
        idx[:, 2] *= -1
    n = idx.shape[0]
    j = (uniform(seed, 3 * n, offset).reshape(n, 3) * 2.0 - 1.0) * jitter
    p = np.asarray(corner, dtype=np.float64)[None, :] + spacing * idx + j
    out = np.zeros((n, 4), dtype=np.float32)
    out[:, :3] = p.astype(np.float32)
    return out


def dam_break(n_side=(32, 32, 32), grid=(64, 64, 64), h=1.0, dt=0.016, iterations=3, seed=0,
              steps=200, name="dam_break") -> Scene:
    """Plain dam break: one lattice block at spacing 0.5h with seeded U(+-0.01h) jitter,
    zero velocity, corner inside the reference wall at (16.25, 0.25, 16.25) h."""
    pos = lattice_block(n_side, 0.5 * h, (16.25 * h, 0.25 * h, 16.25 * h), 0.01 * h, seed)
    vel = np.zeros_like(pos)
    return Scene(name, tuple(grid), config_params(h, dt, iterations), pos, vel, steps,
                 {"seed": seed, "dims": tuple(n_side)})


def cfg_dam_32k(seed=0) -> Scene:
    """32,768 particles, 32^3 lattice, grid 64^3, 3 iterations, dt 0.016, 200 steps."""
    return dam_break((32, 32, 32), (64, 64, 64), 1.0, 0.016, 3, seed, 200, "dam_32k")


def cfg_block_1m(seed=1) -> Scene:
    """1,048,576 particles as a 128x64x128 block, grid 256x128x256, 4 iterations, dt 0.008."""
    return dam_break((128, 64, 128), (256, 128, 256), 1.0, 0.008, 4, seed, 500, "block_1m")


def cfg_stress_4m(seed=2, dt=0.008) -> Scene:
    """4,194,304 particles: a 128x64x256 dam block plus as many uniform in the upper half of
    the walled domain, velocities N(0, 0.5) per axis; grid 384x192x384, 4 iterations."""
    grid = (384, 192, 384)
    h = 1.0
    block = lattice_block((128, 64, 256), 0.5, (16.25, 0.25, 16.25), 0.01, seed)
    m = block.shape[0]
    u = uniform(seed + 1, 3 * m).reshape(m, 3)
    lo = np.array([16.0, grid[1] / 2.0, 16.0])
    hi = np.array([grid[0] - 16.0, float(grid[1]), grid[2] - 16.0])
    rnd = np.zeros((m, 4), dtype=np.float32)
    rnd[:, :3] = (lo + u * (hi - lo)).astype(np.float32)
    pos = np.concatenate([block, rnd])
    n = pos.shape[0]
    vel = np.zeros_like(pos)
    vel[:, :3] = (0.5 * normal(seed + 2, 3 * n).reshape(n, 3)).astype(np.float32)
    return Scene("stress_4m", grid, config_params(h, dt, 4), pos, vel, 200, {"seed": seed})


def cfg_adjoint_16k(seed=4, dt=0.016) -> Scene:
    """16,384 particles in a 32x16x32 lattice, grid 64^3, v0 ~ N(0, 0.3), 32 steps x 3 iterations."""
    pos = lattice_block((32, 16, 32), 0.5, (16.25, 0.25, 16.25), 0.01, seed)
    n = pos.shape[0]
    vel = np.zeros_like(pos)
    vel[:, :3] = (0.3 * normal(seed + 1, 3 * n).reshape(n, 3)).astype(np.float32)
    return Scene("adjoint_16k", (64, 64, 64), config_params(1.0, dt, 3), pos, vel, 32, {"seed": seed})


def batched_scenes(nscenes: int = 256, seed: int = 3, n_lo: int = 2048, n_hi: int = 65536, dt: float = 0.016,
                   iterations: int = 3, steps: int = 100) -> list:
    """Independent scenes of the batched form. Per scene: n log-uniform in
    [n_lo, n_hi], rounded to a multiple of 512 (the plain solver's dispatch
    size); a jittered 0.5h lattice block of ceil(cbrt n) x ceil(n / ceil(cbrt n)^2)
    x ceil(cbrt n) sites (the first n in id order), centred in x and z on the
    floor of a cubic box of side ceil(cbrt(n) 0.5h) 2 cells; walls one cell in
    from the x and z faces; gravity (0, -1, 0) turned by N(0, 0.1) rad about x
    and about z; XSPH c uniform in [0.005, 0.05]. Other constants as the
    dam-break workloads, 3 iterations, dt 0.016."""
    u = uniform(seed, 4 * nscenes).reshape(nscenes, 4)
    g = normal(seed + 1, 2 * nscenes).reshape(nscenes, 2) * 0.1
    out = []
    for s in range(nscenes):
        n = math.exp(math.log(n_lo) + u[s, 0] * (math.log(n_hi) - math.log(n_lo)))
        n = int(min(max(round(n / 512.0) * 512, n_lo), n_hi))
        side = int(math.ceil(n ** (1.0 / 3.0) * 0.5)) * 2
        k = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
        ky = int(math.ceil(n / float(k * k)))
        corner = ((side - k * 0.5) / 2.0 + 0.25, 0.25, (side - k * 0.5) / 2.0 + 0.25)
        pos = lattice_block((k, ky, k), 0.5, corner, 0.01, seed + 10 + s)
        # first n sites in id order (x outer, z, y fastest) keeps whole columns
        pos = pos[:n]
        p = config_params(1.0, dt, iterations)
        ax, az = float(g[s, 0]), float(g[s, 1])
        # (0, -1, 0) turned by az about z, then by ax about x
        gx, gy = math.sin(az), -math.cos(az)
        gdir = (gx, gy * math.cos(ax), gy * math.sin(ax))
        p.gravity_dir = tuple(float(np.float32(x)) for x in gdir)
        p.xsph_viscosity_c = float(np.float32(0.005 + u[s, 1] * 0.045))
        p.wall_lo = (1.0, 0.0, 1.0)
        p.wall_hi_margin = (1.0, 0.0, 1.0)
        out.append(Scene(f"batch{s}", (side, side, side), p, pos, np.zeros_like(pos), steps,
                         {"seed": seed, "index": s}))
    return out


def count_sortbits(v: int) -> int:
    """``count_sortbits`` of src/RadixSort.cpp:24-30: bit length of v, at least 1."""
    r = 1
    while v >> 1:
        v >>= 1
        r += 1
    return r


def key_bits(grid) -> int:
    """Bits the reference's 2-bit LSD passes cover: 2 * ceil(numbits / 2) (src/RadixSort.cpp:127)."""
    nb = count_sortbits(int(grid[0]) * int(grid[1]) * int(grid[2]) - 1)
    return 2 * ((nb + 1) >> 1)
