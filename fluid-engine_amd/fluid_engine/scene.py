"""Scene set-up: the reference's particle generator and parameter defaults, and the
dam-break workloads this engine is measured on.

`generate_volume` follows ParticleGenerator::generate_volume
(src/entities/particle.rs:371-407): counts floor(w/s) x floor(h/s) x floor(d/s) (or given
explicitly), index order x outer / z inner, position origin + r + index * s + one scalar
jitter per particle added to all three axes, mass = rho0 * s^3, all in FP32. The
reference draws the jitter from the unseeded `rand::random`; here it comes from a seeded
SplitMix64 stream so runs are reproducible.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._ffi import FeParams

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """`count` outputs of SplitMix64 started at `seed` (uint64)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    """FP32 uniforms in [0, 1) from the top 24 bits of SplitMix64."""
    return (splitmix64(seed, count) >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)


def generate_volume(origin, width, height, depth, radius, density, spacing, jitter, seed=0,
                    counts=None):
    """Return (positions float32 [N, 3], mass float32)."""
    f = np.float32
    s = f(spacing)
    if counts is None:
        counts = (int(np.floor(f(width) / s)), int(np.floor(f(height) / s)), int(np.floor(f(depth) / s)))
    cx, cy, cz = counts
    off = np.asarray(origin, dtype=np.float32) + f(radius)
    n = cx * cy * cz
    if jitter != 0.0:
        jit = (uniform01(seed, n) - f(0.5)) * f(jitter)
    else:
        jit = np.zeros(n, dtype=np.float32)
    ix, iy, iz = np.meshgrid(np.arange(cx, dtype=np.float32), np.arange(cy, dtype=np.float32),
                             np.arange(cz, dtype=np.float32), indexing="ij")
    pos = np.empty((n, 3), dtype=np.float32)
    pos[:, 0] = (off[0] + ix.reshape(-1) * s) + jit
    pos[:, 1] = (off[1] + iy.reshape(-1) * s) + jit
    pos[:, 2] = (off[2] + iz.reshape(-1) * s) + jit
    mass = f(density) * (s * s * s)
    return pos, mass


def make_params(radius, mass, smoothing_radius, rho0, viscosity, relax, dt, density_iters,
                divergence_iters, gravity, box_min, box_max, grid_res=(128, 128, 128)) -> FeParams:
    """SimulationParams::new (src/entities/particle.rs:333-366)."""
    p = FeParams()
    p.particle_radius = radius
    p.particle_mass = mass
    p.smoothing_radius = smoothing_radius
    p.target_density = rho0
    p.viscosity = viscosity
    p.relax_factor = relax
    p.dt = dt
    p.density_solver_iterations = density_iters
    p.divergence_solver_iterations = divergence_iters
    for k in range(3):
        p.gravity[k] = gravity[k]
        p.box_min[k] = box_min[k]
        p.box_max[k] = box_max[k]
        p.grid_res[k] = grid_res[k]
    return p


@dataclass
class Workload:
    name: str
    positions: np.ndarray
    params: FeParams
    # (amplitude m, frequency Hz) of the box's moving min.x wall, or None for a fixed box
    wall_wave: tuple | None = None

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])


def dam_break(counts, box_max, seed=0, jitter_amp=1e-4, spacing=0.02, rho0=1000.0, viscosity=0.01,
              dt=1e-3, iters=(4, 4), name=None) -> Workload:
    """A block of counts[0] x counts[1] x counts[2] particles at the box's low corner
    (origin = box_min, so the first particle sits at box_min + r), radius = spacing / 2,
    h = 4 r (scene.rs:47), one U(-jitter_amp, jitter_amp) jitter per particle, zero velocity."""
    r = np.float32(spacing) / np.float32(2.0)
    h = r * np.float32(4.0)
    pos, mass = generate_volume((0.0, 0.0, 0.0), 0, 0, 0, r, rho0, spacing, 2.0 * jitter_amp, seed=seed,
                                counts=counts)
    params = make_params(float(r), float(mass), float(h), rho0, viscosity, 0.5, dt, iters[0], iters[1],
                         (0.0, -9.81, 0.0), (0.0, 0.0, 0.0), box_max)
    nm = name or f"dam_break_{counts[0]}x{counts[1]}x{counts[2]}"
    return Workload(nm, pos, params)


def dam_break_32k(seed=0, **kw) -> Workload:
    """32 x 32 x 32 block, spacing 0.02 m, r 0.01 m, h 0.04 m, in a 1 x 1 x 1 m box."""
    return dam_break((32, 32, 32), (1.0, 1.0, 1.0), seed=seed, name="dam_break_32k", **kw)


def dam_break_1m(seed=0, **kw) -> Workload:
    """128 (x) x 128 (y) x 64 (z) block, same spacing and kernel, in a 4 x 4 x 2 m box."""
    return dam_break((128, 128, 64), (4.0, 4.0, 2.0), seed=seed, name="dam_break_1m", **kw)


def block_4m(seed=0, **kw) -> Workload:
    """256 x 128 x 128 block (5.12 x 2.56 x 2.56 m) in a 6 x 3 x 3 m box (2^22 particles)."""
    return dam_break((256, 128, 128), (6.0, 3.0, 3.0), seed=seed, name="block_4m", **kw)


def rest_density_scene() -> Workload:
    """The deterministic set-up of rest_density_diagnostic
    (src/renderer/pipelines/convergence_benchmark.rs:437-481): 20 x 20 x 12 particles from
    spawn (-0.8, 0.5, -0.4), r = 0.02, s = 0.04, h = 0.08, no jitter, viscosity 0.15,
    dt = 5 ms, 4/4 sweeps, box (-1.5, 0, -1) .. (0.8, 4, 1)."""
    r = np.float32(0.02)
    s = r * np.float32(2.0)
    h = r * np.float32(4.0)
    pos, mass = generate_volume((-0.8, 0.5, -0.4), 0.8, 0.8, 0.5, r, 1000.0, s, 0.0)
    params = make_params(float(r), float(mass), float(h), 1000.0, 0.15, 0.5, 0.005, 4, 4, (0.0, -9.81, 0.0),
                         (-1.5, 0.0, -1.0), (0.8, 4.0, 1.0))
    return Workload("rest_density_4800", pos, params)


def app_scene(seed=0) -> Workload:
    """Scene::new (src/core/scene.rs:14-86), the scene the reference app opens: a 1.0 x 2.0 x
    0.8 m block spawned at (-1, 1, -0.8), r = 0.02, s = 0.04 (25 x 50 x 20 = 25000 particles),
    jitter 0.01, h = 0.08, dt = 5 ms, viscosity 0.15, relax 0.5, 4/4 sweeps, box
    (-1.5, 0, -1) .. (0.8, 4, 1) whose min.x wall oscillates with amplitude 0.3 m at 0.5 Hz."""
    r = np.float32(0.02)
    s = r * np.float32(2.0)
    h = r * np.float32(4.0)
    pos, mass = generate_volume((-1.0, 1.0, -0.8), 1.0, 2.0, 0.8, r, 1000.0, s, 0.01, seed=seed)
    params = make_params(float(r), float(mass), float(h), 1000.0, 0.15, 0.5, 0.005, 4, 4, (0.0, -9.81, 0.0),
                         (-1.5, 0.0, -1.0), (0.8, 4.0, 1.0))
    return Workload("app_scene_25000", pos, params, wall_wave=(0.3, 0.5))


def batch_scenes(count=64, seed=0, n_min=4096, n_max=262144, **kw) -> list:
    """Independent dam breaks of uneven sizes: scene k has about N_k = exp(U[ln n_min, ln n_max])
    particles (a block of round(N_k^(1/3))^2 x round(N_k / that^2)), a cubic box of edge
    (N_k / 32768)^(1/3) m, rest density U[950, 1050] and viscosity U[0.005, 0.05] (mass =
    rho0 s^3), all drawn from one seeded stream; spacing, h, dt and sweeps as dam_break."""
    u = uniform01(seed * 7919 + 17, 3 * count).astype(np.float64).reshape(count, 3)
    out = []
    for k in range(count):
        target = float(np.exp(np.log(n_min) + u[k, 0] * (np.log(n_max) - np.log(n_min))))
        side = max(1, int(round(target ** (1.0 / 3.0))))
        depth = max(1, int(round(target / (side * side))))
        n = side * side * depth
        edge = float((n / 32768.0) ** (1.0 / 3.0))
        rho0 = 950.0 + 100.0 * float(u[k, 1])
        nu = 0.005 + 0.045 * float(u[k, 2])
        w = dam_break((side, side, depth), (edge, edge, edge), seed=seed * 1000 + k, rho0=rho0, viscosity=nu,
                      name=f"batch{seed}_scene{k}", **kw)
        out.append(w)
    return out
