"""The two Jacobi sweep variants on the GPU, through the C ABI: the staged-window sweeps
(pressure_force / pressure_update tuned to items 2: one 64-particle window per wave, the
window's neighbour union staged in LDS, 16-bit lists) give the same fields, grid tables,
stats and sweep counts bit for bit as the per-particle list sweeps (items 1), which are
themselves pinned to the CPU oracle (test_gpu_parity.py). Covered: fixed and converging
solves, every window block size, the 1M dam break (where no window falls back), a batch
whose windows straddle scene boundaries, and the staged sweeps against the oracle directly.
"""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

FIELDS = ("position_a", "velocity_a", "velocity_b", "position_b", "densities", "factors", "source_terms",
          "pressures", "pressure_accelerations", "colors")


def opts(mode, iters=100):
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_RADIX
    o.solver_mode = mode
    o.density_tol = 1e-3
    o.divergence_tol = 1e-2
    o.max_iterations = iters
    return o


def same(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, name
    va = a.view(np.uint32) if a.dtype == np.float32 else a
    vb = b.view(np.uint32) if b.dtype == np.float32 else b
    if not np.array_equal(va, vb):
        bad = np.nonzero(va.reshape(-1) != vb.reshape(-1))[0]
        raise AssertionError(f"{name}: {len(bad)} of {va.size} differ, first at {bad[:5].tolist()}")


def staged(sim, block):
    sim.set_tune("pressure_force", block, 2)
    sim.set_tune("pressure_update", block, 2)


def listed(sim):
    sim.set_tune("pressure_force", 256, 1)
    sim.set_tune("pressure_update", 256, 1)


@pytest.mark.parametrize("mode", [fe.FE_SOLVER_FIXED, fe.FE_SOLVER_CONVERGE])
@pytest.mark.parametrize("block", [64, 128, 256])
def test_staged_matches_list_32k(mode, block):
    w = scene.dam_break_32k(seed=7)
    frames = 5
    with fe.Simulation(w.params, w.positions, device=0, options=opts(mode)) as a, \
            fe.Simulation(w.params, w.positions, device=0, options=opts(mode)) as b:
        listed(a)
        staged(b, block)
        a.run(frames)
        b.run(frames)
        for name in FIELDS + ("grid_entries", "grid_start"):
            same(name, a.download(name), b.download(name))
        same("stats", a.read_stats(), b.read_stats())
        assert a.iteration_totals() == b.iteration_totals()


def test_staged_matches_oracle(oracle_mod):
    w = scene.rest_density_scene()
    o = opts(fe.FE_SOLVER_FIXED)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as s:
        staged(s, 128)
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            for _ in range(3):
                s.step()
                orc.step()
            for name in ("position_a", "velocity_a", "densities", "pressures", "pressure_accelerations"):
                same(f"{name} vs oracle", s.download(name), orc.download(name))
        finally:
            orc.close()


def test_staged_matches_list_1m():
    w = scene.dam_break_1m()
    o = opts(fe.FE_SOLVER_CONVERGE)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as a, \
            fe.Simulation(w.params, w.positions, device=0, options=o) as b:
        listed(a)
        staged(b, 256)
        a.run(3)
        b.run(3)
        for name in ("position_a", "velocity_a", "pressures", "pressure_accelerations", "source_terms"):
            same(name, a.download(name), b.download(name))
        assert a.iteration_totals() == b.iteration_totals()


def test_staged_batch_matches_list():
    ws = scene.batch_scenes(count=6, seed=11, n_min=3000, n_max=20000)   # scene sizes not multiples of 64
    o = opts(fe.FE_SOLVER_CONVERGE)
    a = fe.Batch(ws, device=0, options=o)
    b = fe.Batch(ws, device=0, options=o)
    try:
        listed(a)
        staged(b, 128)
        a.run(4)
        b.run(4)
        for name in FIELDS:
            same(name, a.download(name), b.download(name))
        for k in range(len(ws)):
            assert a.scene_iterations(k) == b.scene_iterations(k), k
    finally:
        a.close()
        b.close()
