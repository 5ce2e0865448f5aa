"""Frames at edge particle counts and layouts, on the GPU through the C ABI, equal the oracle's
bit for bit: a single particle (T = 1, no sort pass runs), two and three particles, counts at
and just past one 64-lane window (64, 65), odd counts that leave ragged sort tiles, windows and
padded hash entries (1001, 4335), a sparse scatter where most particles have no neighbour and
most buckets are empty, and a cluster with coincident particles (the kernel gradient is zero for
r < 1e-6, common.glsl:58). Each runs in both sorts, in fixed and in converge mode.
"""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

F = np.float32
FIELDS = ("position_a", "velocity_a", "velocity_b", "position_b", "densities", "factors", "source_terms",
          "pressures", "pressure_accelerations", "colors", "grid_entries", "grid_start")


def opts(sort, mode):
    o = fe.default_options()
    o.sort_algorithm = sort
    o.solver_mode = mode
    o.density_tol = 1e-3
    o.divergence_tol = 1e-2
    o.max_iterations = 100
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


def sparse(n=777, seed=3):
    """n particles uniform over the 1 m box (h = 0.04): most have no neighbour."""
    w = scene.dam_break((1, 1, 1), (1.0, 1.0, 1.0))
    r = float(w.params.particle_radius)
    u = scene.uniform01(seed, 3 * n).reshape(n, 3)
    pos = (F(r) + u * F(1.0 - 2 * r)).astype(np.float32)
    return scene.Workload(f"sparse_{n}", pos, w.params)


def cluster():
    """A 3 x 3 x 3 block with two more particles on top of existing ones (exactly coincident)."""
    w = scene.dam_break((3, 3, 3), (1.0, 1.0, 1.0), jitter_amp=0.0)
    pos = np.concatenate([w.positions, w.positions[[4, 13]]]).astype(np.float32)
    return scene.Workload("cluster_29", pos, w.params)


CASES = {
    "single": lambda: scene.dam_break((1, 1, 1), (1.0, 1.0, 1.0)),
    "pair": lambda: scene.dam_break((1, 1, 2), (1.0, 1.0, 1.0)),
    "three": lambda: scene.dam_break((3, 1, 1), (1.0, 1.0, 1.0)),
    "window_64": lambda: scene.dam_break((4, 4, 4), (1.0, 1.0, 1.0)),
    "window_65": lambda: scene.dam_break((5, 13, 1), (1.0, 1.0, 1.0)),
    "ragged_1001": lambda: scene.dam_break((7, 11, 13), (1.0, 1.0, 1.0), seed=1),
    "ragged_4335": lambda: scene.dam_break((17, 15, 17), (1.0, 1.0, 1.0), seed=2),
    "sparse_777": sparse,
    "cluster_29": cluster,
}


@pytest.mark.parametrize("mode", [fe.FE_SOLVER_FIXED, fe.FE_SOLVER_CONVERGE])
@pytest.mark.parametrize("sort", [fe.FE_SORT_RADIX, fe.FE_SORT_BITONIC])
@pytest.mark.parametrize("case", list(CASES))
def test_edge_frames_match_oracle(case, sort, mode, oracle_mod):
    w = CASES[case]()
    o = opts(sort, mode)
    print(f"[gpu edge] case={case} n={w.n} sort={sort} mode={mode}", flush=True)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as sim:
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            for frame in range(3):
                sim.step()
                orc.step()
                for name in FIELDS:
                    same(f"{case} frame {frame} {name}", sim.download(name), orc.download(name))
                same(f"{case} frame {frame} stats", sim.read_stats(), orc.read_stats())
                assert sim.last_iterations() == orc.last_iterations(), frame
            assert np.isfinite(sim.download("position_a")).all()
        finally:
            orc.close()
