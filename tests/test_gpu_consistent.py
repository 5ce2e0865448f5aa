"""The consistent divergence order on the GPU, through the C ABI: frames from a moving state
(where the post-integrate sort permutes most particles) equal the oracle's consistent-order
frames bit for bit, in fixed and converge modes; batch scenes equal single contexts; the
autotuner leaves the next frame unchanged; the stand-alone integration refuses to run before
a neighbour search has produced the permutation it scatters through.
"""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

FIELDS = ("position_a", "velocity_a", "velocity_b", "position_b", "densities", "factors", "source_terms",
          "pressures", "pressure_accelerations", "colors", "grid_entries", "grid_start")


def opts(order, mode=fe.FE_SOLVER_FIXED):
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_RADIX
    o.solver_mode = mode
    o.density_tol = 1e-3
    o.divergence_tol = 1e-2
    o.max_iterations = 100
    o.divergence_order = order
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


@pytest.fixture(scope="module")
def moving():
    """The 32k dam break after 40 reference frames on the GPU: (workload, position_a, velocity_a)."""
    w = scene.dam_break_32k(seed=3)
    with fe.Simulation(w.params, w.positions, device=0, options=opts(fe.FE_DIVERGENCE_REFERENCE)) as s:
        s.run(40)
        return w, s.download("position_a"), s.download("velocity_a")


@pytest.mark.parametrize("mode", [fe.FE_SOLVER_FIXED, fe.FE_SOLVER_CONVERGE])
def test_consistent_frames_match_oracle(moving, mode, oracle_mod):
    w, pos, vel = moving
    o = opts(fe.FE_DIVERGENCE_CONSISTENT, mode)
    print(f"[gpu consistent] mode={mode} n={w.n}", flush=True)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as sim:
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            for t in (sim, orc):
                t.upload("position_a", pos)
                t.upload("velocity_a", vel)
            for frame in range(4):
                sim.step()
                orc.step()
                for name in FIELDS:
                    same(f"frame {frame} {name}", sim.download(name), orc.download(name))
                same(f"frame {frame} stats", sim.read_stats(), orc.read_stats())
                assert sim.last_iterations() == orc.last_iterations(), frame
            moved = np.mean(sim.download("grid_entries")[: w.n, 1] != np.arange(w.n))
            assert moved > 0.01
        finally:
            orc.close()


def test_consistent_differs_from_reference(moving):
    """Where the order changes the two orders give different velocities (the mode is live)."""
    w, pos, vel = moving
    out = []
    for order in (fe.FE_DIVERGENCE_REFERENCE, fe.FE_DIVERGENCE_CONSISTENT):
        with fe.Simulation(w.params, w.positions, device=0, options=opts(order)) as s:
            s.upload("position_a", pos)
            s.upload("velocity_a", vel)
            s.step()
            out.append(s.download("velocity_a"))
    assert not np.array_equal(out[0], out[1])


def test_consistent_batch_matches_single_contexts():
    ws = scene.batch_scenes(count=4, seed=5, n_min=3000, n_max=12000)
    o = opts(fe.FE_DIVERGENCE_CONSISTENT, fe.FE_SOLVER_CONVERGE)
    b = fe.Batch(ws, device=0, options=o)
    try:
        b.run(12)
        for k, wk in enumerate(ws):
            with fe.Simulation(wk.params, wk.positions, device=0, options=o) as s:
                s.run(12)
                for name in ("position_a", "velocity_a", "velocity_b", "pressures", "colors"):
                    same(f"scene {k} {name}", b.download_scene(k, name), s.download(name))
    finally:
        b.close()


def test_consistent_autotune_keeps_the_next_frame(moving):
    w, pos, vel = moving
    o = opts(fe.FE_DIVERGENCE_CONSISTENT)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as a, \
            fe.Simulation(w.params, w.positions, device=0, options=o) as b:
        for s in (a, b):
            s.upload("position_a", pos)
            s.upload("velocity_a", vel)
            s.step()
        a.autotune_now(reps=3)
        a.step()
        b.step()
        for name in ("position_a", "velocity_a", "velocity_b", "colors", "pressures"):
            same(name, a.download(name), b.download(name))


def test_stage_integration_needs_a_search():
    w = scene.rest_density_scene()
    with fe.Simulation(w.params, w.positions, device=0, options=opts(fe.FE_DIVERGENCE_CONSISTENT)) as s:
        with pytest.raises(fe.FluidEngineError):
            s.divergence_integration()
        s.neighbor_search()
        s.divergence_integration()
        s.sync()
