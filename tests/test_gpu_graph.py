"""Single-scene frames replayed from a captured HIP graph (fe_use_graph), through the C ABI.
The reference records each physics frame into one command buffer (src/renderer/mod.rs:
279-363); here the steady-state FIXED frame is captured once and replayed. Every replay must
equal the eager frame bit for bit: all fields, the grid tables, the stats and the moving
wall. Each replay is checked on its own, with downloads (eager work and synchronisation)
between replays. Covered: the reference app's scene with its oscillating wall, the 32k dam
break against the oracle, a parameter change (which drops the graph), and converge mode
(which runs eagerly)."""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

FIELDS = ("position_a", "velocity_a", "velocity_b", "position_b", "densities", "factors", "source_terms",
          "pressures", "pressure_accelerations", "colors")


def same(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, name
    va = a.view(np.uint32) if a.dtype == np.float32 else a
    vb = b.view(np.uint32) if b.dtype == np.float32 else b
    if not np.array_equal(va, vb):
        bad = np.nonzero(va.reshape(-1) != vb.reshape(-1))[0]
        raise AssertionError(f"{name}: {len(bad)} of {va.size} differ, first at {bad[:5].tolist()}")


def compare(a, b, label, fields=FIELDS):
    for name in fields + ("grid_entries", "grid_start"):
        same(f"{label} {name}", a.download(name), b.download(name))
    same(f"{label} stats", a.read_stats(), b.read_stats())
    assert a.iteration_totals() == b.iteration_totals(), label


def app_pair():
    w = scene.app_scene()
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_RADIX
    sims = []
    for _ in range(2):
        s = fe.Simulation(w.params, w.positions, device=0, options=o)
        s.set_wall(fe.CollisionBox.from_params(w.params, *w.wall_wave))
        sims.append(s)
    return w, sims


def test_app_scene_graph_frames_equal_eager():
    w, (eager, graph) = app_pair()
    try:
        graph.use_graph(True)
        for f in range(12):
            eager.step()
            graph.step()
            compare(eager, graph, f"frame {f}")
            we, wg = eager.wall(), graph.wall()
            same(f"frame {f} wall", np.array(we.min, np.float32), np.array(wg.min, np.float32))
        # many replays without checks in between, then one check
        eager.run(40)
        graph.run(40)
        compare(eager, graph, "after 52 frames")
    finally:
        eager.close()
        graph.close()


def test_graph_dropped_on_parameter_change():
    w, (eager, graph) = app_pair()
    try:
        graph.use_graph(True)
        eager.run(5)
        graph.run(5)
        for s in (eager, graph):
            p = fe.FeParams.from_buffer_copy(s.params)
            p.viscosity = 0.05
            p.density_solver_iterations = 3
            s.set_params(p)
        for f in range(4):
            eager.step()
            graph.step()
            compare(eager, graph, f"after change, frame {f}")
        graph.use_graph(False)
        eager.step()
        graph.step()
        compare(eager, graph, "graph off")
    finally:
        eager.close()
        graph.close()


def test_graph_frames_match_oracle(oracle_mod):
    w = scene.dam_break_32k(seed=2)
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_BITONIC
    with fe.Simulation(w.params, w.positions, device=0, options=o) as s:
        s.use_graph(True)
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            for f in range(4):
                s.step()
                orc.step()
                for name in FIELDS + ("grid_entries", "grid_start"):
                    same(f"frame {f} {name} vs oracle", s.download(name), orc.download(name))
                same(f"frame {f} stats vs oracle", s.read_stats(), orc.read_stats())
        finally:
            orc.close()


def test_converge_mode_runs_eagerly_under_graph_flag():
    w = scene.dam_break_32k(seed=5)
    o = fe.default_options()
    o.solver_mode = fe.FE_SOLVER_CONVERGE
    o.density_tol = 1e-3
    o.divergence_tol = 1e-2
    o.max_iterations = 100
    with fe.Simulation(w.params, w.positions, device=0, options=o) as a, \
            fe.Simulation(w.params, w.positions, device=0, options=o) as b:
        b.use_graph(True)
        a.run(6)
        b.run(6)
        compare(a, b, "converge")
