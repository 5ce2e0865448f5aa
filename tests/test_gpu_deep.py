"""The small-scene launch variants on the GPU, through the C ABI:
- deep list gathers (tune items 4): 16 listed neighbours' records in flight per work-item
  instead of 4 (one scene, blocks of 64 to 256);
- four lanes per particle (items 8) in the two Jacobi sweeps and the other list gathers: each
  lane forms the terms of every fourth listed slot and the quad adds them in slot order;
- build_list over windows of 8 schedule positions per wave (items 8) instead of 64.
Each visits every particle's pairs in the same list order and builds the same lists, so
fields, grid tables, stats and sweep counts equal those of the default launches bit for bit,
and the oracle's directly. Covered: fixed and converging frames at every block size, the
reference app's scene with its moving wall replayed from a graph, particles that overflow
the list (bucket walk), and a batch (which keeps the default gathers and sweeps)."""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

FIELDS = ("position_a", "velocity_a", "velocity_b", "position_b", "densities", "factors", "source_terms",
          "pressures", "pressure_accelerations", "colors")
GATHERS = ("density_alpha", "viscosity", "density_source", "divergence_source", "pressure_force", "pressure_update",
           "density_alpha_divergence_source")


def opts(mode):
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_RADIX
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


def deep(sim, block):
    for k in GATHERS:
        sim.set_tune(k, block, 4)


def small(sim, block, gathers=4):
    """Every small-scene variant: deep (gathers=4) or four-lane (gathers=8) list gathers, quad
    sweeps, 8-position list windows."""
    for k in GATHERS:
        sim.set_tune(k, block, gathers)
    sim.set_tune("pressure_force", block, 8)
    sim.set_tune("pressure_update", block, 8)
    sim.set_tune("build_list", block, 8)


def compare(a, b, label):
    for name in FIELDS + ("grid_entries", "grid_start"):
        same(f"{label} {name}", a.download(name), b.download(name))
    same(f"{label} stats", a.read_stats(), b.read_stats())
    assert a.iteration_totals() == b.iteration_totals(), label


@pytest.mark.parametrize("mode", [fe.FE_SOLVER_FIXED, fe.FE_SOLVER_CONVERGE])
@pytest.mark.parametrize("block", [64, 128, 256])
def test_deep_matches_default_32k(mode, block):
    w = scene.dam_break_32k(seed=9)
    with fe.Simulation(w.params, w.positions, device=0, options=opts(mode)) as a, \
            fe.Simulation(w.params, w.positions, device=0, options=opts(mode)) as b:
        deep(b, block)
        assert all(e["items"] == 4 and e["block"] == block for k, e in b.tune_entries().items() if k in GATHERS)
        a.run(5)
        b.run(5)
        compare(a, b, f"mode {mode} block {block}")


def test_deep_matches_oracle(oracle_mod):
    w = scene.rest_density_scene()
    o = opts(fe.FE_SOLVER_FIXED)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as s:
        deep(s, 128)
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            for f in range(3):
                s.step()
                orc.step()
                for name in FIELDS:
                    same(f"frame {f} {name} vs oracle", s.download(name), orc.download(name))
                same(f"frame {f} stats vs oracle", s.read_stats(), orc.read_stats())
        finally:
            orc.close()


def test_deep_app_scene_graph_frames():
    w = scene.app_scene()
    o = opts(fe.FE_SOLVER_FIXED)
    sims = []
    try:
        for _ in range(2):
            s = fe.Simulation(w.params, w.positions, device=0, options=o)
            s.set_wall(fe.CollisionBox.from_params(w.params, *w.wall_wave))
            sims.append(s)
        a, b = sims
        deep(b, 64)
        b.use_graph(True)
        for f in range(6):
            a.step()
            b.step()
            compare(a, b, f"app frame {f}")
        a.run(30)
        b.run(30)
        compare(a, b, "app after 36 frames")
    finally:
        for s in sims:
            s.close()


def test_deep_overflowing_lists_take_the_walk(oracle_mod):
    """A clump of 300 particles in half a smoothing length (about 299 pairs each, past the
    list's 64) beside ordinary particles: the clump walks the buckets, the rest read their
    lists 16 at a time, stage by stage against the oracle."""
    rng = np.random.default_rng(5)
    w = scene.dam_break_32k(seed=4)
    pos = w.positions.copy()
    pos[:300] = np.array([0.5, 0.3, 0.5], np.float32) + rng.uniform(0, 0.02, (300, 3)).astype(np.float32)
    o = opts(fe.FE_SOLVER_FIXED)
    with fe.Simulation(w.params, pos, device=0, options=o) as s:
        deep(s, 256)
        orc = oracle_mod.Oracle(w.params, pos, options=o)
        try:
            for x in (s, orc):
                x.neighbor_search()
                x.density_alpha()
                x.viscosity()
                x.density_source()
                x.pressure_force()
                x.pressure_update()
            for name in ("densities", "factors", "velocity_a", "source_terms", "pressure_accelerations", "pressures"):
                same("clump " + name, s.download(name), orc.download(name))
        finally:
            orc.close()


def test_batch_keeps_four_deep_gathers():
    ws = scene.batch_scenes(count=4, seed=3, n_min=3000, n_max=12000)
    o = opts(fe.FE_SOLVER_CONVERGE)
    a = fe.Batch(ws, device=0, options=o)
    b = fe.Batch(ws, device=0, options=o)
    try:
        deep(b, 128)
        a.run(3)
        b.run(3)
        for name in FIELDS:
            same(name, a.download(name), b.download(name))
        for k in range(len(ws)):
            assert a.scene_iterations(k) == b.scene_iterations(k), k
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("mode", [fe.FE_SOLVER_FIXED, fe.FE_SOLVER_CONVERGE])
@pytest.mark.parametrize("block", [64, 128, 256])
def test_quad_sweeps_and_small_windows_match_default_32k(mode, block):
    w = scene.dam_break_32k(seed=10)
    with fe.Simulation(w.params, w.positions, device=0, options=opts(mode)) as a, \
            fe.Simulation(w.params, w.positions, device=0, options=opts(mode)) as b:
        small(b, block)
        e = b.tune_entries()
        assert e["pressure_update"]["items"] == 8 and e["build_list"]["items"] == 8
        a.run(5)
        b.run(5)
        compare(a, b, f"mode {mode} block {block}")


def test_small_variants_match_oracle(oracle_mod):
    w = scene.rest_density_scene()
    o = opts(fe.FE_SOLVER_CONVERGE)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as s:
        small(s, 64)
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            for f in range(3):
                s.step()
                orc.step()
                assert s.last_iterations() == orc.last_iterations(), f
                for name in FIELDS:
                    same(f"frame {f} {name} vs oracle", s.download(name), orc.download(name))
                same(f"frame {f} stats vs oracle", s.read_stats(), orc.read_stats())
        finally:
            orc.close()


def test_small_variants_app_scene_graph_frames():
    w = scene.app_scene()
    o = opts(fe.FE_SOLVER_FIXED)
    sims = []
    try:
        for _ in range(2):
            x = fe.Simulation(w.params, w.positions, device=0, options=o)
            x.set_wall(fe.CollisionBox.from_params(w.params, *w.wall_wave))
            sims.append(x)
        a, b = sims
        small(b, 64)
        b.use_graph(True)
        for f in range(6):
            a.step()
            b.step()
            compare(a, b, f"app frame {f}")
        a.run(30)
        b.run(30)
        compare(a, b, "app after 36 frames")
    finally:
        for x in sims:
            x.close()


def test_quad_sweeps_overflowing_lists(oracle_mod):
    """The clump beside ordinary particles with quad sweeps and 8-position list windows: the
    clump's lane 0 walks the buckets alone, the rest read their lists four lanes at a time."""
    rng = np.random.default_rng(6)
    w = scene.dam_break_32k(seed=4)
    pos = w.positions.copy()
    pos[:300] = np.array([0.5, 0.3, 0.5], np.float32) + rng.uniform(0, 0.02, (300, 3)).astype(np.float32)
    o = opts(fe.FE_SOLVER_FIXED)
    with fe.Simulation(w.params, pos, device=0, options=o) as s:
        small(s, 128)
        orc = oracle_mod.Oracle(w.params, pos, options=o)
        try:
            for x in (s, orc):
                x.neighbor_search()
                x.density_alpha()
                x.viscosity()
                x.density_source()
                for _ in range(2):
                    x.pressure_force()
                    x.pressure_update()
            for name in ("densities", "factors", "velocity_a", "source_terms", "pressure_accelerations", "pressures"):
                same("clump " + name, s.download(name), orc.download(name))
        finally:
            orc.close()


def test_batch_small_windows_and_default_sweeps():
    ws = scene.batch_scenes(count=5, seed=8, n_min=3000, n_max=15000)
    o = opts(fe.FE_SOLVER_CONVERGE)
    a = fe.Batch(ws, device=0, options=o)
    b = fe.Batch(ws, device=0, options=o)
    try:
        small(b, 64)
        a.run(3)
        b.run(3)
        for name in FIELDS:
            same(name, a.download(name), b.download(name))
        for k in range(len(ws)):
            assert a.scene_iterations(k) == b.scene_iterations(k), k
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("mode", [fe.FE_SOLVER_FIXED, fe.FE_SOLVER_CONVERGE])
@pytest.mark.parametrize("block", [64, 256])
def test_four_lane_gathers_match_default_32k(mode, block):
    w = scene.dam_break_32k(seed=12)
    with fe.Simulation(w.params, w.positions, device=0, options=opts(mode)) as a, \
            fe.Simulation(w.params, w.positions, device=0, options=opts(mode)) as b:
        small(b, block, gathers=8)
        assert all(e["items"] == 8 for k, e in b.tune_entries().items() if k in GATHERS)
        a.run(5)
        b.run(5)
        compare(a, b, f"mode {mode} block {block}")


def test_four_lane_gathers_match_oracle_with_overflow(oracle_mod):
    """Every list gather four lanes per particle, stage by stage against the oracle, with the
    clump whose lists overflow (walked by lane 0 alone), then two whole frames."""
    rng = np.random.default_rng(13)
    w = scene.dam_break_32k(seed=4)
    pos = w.positions.copy()
    pos[:300] = np.array([0.5, 0.3, 0.5], np.float32) + rng.uniform(0, 0.02, (300, 3)).astype(np.float32)
    o = opts(fe.FE_SOLVER_FIXED)
    with fe.Simulation(w.params, pos, device=0, options=o) as s:
        small(s, 128, gathers=8)
        orc = oracle_mod.Oracle(w.params, pos, options=o)
        try:
            for x in (s, orc):
                x.neighbor_search()
                x.density_alpha()
                x.viscosity()
                x.density_source()
                for _ in range(2):
                    x.pressure_force()
                    x.pressure_update()
                x.divergence_source()
            for name in ("densities", "factors", "velocity_a", "source_terms", "pressure_accelerations", "pressures"):
                same("clump " + name, s.download(name), orc.download(name))
            for f in range(2):
                s.step()
                orc.step()
                for name in FIELDS:
                    same(f"frame {f} {name} vs oracle", s.download(name), orc.download(name))
        finally:
            orc.close()


def test_four_lane_gathers_app_scene_graph_frames():
    w = scene.app_scene()
    o = opts(fe.FE_SOLVER_FIXED)
    sims = []
    try:
        for _ in range(2):
            x = fe.Simulation(w.params, w.positions, device=0, options=o)
            x.set_wall(fe.CollisionBox.from_params(w.params, *w.wall_wave))
            sims.append(x)
        a, b = sims
        small(b, 64, gathers=8)
        b.use_graph(True)
        for f in range(4):
            a.step()
            b.step()
            compare(a, b, f"app frame {f}")
        a.run(20)
        b.run(20)
        compare(a, b, "app after 24 frames")
    finally:
        for x in sims:
            x.close()


def test_four_lane_gathers_with_a_stale_list(oracle_mod):
    """position_b uploaded after the list was built: the list is stale, so every four-lane
    gather walks the buckets in lane 0, with the moved positions, as the oracle does."""
    rng = np.random.default_rng(14)
    w = scene.dam_break_32k(seed=5)
    o = opts(fe.FE_SOLVER_FIXED)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as s:
        small(s, 64, gathers=8)
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            for x in (s, orc):
                x.neighbor_search()
                x.density_alpha()
            pb = orc.download("position_b")
            pb[:, :3] += rng.normal(0, 0.002, (pb.shape[0], 3)).astype(np.float32)
            for x in (s, orc):
                x.upload("position_b", pb)
                x.viscosity()
                x.density_source()
                x.pressure_force()
                x.pressure_update()
            for name in ("velocity_a", "source_terms", "pressure_accelerations", "pressures"):
                same("stale-list " + name, s.download(name), orc.download(name))
        finally:
            orc.close()
