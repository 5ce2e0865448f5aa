"""The moving wall on the GPU, through the C ABI: frames with an attached CollisionBox
(src/entities/collision.rs:31-35, the box entering the parameters as in
src/renderer/ui.rs:125-126 and the wall advancing as in src/renderer/mod.rs:269-270) equal
the oracle's frames bit for bit, with and without the host flags; the wall really moves the
fluid; a wall of amplitude 0 is the fixed box; and batch scenes with their own walls (eager
and replayed from a captured graph) equal single contexts with the same walls.
"""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

F = np.float32
STEP_FIELDS = ("position_a", "velocity_a", "velocity_b", "densities", "factors", "source_terms", "pressures",
               "pressure_accelerations", "colors", "grid_entries", "grid_start")


def opts(mode=fe.FE_SOLVER_FIXED, **kw):
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_RADIX
    o.solver_mode = mode
    o.density_tol = 1e-3
    o.divergence_tol = 1e-2
    o.max_iterations = 100
    for k, v in kw.items():
        setattr(o, k, v)
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


# (workload, wall base min.x just left of the fluid, amplitude m, frequency Hz)
CASES = {
    "rest_density_4800": (scene.rest_density_scene, -0.79, 0.05, 5.0),
    "dam_break_32k": (lambda: scene.dam_break_32k(seed=4), -0.002, 0.03, 4.0),
}


def walls_for(w, base, amp, freq, oracle_mod):
    box_min = (base, float(w.params.box_min[1]), float(w.params.box_min[2]))
    box_max = tuple(float(x) for x in w.params.box_max[:3])
    b = fe.CollisionBox(box_min, box_max)
    b.wave_amplitude = amp
    b.wave_frequency = freq
    return b, oracle_mod.make_wall(box_min, box_max, amp, freq)


@pytest.mark.parametrize("case", list(CASES))
def test_wall_frames_match_oracle(case, oracle_mod):
    make, base, amp, freq = CASES[case]
    w = make()
    print(f"[gpu wall] case={case} n={w.n}", flush=True)
    o = opts()
    box, wall = walls_for(w, base, amp, freq, oracle_mod)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as sim:
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            sim.set_wall(box)
            orc.set_wall(wall)
            for frame in range(6):
                sim.step()
                orc.step()
                same(f"frame {frame} box_min", np.array(sim.params.box_min[:]), np.array(orc.params.box_min[:]))
                same(f"frame {frame} box_max", np.array(sim.params.box_max[:]), np.array(orc.params.box_max[:]))
                for name in STEP_FIELDS:
                    same(f"frame {frame} {name}", sim.download(name), orc.download(name))
                same(f"frame {frame} stats", sim.read_stats(), orc.read_stats())
                g = sim.wall()
                ow = orc.wall()
                assert F(g.time) == F(ow.time) and F(g.min[0]) == F(ow.min[0]), frame
        finally:
            orc.close()


def test_wall_with_host_flags(oracle_mod):
    """CFL, early exit and the converge mode act on frames whose box moves."""
    w = scene.rest_density_scene()
    o = opts(mode=fe.FE_SOLVER_CONVERGE, use_cfl=1, use_solver_error_threshold=1)
    box, wall = walls_for(w, -0.79, 0.05, 5.0, oracle_mod)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as sim:
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            sim.set_wall(box)
            orc.set_wall(wall)
            for frame in range(5):
                sim.step(w.params.dt)
                orc.step(w.params.dt)
                assert sim.last_iterations() == orc.last_iterations(), frame
                assert sim.params.dt == orc.params.dt, frame
                for name in ("position_a", "velocity_a", "pressures"):
                    same(f"flags frame {frame} {name}", sim.download(name), orc.download(name))
        finally:
            orc.close()


def test_wall_moves_the_fluid(oracle_mod):
    """The same frames without the wall differ: the wall pushed particles."""
    w = scene.rest_density_scene()
    o = opts()
    box, _ = walls_for(w, -0.79, 0.05, 5.0, oracle_mod)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as a, \
            fe.Simulation(w.params, w.positions, device=0, options=o) as b:
        a.set_wall(box)
        a.run(6)
        b.run(6)
        pa = a.download("position_a")
        pb = b.download("position_a")
        moved = np.abs(pa[:, 0] - pb[:, 0]) > 1e-4
        assert moved.sum() > 100, moved.sum()
        # the wall advanced by 6 max_dt, from the start of the wave
        g = a.wall()
        assert abs(g.time - 6 * float(w.params.dt)) < 1e-6
        assert g.min[0] > -0.79 + 0.02


def test_static_wall_is_the_fixed_box():
    """A wall of amplitude 0 with the parameter block's own box changes nothing."""
    w = scene.dam_break_32k(seed=5)
    o = opts(mode=fe.FE_SOLVER_CONVERGE)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as a, \
            fe.Simulation(w.params, w.positions, device=0, options=o) as b:
        a.set_wall(fe.CollisionBox.from_params(w.params))
        a.run(4)
        b.run(4)
        for name in STEP_FIELDS:
            same(name, a.download(name), b.download(name))
        assert a.wall().time > 0.0
        a.set_wall(None)
        assert a.wall() is None


@pytest.mark.parametrize("graph", [False, True])
def test_batch_walls_match_single_contexts(graph, oracle_mod):
    ws = scene.batch_scenes(count=4, seed=9, n_min=3000, n_max=12000)
    mode = fe.FE_SOLVER_FIXED if graph else fe.FE_SOLVER_CONVERGE
    o = opts(mode=mode)
    waves = {0: (0.02, 6.0), 2: (0.04, 3.0)}
    frames = 6
    b = fe.Batch(ws, device=0, options=o)
    try:
        if graph:
            b.use_graph(True)
        for k, (amp, freq) in waves.items():
            box, _ = walls_for(ws[k], -0.002, amp, freq, oracle_mod)
            b.set_scene_wall(k, box)
        b.run(frames)
        for k, wk in enumerate(ws):
            with fe.Simulation(wk.params, wk.positions, device=0, options=o) as s:
                if k in waves:
                    box, _ = walls_for(wk, -0.002, *waves[k], oracle_mod)
                    s.set_wall(box)
                s.run(frames)
                for name in ("position_a", "velocity_a", "densities", "pressures", "colors", "grid_start"):
                    same(f"scene {k} {name}", b.download_scene(k, name), s.download(name))
                same(f"scene {k} stats", b.read_scene_stats(k), s.read_stats())
                bw = b.scene_wall(k)
                sw = s.wall()
                if k in waves:
                    assert F(bw.min[0]) == F(sw.min[0]) and F(bw.time) == F(sw.time), k
                else:
                    assert bw is None and sw is None
    finally:
        b.close()


def test_wall_entry_points_reject_misuse():
    ws = scene.batch_scenes(count=2, seed=1, n_min=3000, n_max=4000)
    b = fe.Batch(ws, device=0)
    try:
        box = fe.CollisionBox((0, 0, 0), (1, 1, 1))
        with pytest.raises(fe.FluidEngineError):
            b.set_wall(box)
        with pytest.raises(fe.FluidEngineError):
            b.set_scene_wall(5, box)
        bad = fe.CollisionBox((0, 0, 0), (1, 1, 1))
        bad.wave_amplitude = float("nan")
        with pytest.raises(fe.FluidEngineError):
            b.set_scene_wall(0, bad)
        assert b.scene_wall(0) is None
    finally:
        b.close()
