"""Wall boundary particles on the GPU, through the C ABI: frames of the 32k dam break (a block in
the box corner, on the floor and against two walls) with FE_BOUNDARY_PARTICLES equal the
oracle's frames bit for bit, in fixed and converge modes, combined with the consistent
divergence order, and under the moving wall (whose box changes the samples every frame); the
option is refused where it does not apply.
"""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

FIELDS = ("position_a", "velocity_a", "velocity_b", "position_b", "densities", "factors", "source_terms",
          "pressures", "pressure_accelerations", "colors", "grid_entries", "grid_start")


def opts(mode=fe.FE_SOLVER_FIXED, order=fe.FE_DIVERGENCE_REFERENCE):
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_RADIX
    o.solver_mode = mode
    o.density_tol = 1e-3
    o.divergence_tol = 1e-2
    o.max_iterations = 100
    o.divergence_order = order
    o.boundary = fe.FE_BOUNDARY_PARTICLES
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


@pytest.mark.parametrize("mode,order", [(fe.FE_SOLVER_FIXED, fe.FE_DIVERGENCE_REFERENCE),
                                        (fe.FE_SOLVER_CONVERGE, fe.FE_DIVERGENCE_REFERENCE),
                                        (fe.FE_SOLVER_CONVERGE, fe.FE_DIVERGENCE_CONSISTENT)])
def test_boundary_frames_match_oracle(mode, order, oracle_mod):
    w = scene.dam_break_32k(seed=2)
    o = opts(mode, order)
    print(f"[gpu boundary] mode={mode} order={order} n={w.n}", flush=True)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as sim:
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            for frame in range(5):
                sim.step()
                orc.step()
                for name in FIELDS:
                    same(f"frame {frame} {name}", sim.download(name), orc.download(name))
                same(f"frame {frame} stats", sim.read_stats(), orc.read_stats())
                assert sim.last_iterations() == orc.last_iterations(), frame
        finally:
            orc.close()


def test_boundary_changes_the_wall_layer():
    """The samples act: the first frame's densities differ from the clamp-only run near the
    walls and agree far from them."""
    w = scene.dam_break_32k(seed=2)
    out = []
    for b in (fe.FE_BOUNDARY_CLAMP, fe.FE_BOUNDARY_PARTICLES):
        o = opts()
        o.boundary = b
        with fe.Simulation(w.params, w.positions, device=0, options=o) as s:
            s.neighbor_search()
            s.density_alpha()
            out.append((s.download("position_b"), s.download("densities")))
    (pos, rc), (_, rw) = out
    # The sample planes lie r outside the walls: a particle reaches them within h - r of a wall.
    # Every particle of the wall layers (within r + r / 2 of a wall) gains density; beyond the
    # reach nothing changes. (Between, the one sample in reach adds ~1e-5, which rounds away
    # against rho ~ 1000 for some particles.)
    r = np.float32(w.params.particle_radius)
    reach = np.float32(w.params.smoothing_radius) - r
    layer = (pos[:, 0] < np.float32(1.5) * r) | (pos[:, 1] < np.float32(1.5) * r) | (pos[:, 2] < np.float32(1.5) * r)
    far = (pos[:, 0] >= reach) & (pos[:, 1] >= reach) & (pos[:, 2] >= reach)
    assert layer.sum() > 1000 and far.sum() > 1000
    assert (rw[layer] > rc[layer]).all() and (rw >= rc).all()
    same("far from the walls", rw[far], rc[far])


def test_boundary_with_moving_wall(oracle_mod):
    w = scene.dam_break_32k(seed=4)
    o = opts(fe.FE_SOLVER_CONVERGE)
    box = fe.CollisionBox((-0.002, 0.0, 0.0), tuple(float(x) for x in w.params.box_max[:3]))
    box.wave_amplitude = 0.03
    box.wave_frequency = 4.0
    wall = oracle_mod.make_wall((-0.002, 0.0, 0.0), tuple(float(x) for x in w.params.box_max[:3]), 0.03, 4.0)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as sim:
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            sim.set_wall(box)
            orc.set_wall(wall)
            for frame in range(5):
                sim.step()
                orc.step()
                for name in ("position_a", "velocity_a", "densities", "pressures"):
                    same(f"wall frame {frame} {name}", sim.download(name), orc.download(name))
        finally:
            orc.close()


@pytest.mark.parametrize("change", ["mass", "boundary_on", "boundary_off"])
def test_frame_start_reuse_follows_changes(change, oracle_mod):
    """fe_step reuses the last search and densities at a frame's start when the particles have
    not moved; a new particle mass or a boundary switch between frames changes those densities,
    and the frames must still equal the oracle's (which always recomputes them)."""
    w = scene.dam_break_32k(seed=6)
    o = opts()
    o.boundary = fe.FE_BOUNDARY_PARTICLES if change == "boundary_off" else fe.FE_BOUNDARY_CLAMP
    with fe.Simulation(w.params, w.positions, device=0, options=o) as sim:
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            for frame in range(4):
                if frame == 2:
                    if change == "mass":
                        p = fe.FeParams.from_buffer_copy(sim.params)
                        p.particle_mass = p.particle_mass * 1.05
                        sim.set_params(p)
                        orc.set_params(p)
                    else:
                        o.boundary = fe.FE_BOUNDARY_PARTICLES if change == "boundary_on" else fe.FE_BOUNDARY_CLAMP
                        sim.set_options(o)
                        orc.set_options(o)
                sim.step()
                orc.step()
                for name in ("position_a", "velocity_a", "densities", "pressures"):
                    same(f"{change} frame {frame} {name}", sim.download(name), orc.download(name))
        finally:
            orc.close()


def test_boundary_option_refused_where_it_does_not_apply():
    ws = scene.batch_scenes(count=2, seed=1, n_min=3000, n_max=4000)
    b = fe.Batch(ws, device=0)
    try:
        with pytest.raises(fe.FluidEngineError):
            b.set_options(opts())
    finally:
        b.close()
    w = scene.rest_density_scene()
    p = fe.FeParams.from_buffer_copy(w.params)
    p.smoothing_radius = 40 * p.particle_radius
    with fe.Simulation(p, w.positions, device=0) as s:
        with pytest.raises(fe.FluidEngineError):
            s.set_options(opts())
