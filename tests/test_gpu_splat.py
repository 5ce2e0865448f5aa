"""The density splat on the GPU, through the C ABI: fe_splat_density's texture equals the
oracle's splat_density.comp restatement bit for bit on the same position_a (the app's scene,
the rest-density scene, the 32k and 1M dam breaks after frames, a moving box, non-cubic and
changing grid_res, particles on the far face, outside the box and NaN); at 1M the texture is
also additive over a split of the particles. The render copies of position_a / colors
(fe_copy_field_device) are checked against downloads.
"""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

F = np.float32


def same(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (name, a.shape, b.shape)
    if not np.array_equal(a, b):
        bad = np.nonzero(a.reshape(-1) != b.reshape(-1))[0]
        raise AssertionError(f"{name}: {len(bad)} of {a.size} texels differ, first at {bad[:5].tolist()}")


def oracle_texture(oracle_mod, sim, positions=None):
    """The oracle's texture for the simulation's current position_a and parameters."""
    p = sim.params
    pos = sim.download("position_a") if positions is None else positions
    orc = oracle_mod.Oracle(p, np.zeros((pos.shape[0], 3), F))
    try:
        orc.set_params(p)
        orc.upload("position_a", pos)
        return orc.splat_density()
    finally:
        orc.close()


CASES = {
    "app_scene_25000": (scene.app_scene, 3),
    "rest_density_4800": (scene.rest_density_scene, 2),
    "dam_break_32k": (lambda: scene.dam_break_32k(seed=6), 5),
}


@pytest.mark.parametrize("case", list(CASES))
def test_splat_matches_oracle(case, oracle_mod):
    make, frames = CASES[case]
    w = make()
    print(f"[gpu splat] case={case} n={w.n}", flush=True)
    with fe.Simulation(w.params, w.positions, device=0) as sim:
        sim.splat_density()
        t0 = sim.density_texture()
        same("frame 0", t0, oracle_texture(oracle_mod, sim))
        assert t0.any()
        sim.run(frames)
        sim.splat_density()
        same(f"after {frames} frames", sim.density_texture(), oracle_texture(oracle_mod, sim))


def test_splat_follows_the_moving_wall(oracle_mod):
    """The texture maps the current box (the app's moving wall) onto grid_res."""
    w = scene.app_scene()
    with fe.Simulation(w.params, w.positions, device=0) as sim:
        sim.set_wall(fe.CollisionBox.from_params(w.params, 0.3, 5.0))
        for k in range(3):
            sim.run(4)
            sim.splat_density()
            same(f"wall {k}", sim.density_texture(), oracle_texture(oracle_mod, sim))


def test_grid_res_changes_and_edges(oracle_mod):
    """Non-cubic and changing grid_res (the texture is reallocated), particles on the far
    face, outside the box and NaN."""
    w = scene.dam_break_32k(seed=8)
    pos = np.concatenate([w.positions, np.array([[1.0, 1.0, 1.0], [0.0, 0.0, 0.0], [-0.01, 0.5, 0.5],
                                                  [0.5, 1.01, 0.5], [np.nan, 0.5, 0.5]], F)])
    with fe.Simulation(w.params, pos, device=0) as sim:
        for res in [(64, 64, 64), (40, 50, 30), (7, 3, 129), (128, 128, 128)]:
            p = fe.FeParams.from_buffer_copy(sim.params)
            for a in range(3):
                p.grid_res[a] = res[a]
            sim.set_params(p)
            sim.splat_density()
            t = sim.density_texture()
            assert t.shape == (res[2], res[1], res[0])
            same(f"grid {res}", t, oracle_texture(oracle_mod, sim))


def test_splat_1m_matches_oracle_and_is_additive(oracle_mod):
    w = scene.dam_break_1m()
    n = w.n
    with fe.Simulation(w.params, w.positions, device=0) as sim:
        sim.run(3)
        sim.splat_density()
        tex = sim.density_texture()
        pos = sim.download("position_a")
        same("1M vs oracle", tex, oracle_texture(oracle_mod, sim, pos))
        with fe.Simulation(w.params, pos[: n // 3, :3], device=0) as a, \
                fe.Simulation(w.params, pos[n // 3:, :3], device=0) as b:
            a.splat_density()
            b.splat_density()
            same("1M additive", tex, a.density_texture() + b.density_texture())


def test_render_copies():
    """fe_copy_field_device into another context's device buffers (as a renderer's imported
    buffers would be), read back through that context."""
    w = scene.rest_density_scene()
    with fe.Simulation(w.params, w.positions, device=0) as sim, \
            fe.Simulation(w.params, np.zeros_like(w.positions), device=0) as dst:
        sim.run(2)
        for name in ("position_a", "colors"):
            ptr = dst.device_ptr(name)
            sim.copy_field_device(name, ptr, w.n * 16)
            sim.sync()
            same(name, dst.download(name).view(np.uint32), sim.download(name).view(np.uint32))
        with pytest.raises(fe.FluidEngineError):
            sim.copy_field_device("grid_entries", dst.device_ptr("position_a"), 8)
        with pytest.raises(fe.FluidEngineError):
            sim.copy_field_device("position_a", dst.device_ptr("position_a"), 16)


def test_texture_misuse():
    w = scene.rest_density_scene()
    with fe.Simulation(w.params, w.positions, device=0) as sim:
        with pytest.raises(fe.FluidEngineError):
            sim.density_texture()
        p = fe.FeParams.from_buffer_copy(sim.params)
        p.grid_res[1] = 0
        sim.set_params(p)
        with pytest.raises(fe.FluidEngineError):
            sim.splat_density()
    ws = scene.batch_scenes(count=2, seed=1, n_min=3000, n_max=4000)
    b = fe.Batch(ws, device=0)
    try:
        with pytest.raises(fe.FluidEngineError):
            b.splat_density()
    finally:
        b.close()
