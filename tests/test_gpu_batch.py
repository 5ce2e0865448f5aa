"""Batched scenes on the GPU, through the C ABI: every scene of a batch evolves exactly as the
same scene does in its own context (bit for bit: fields, grid tables, stats, sweep counts),
the frame replayed from a captured HIP graph equals the eager frame, and a batch scene also
matches the CPU oracle directly. Single-scene contexts are themselves pinned to the oracle
(test_gpu_parity.py), so the batch inherits that parity.
"""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

FIELDS = ("position_a", "velocity_a", "velocity_b", "position_b", "densities", "factors", "source_terms",
          "pressures", "pressure_accelerations", "colors", "grid_entries", "grid_start")


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


@pytest.fixture(scope="module")
def small_batch():
    return scene.batch_scenes(count=5, seed=3, n_min=4096, n_max=20000)


@pytest.mark.parametrize("mode", [fe.FE_SOLVER_FIXED, fe.FE_SOLVER_CONVERGE])
def test_batch_scenes_match_single_contexts(small_batch, mode):
    frames = 4
    b = fe.Batch(small_batch, device=0, options=opts(mode))
    try:
        b.run(frames)
        b.sync()
        for k, w in enumerate(small_batch):
            with fe.Simulation(w.params, w.positions, device=0, options=opts(mode)) as s:
                s.run(frames)
                for name in FIELDS:
                    same(f"scene {k} {name}", b.download_scene(k, name), s.download(name))
                same(f"scene {k} stats", b.read_scene_stats(k), s.read_stats())
                (d, v), (dt, vt) = b.scene_iterations(k)
                sd, sv = s.last_iterations()
                std, stv, _ = s.iteration_totals()
                assert (d, v) == (sd, sv), k
                assert (dt, vt) == (std, stv), k
        if mode == fe.FE_SOLVER_CONVERGE:
            its = [b.scene_iterations(k)[1][0] for k in range(len(small_batch))]
            print("density sweeps per scene over", frames, "frames:", its)
    finally:
        b.close()


def test_batch_scene_matches_oracle(small_batch, oracle_mod):
    o = opts(fe.FE_SOLVER_FIXED)
    b = fe.Batch(small_batch, device=0, options=o)
    try:
        b.run(2)
        for k in (0, len(small_batch) - 1):
            w = small_batch[k]
            orc = oracle_mod.Oracle(w.params, w.positions, options=o)
            orc.step()
            orc.step()
            for name in ("position_a", "velocity_a", "densities", "pressures", "grid_entries", "grid_start"):
                same(f"scene {k} {name} vs oracle", b.download_scene(k, name), orc.download(name))
            orc.close()
    finally:
        b.close()


def test_graph_replay_matches_eager(small_batch):
    o = opts(fe.FE_SOLVER_FIXED)
    eager = fe.Batch(small_batch, device=0, options=o)
    graph = fe.Batch(small_batch, device=0, options=o)
    try:
        graph.use_graph(True)
        eager.run(6)
        graph.run(6)
        for k in range(len(small_batch)):
            for name in ("position_a", "velocity_a", "densities", "pressures", "colors"):
                same(f"scene {k} {name}", graph.download_scene(k, name), eager.download_scene(k, name))
            assert graph.scene_iterations(k) == eager.scene_iterations(k)
        assert graph.iteration_totals() == eager.iteration_totals()
        # a parameter change drops the graph; the next frames still agree
        p = fe.FeParams.from_buffer_copy(small_batch[1].params)
        p.viscosity = 0.02
        eager.set_scene_params(1, p)
        graph.set_scene_params(1, p)
        eager.run(3)
        graph.run(3)
        for k in range(len(small_batch)):
            same(f"scene {k} after change", graph.download_scene(k, "velocity_a"),
                 eager.download_scene(k, "velocity_a"))
    finally:
        eager.close()
        graph.close()


def test_batch_rejects_unsupported():
    ws = scene.batch_scenes(count=3, seed=1, n_min=4096, n_max=8192)
    b = fe.Batch(ws, device=0)
    try:
        o = opts(fe.FE_SOLVER_FIXED)
        o.sort_algorithm = fe.FE_SORT_BITONIC
        with pytest.raises(fe.FluidEngineError):
            b.set_options(o)
        with pytest.raises(fe.FluidEngineError):
            b.sort()
        with pytest.raises(fe.FluidEngineError):
            b.download("grid_entries")
        with pytest.raises(fe.FluidEngineError):
            b.learn_collect(16, 1024)
    finally:
        b.close()


def test_batch_64_scenes():
    """The 64-scene workload (particle counts log-uniform in [4096, 262144]): converging
    frames run, every scene converges on its own, and the smallest and the largest scene
    equal their own contexts after two frames."""
    ws = scene.batch_scenes(count=64, seed=0)
    b = fe.Batch(ws, device=0, options=opts(fe.FE_SOLVER_CONVERGE))
    try:
        b.run(2)
        b.sync()
        n = np.array(b.scene_n)
        print(f"64 scenes: {n.sum()} particles, {n.min()}..{n.max()} per scene")
        for k in (int(n.argmin()), int(n.argmax())):
            w = ws[k]
            with fe.Simulation(w.params, w.positions, device=0, options=opts(fe.FE_SOLVER_CONVERGE)) as s:
                s.run(2)
                for name in ("position_a", "velocity_a", "pressures"):
                    same(f"scene {k} {name}", b.download_scene(k, name), s.download(name))
                assert b.scene_iterations(k)[1] == s.iteration_totals()[:2]
        assert np.isfinite(b.download("position_a")).all()
    finally:
        b.close()
