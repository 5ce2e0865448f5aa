"""The autotuner on the GPU: every tunable kernel gets a configuration it is built for and a
measured median; tuning on a stepped state leaves the following frames bit-identical to an
untuned context (the timed launches' effects are undone, and configurations never change
results); an explicit non-default configuration gives the same fields; the JSON cache
round-trips."""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

FIELDS = ("position_a", "velocity_a", "densities", "pressures", "factors", "colors")


def conv():
    o = fe.default_options()
    o.solver_mode = fe.FE_SOLVER_CONVERGE
    return o


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_autotune_keeps_results(tmp_path):
    w = scene.dam_break_32k(seed=4)
    ref = fe.Simulation(w.params, w.positions, device=0, options=conv())
    tun = fe.Simulation(w.params, w.positions, device=0, options=conv())
    try:
        ref.run(2)
        tun.run(2)
        res = tun.autotune(str(tmp_path / "tune.json"), reps=5)
        assert set(res) == {"density_alpha", "viscosity", "density_source", "divergence_source", "pressure_force",
                            "pressure_update", "build_list", "pressure_integration", "divergence_integration",
                            "density_alpha_divergence_source"}
        for k, v in res.items():
            assert v["median_us"] > 0 and v["default_us"] > 0, k
            assert v["median_us"] <= v["default_us"] * 1.0001, k
        print({k: (v["block"], v["items"], v["median_us"], v["default_us"]) for k, v in res.items()})
        ref.run(5)
        tun.run(5)
        for name in FIELDS:
            assert same(tun.download(name), ref.download(name)), name
        assert tun.iteration_totals() == ref.iteration_totals()
        # the cache applies the same configuration to a new context without measuring
        fresh = fe.Simulation(w.params, w.positions, device=0, options=conv())
        got = fresh.autotune(str(tmp_path / "tune.json"))
        assert {k: (v["block"], v["items"]) for k, v in got.items()} == \
            {k: (v["block"], v["items"]) for k, v in res.items()}
        assert {k: (v["block"], v["items"]) for k, v in fresh.tune_entries().items()} == \
            {k: (v["block"], v["items"]) for k, v in res.items()}
        fresh.close()
    finally:
        ref.close()
        tun.close()


def test_explicit_configurations_same_fields():
    w = scene.dam_break_32k(seed=5)
    a = fe.Simulation(w.params, w.positions, device=0, options=conv())
    b = fe.Simulation(w.params, w.positions, device=0, options=conv())
    try:
        for k, blk, it in (("pressure_force", 64, 2), ("pressure_update", 1024, 1), ("density_alpha", 512, 1),
                           ("viscosity", 128, 1), ("build_list", 64, 1), ("pressure_integration", 1024, 4),
                           ("divergence_integration", 64, 2), ("density_source", 1024, 1),
                           ("divergence_source", 64, 1)):
            b.set_tune(k, blk, it)
        a.run(4)
        b.run(4)
        for name in FIELDS:
            assert same(a.download(name), b.download(name)), name
        assert a.iteration_totals() == b.iteration_totals()
        with pytest.raises(fe.FluidEngineError):
            b.set_tune("pressure_force", 96, 1)
        with pytest.raises(fe.FluidEngineError):
            b.set_tune("build_list", 512, 1)
        with pytest.raises(fe.FluidEngineError):   # staged sweeps run 1, 2 or 4 windows per block
            b.set_tune("pressure_update", 512, 2)
        with pytest.raises(fe.FluidEngineError):
            b.set_tune("no_such_kernel", 256, 1)
    finally:
        a.close()
        b.close()


def test_autotune_batch():
    ws = scene.batch_scenes(count=4, seed=2, n_min=4096, n_max=16384)
    o = fe.default_options()
    ref = fe.Batch(ws, device=0, options=o)
    tun = fe.Batch(ws, device=0, options=o)
    try:
        ref.run(2)
        tun.run(2)
        tun.autotune_now(reps=3)
        tun.use_graph(True)
        ref.run(4)
        tun.run(4)
        for k in range(len(ws)):
            assert same(tun.download_scene(k, "position_a"), ref.download_scene(k, "position_a")), k
    finally:
        ref.close()
        tun.close()


@pytest.mark.parametrize("block", [64, 128, 512, 1024])
def test_walled_gathers_take_tuned_blocks(block):
    """With wall boundary particles the gathers run the walled instances in the block size
    tuned for each kernel: every block size gives the default configuration's frames, and
    tuning the walled context keeps its next frames."""
    w = scene.dam_break_32k(seed=8)
    o = conv()
    o.boundary = fe.FE_BOUNDARY_PARTICLES
    names = ("density_alpha", "density_source", "divergence_source", "pressure_force", "pressure_update",
             "density_alpha_divergence_source")
    with fe.Simulation(w.params, w.positions, device=0, options=o) as ref, \
            fe.Simulation(w.params, w.positions, device=0, options=o) as alt:
        for k in names:
            alt.set_tune(k, block, 1)
        ref.run(3)
        alt.run(3)
        for name in FIELDS:
            assert same(alt.download(name), ref.download(name)), name
        assert alt.iteration_totals() == ref.iteration_totals()
        if block == 64:
            alt.autotune(None, reps=3)
            ref.run(2)
            alt.run(2)
            for name in FIELDS:
                assert same(alt.download(name), ref.download(name)), name
