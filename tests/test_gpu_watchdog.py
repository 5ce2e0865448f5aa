"""The watchdog (fe_set_watchdog / fe_get_health; no reference counterpart, whose only health
signal is the max-speed atomic of stats.comp:22-23), through the C ABI. A healthy frame
reports nothing and leaves every result unchanged; a non-finite value or an over-fast
particle put into the state from the host (fe_upload) is reported for the frame that ran
with it, and with halt set the next frame refuses to run with FE_ERR_UNSTABLE."""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

FIELDS = ("position_a", "velocity_a", "densities", "pressures", "source_terms", "factors")


def same(name, a, b):
    va = np.asarray(a).view(np.uint32)
    vb = np.asarray(b).view(np.uint32)
    assert np.array_equal(va, vb), name


def make():
    w = scene.dam_break_32k(seed=6)
    o = fe.default_options()
    return fe.Simulation(w.params, w.positions, device=0, options=o)


def test_healthy_frames_report_nothing_and_change_nothing():
    with make() as a, make() as b:
        b.set_watchdog(True, speed_limit=50.0, halt=True)
        for f in range(5):
            a.step()
            b.step()
            h = b.health()
            assert h["frame"] == f + 1
            assert h["nonfinite"] == 0 and h["first_nonfinite"] == 0xFFFFFFFF and h["flags"] == 0
            assert h["sort_error"] == 0
            stats = b.read_stats()
            assert np.float32(h["max_speed"]).view(np.uint32) == stats[0]
        for name in FIELDS:
            same(name, a.download(name), b.download(name))
        same("stats", a.read_stats(), b.read_stats())


def test_nonfinite_state_is_reported_and_halts():
    with make() as s:
        s.set_watchdog(True, halt=True)
        s.run(2)
        v = s.download("velocity_a")
        v[123, 1] = np.nan
        s.upload("velocity_a", v)
        s.step()                      # runs: the frame before it was healthy
        h = s.health()
        assert h["flags"] & fe.FE_HEALTH_NONFINITE
        assert h["nonfinite"] >= 1 and h["first_nonfinite"] < s.n
        with pytest.raises(fe.FluidEngineError) as err:
            s.step()
        assert err.value.status == fe.FE_ERR_UNSTABLE
        assert "non-finite" in str(err.value)


def test_speed_limit_is_reported_and_halts():
    with make() as s:
        s.set_watchdog(True, speed_limit=2.0, halt=True)
        s.run(2)
        assert s.health()["flags"] == 0
        v = s.download("velocity_a")
        v[5, :3] = (40.0, 0.0, 0.0)
        s.upload("velocity_a", v)
        s.step()
        h = s.health()
        assert h["flags"] == fe.FE_HEALTH_SPEED and h["max_speed"] > 2.0 and h["nonfinite"] == 0
        with pytest.raises(fe.FluidEngineError) as err:
            s.step()
        assert err.value.status == fe.FE_ERR_UNSTABLE


def test_without_halt_frames_continue():
    with make() as s:
        s.set_watchdog(True, speed_limit=2.0, halt=False)
        s.run(1)
        v = s.download("velocity_a")
        v[7, :3] = (40.0, 0.0, 0.0)
        s.upload("velocity_a", v)
        s.run(3)
        assert s.health()["frame"] == 4
        s.set_watchdog(False)
        with pytest.raises(fe.FluidEngineError):
            s.health()


def test_batch_refuses_the_watchdog():
    ws = scene.batch_scenes(count=3, seed=1, n_min=4096, n_max=8192)
    with fe.Batch(ws, device=0, options=fe.default_options()) as b:
        with pytest.raises(fe.FluidEngineError):
            b.set_watchdog(True)
