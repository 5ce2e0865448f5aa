"""The moving wall (CollisionBox, src/entities/collision.rs:11-35) on the CPU: the oracle's
restatement pinned by closed forms and by a float64 evaluation of the wave, the frame order
of the box (src/renderer/ui.rs:125-126 before the physics, src/renderer/mod.rs:270 after the
parameter upload) on an oracle run, and the library's host functions fe_wall_init /
fe_wall_update against the oracle bit for bit (no GPU call is made).
"""
import math

import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

F = np.float32


def oracle_seq(oracle_mod, base, amp, freq, dts):
    w = oracle_mod.make_wall((base, 0.0, -1.0), (0.8, 4.0, 1.0), amp, freq)
    out = []
    for dt in dts:
        oracle_mod.wall_update(w, dt)
        out.append((w.time, w.min[0]))
    return np.array(out, dtype=np.float64)


def test_new_defaults():
    """CollisionBox::new: amplitude 0, frequency 1, time 0, base_min_x = min.x (collision.rs:20-29)."""
    b = fe.CollisionBox((-1.5, 0.0, -1.0), (0.8, 4.0, 1.0))
    assert b.wave_amplitude == 0.0 and b.wave_frequency == 1.0 and b.time == 0.0
    assert b.base_min_x == F(-1.5)
    assert b.min.tolist() == [F(-1.5), 0.0, -1.0] and b.max.tolist() == [F(0.8), 4.0, 1.0]


def test_zero_amplitude_is_static(oracle_mod):
    seq = oracle_seq(oracle_mod, -1.5, 0.0, 0.5, [0.005] * 2000)
    assert (seq[:, 1] == F(-1.5)).all()


def test_phase_closed_forms(oracle_mod):
    """frequency 1/4 Hz and dt 1 s put the phase at k pi/2 in float (TAU/4 is the float of
    pi/2 exactly): sin is then 1, ~0, -1, ~0, so min.x = base + A, base, base - A, base."""
    base, amp = -1.5, 0.25
    seq = oracle_seq(oracle_mod, base, amp, 0.25, [1.0] * 4)
    assert seq[:, 0].tolist() == [1.0, 2.0, 3.0, 4.0]
    assert seq[0, 1] == -1.25
    assert abs(seq[1, 1] - base) < 1e-6
    assert seq[2, 1] == -1.75
    assert abs(seq[3, 1] - base) < 1e-6


def test_wave_against_float64(oracle_mod):
    """time is the float sum of the steps; min.x must follow base + A sin(2 pi f t) within the
    float rounding of the phase (~|phase| 2^-23) and of the two final operations."""
    rng = np.random.default_rng(3)
    for amp, freq, dt, frames in [(0.3, 0.5, 0.005, 2000), (0.05, 5.0, 0.001, 3000), (-0.2, 1.7, 0.0123, 500)]:
        dts = [dt] * frames
        if amp < 0:
            dts = list(rng.uniform(0.001, 0.02, frames))
        seq = oracle_seq(oracle_mod, -1.5, amp, freq, dts)
        t = np.float32(0.0)
        times = []
        for d in dts:
            t = np.float32(t + np.float32(d))
            times.append(float(t))
        assert np.array_equal(seq[:, 0], np.array(times))
        phase = 2.0 * math.pi * float(F(freq)) * np.array(times)
        want = -1.5 + float(F(amp)) * np.sin(phase)
        tol = abs(amp) * (np.abs(phase) + 1.0) * 2.5e-7 + 2.5e-7
        err = np.abs(seq[:, 1] - want)
        assert (err <= tol).all(), (amp, freq, err.max())
        # the wave actually moves: it spans most of [base - A, base + A]
        assert seq[:, 1].max() - seq[:, 1].min() > 1.5 * abs(amp)


def test_library_update_matches_oracle(oracle_mod):
    """fe_wall_update (the product's host sequencer) equals the oracle's bit for bit."""
    rng = np.random.default_rng(5)
    cases = [(-1.5, 0.3, 0.5, [0.005] * 3000)]
    for _ in range(5):
        cases.append((float(rng.uniform(-3, 0)), float(rng.uniform(-0.5, 0.5)), float(rng.uniform(0.05, 8.0)),
                      list(rng.uniform(1e-4, 0.05, 400))))
    for base, amp, freq, dts in cases:
        b = fe.CollisionBox((base, 0.0, -1.0), (0.8, 4.0, 1.0))
        b.wave_amplitude = amp
        b.wave_frequency = freq
        w = oracle_mod.make_wall((base, 0.0, -1.0), (0.8, 4.0, 1.0), amp, freq)
        for dt in dts:
            b.update(dt)
            oracle_mod.wall_update(w, dt)
            assert F(b.time).view(np.uint32) == F(w.time).view(np.uint32)
            assert F(b.min[0]).view(np.uint32) == F(w.min[0]).view(np.uint32), (base, amp, freq, dt)


def test_contains():
    b = fe.CollisionBox((-1.5, 0.0, -1.0), (0.8, 4.0, 1.0))
    assert b.contains((0.0, 1.0, 0.0))
    assert b.contains((-1.5, 0.0, -1.0)) and b.contains((0.8, 4.0, 1.0))
    assert not b.contains((0.81, 1.0, 0.0))


def test_frame_order_in_oracle(oracle_mod):
    """Frame k runs with the wall of time t_k: the first frame with base_min_x, the second with
    base + A sin(dt f TAU); box w = 0; the wall's time is the sum of the frames' max_dt."""
    w = scene.rest_density_scene()
    dt = float(w.params.dt)
    orc = oracle_mod.Oracle(w.params, w.positions)
    try:
        wall = oracle_mod.make_wall((-0.79, -0.01, -1.0), (0.8, 4.0, 1.0), 0.05, 5.0)
        orc.set_wall(wall)
        expect = oracle_mod.make_wall((-0.79, -0.01, -1.0), (0.8, 4.0, 1.0), 0.05, 5.0)
        for k in range(3):
            orc.step()
            p = orc.params
            assert list(p.box_min) == [expect.min[0], F(-0.01), -1.0, 0.0], k
            assert list(p.box_max) == [F(0.8), 4.0, 1.0, 0.0], k
            oracle_mod.wall_update(expect, dt)
            got = orc.wall()
            assert got.time == expect.time and got.min[0] == expect.min[0]
        assert expect.min[0] != F(-0.79)
        orc.set_wall(None)
        assert orc.wall() is None
    finally:
        orc.close()


def test_app_scene():
    """Scene::new: 25 x 50 x 20 particles, mass rho0 s^3, the wave 0.3 m / 0.5 Hz."""
    w = scene.app_scene()
    assert w.n == 25000
    s = F(0.04)
    assert w.params.particle_mass == F(1000.0) * (s * s * s)
    assert w.wall_wave == (0.3, 0.5)
    assert list(w.params.box_min)[:3] == [F(-1.5), 0.0, -1.0]
    b = fe.CollisionBox.from_params(w.params, *w.wall_wave)
    assert b.base_min_x == F(-1.5) and b.wave_amplitude == F(0.3)


def test_wall_null_arguments():
    """The wall entry points reject null pointers with FE_ERR_INVALID_ARGUMENT."""
    lib = fe.load_library()
    w = fe.FeWall()
    assert lib.fe_set_wall(None, w) == 1
    assert lib.fe_wall_update(None, 0.1) == 1
    assert lib.fe_wall_init(None, None, None) == 1


@pytest.mark.parametrize("freq", [0.5, 3.0])
def test_library_update_large_time(oracle_mod, freq):
    """Long runs: time grows to ~100 s in float and both sides still agree bit for bit."""
    b = fe.CollisionBox((-1.5, 0.0, -1.0), (0.8, 4.0, 1.0))
    b.wave_amplitude = 0.3
    b.wave_frequency = freq
    w = oracle_mod.make_wall((-1.5, 0.0, -1.0), (0.8, 4.0, 1.0), 0.3, freq)
    for _ in range(20000):
        b.update(0.005)
        oracle_mod.wall_update(w, 0.005)
    assert F(b.min[0]).view(np.uint32) == F(w.min[0]).view(np.uint32)
    assert b.time > 90.0
