"""Host-side checks of the C-ABI library (no GPU compute): it loads, it exports every entry
point include/fluid_engine.h declares, the parameter block matches SimulationParams, the
host frame control follows Renderer::step, and without a device it fails loudly."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene
from fluid_engine._ffi import FeOptions, FeParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fluid_engine.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fe_[a-z0-9_]+)\s*\(", text)))


def test_library_built():
    assert os.path.exists(fe.LIB_PATH), "build libfluid_engine.so first (make -C fluid-engine_amd)"


def test_every_declared_symbol_is_exported():
    lib = fe.load_library()
    syms = declared_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(fe.EXPORTED_SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", fe.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (fe_[a-z0-9_]+)$", out, flags=re.M))
    assert set(syms) <= exported


def test_library_is_gfx950():
    out = subprocess.run(["strings", fe.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "gfx950" in out


def test_params_layout_matches_simulation_params():
    # SimulationParams (src/entities/particle.rs:310-331): repr(C, align(16)), 112 bytes
    assert ctypes.sizeof(FeParams) == 112
    off = {name: getattr(FeParams, name).offset for name, _ in FeParams._fields_}
    assert off["dt"] == 24 and off["density_solver_iterations"] == 28
    assert off["gravity"] == 48 and off["box_min"] == 64 and off["box_max"] == 80 and off["grid_res"] == 96


def _opts(**kw):
    o = FeOptions()
    fe.load_library().fe_default_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def test_default_options():
    o = _opts()
    assert o.sort_algorithm == fe.FE_SORT_RADIX        # ui.rs:36
    assert o.use_cfl == 0 and o.use_solver_error_threshold == 0
    assert abs(o.density_error_pct - 0.1) < 1e-7 and abs(o.divergence_error_pct - 0.5) < 1e-7
    assert o.max_iterations == 100


def test_frame_control_early_exit_and_cfl():
    p = scene.make_params(0.02, 0.064, 0.08, 1000.0, 0.15, 0.5, 0.005, 4, 4, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    n = 1000
    # flags off: the scene's sweep counts, dt unchanged
    assert fe.frame_control([0, 0, 0], n, p, _opts()) == (pytest.approx(0.005), 4, 4)
    # stats start as zeros (particle.rs:270-281): with the early-exit flag both errors are
    # below threshold on the first frame -> 2 / 1 sweeps (mod.rs:251-265)
    assert fe.frame_control([0, 0, 0], n, p, _opts(use_solver_error_threshold=1))[1:] == (2, 1)
    # avg density error = 1500/1000 = 1.5 kg/m3 > 0.1 % of 1000 -> keep 4; divergence
    # 40/(10*1000) = 0.004 < 5 -> 1
    assert fe.frame_control([0, 1500, 40], n, p, _opts(use_solver_error_threshold=1))[1:] == (4, 1)
    # CFL: max speed 2 m/s -> 0.4 h / v = 0.016, capped by 1.1 * dt = 0.0055
    v2 = int(np.float32(2.0).view(np.uint32))
    dt, _, _ = fe.frame_control([v2, 0, 0], n, p, _opts(use_cfl=1))
    assert dt == pytest.approx(0.0055)
    # fast flow: 0.4 * 0.08 / 40 = 0.0008 -> clamped to the 1 ms floor
    v40 = int(np.float32(40.0).view(np.uint32))
    assert fe.frame_control([v40, 0, 0], n, p, _opts(use_cfl=1))[0] == pytest.approx(0.001)
    # below 0.01 m/s the CFL update is skipped
    vs = int(np.float32(0.005).view(np.uint32))
    assert fe.frame_control([vs, 0, 0], n, p, _opts(use_cfl=1))[0] == pytest.approx(0.005)


def test_frame_control_rejects_bad_arguments():
    p = scene.make_params(0.02, 0.064, 0.08, 1000.0, 0.15, 0.5, 0.005, 4, 4, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    with pytest.raises(fe.FluidEngineError):
        fe.frame_control([0, 0, 0], 0, p, _opts())


def test_create_without_device_fails_loudly():
    w = scene.rest_density_scene()
    try:
        sim = fe.Simulation(w.params, w.positions)
    except fe.FluidEngineError as e:
        assert e.status == 3, str(e)      # FE_ERR_DEVICE: no silent CPU fallback
        assert "device" in str(e).lower()
        return
    sim.close()
    pytest.skip("a HIP device is present")
