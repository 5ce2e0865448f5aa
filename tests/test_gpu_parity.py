"""Parity of the HIP path (through the C ABI) with the CPU oracle, stage by stage and
frame by frame, from identical inputs.

Bars: integer work (hashes, sort permutation, grid_start, the reorder gather, stats) is
compared bit for bit. FP32 fields must satisfy |gpu - oracle| <= 1e-5 * max(|oracle|,
rms(oracle field)) element-wise (the relative tolerance stated for this path, made robust
for fields that cross zero).
"""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene
from fluid_engine._ffi import FeOptions

pytestmark = pytest.mark.gpu

TOL = 1e-5
F = np.float32


def opts(sort=fe.FE_SORT_RADIX, mode=fe.FE_SOLVER_FIXED, maxit=100, **kw):
    o = fe.default_options()
    o.sort_algorithm = sort
    o.solver_mode = mode
    o.max_iterations = maxit
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def assert_close(name, gpu, ref, tol=TOL):
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert gpu.shape == ref.shape, name
    assert np.isfinite(gpu).all(), f"{name}: non-finite values"
    rms = np.sqrt(np.mean(ref * ref)) if ref.size else 0.0
    bound = tol * np.maximum(np.abs(ref), rms) + 1e-30
    err = np.abs(gpu - ref)
    bad = err > bound
    assert not bad.any(), (f"{name}: {bad.sum()} of {bad.size} outside tolerance, "
                           f"max err {err.max():.3e}, worst rel {(err / np.maximum(np.abs(ref), rms + 1e-30)).max():.3e}")


def assert_exact(name, gpu, ref):
    gpu = np.asarray(gpu)
    ref = np.asarray(ref)
    assert gpu.shape == ref.shape, name
    if not np.array_equal(gpu.view(np.uint32) if gpu.dtype == np.float32 else gpu,
                          ref.view(np.uint32) if ref.dtype == np.float32 else ref):
        diff = np.nonzero(gpu.reshape(-1) != ref.reshape(-1))[0]
        raise AssertionError(f"{name}: {len(diff)} mismatches, first at {diff[:5].tolist()}")


def assert_both(name, gpu, ref):
    assert_close(name, gpu, ref)
    assert_exact(name, gpu, ref)


def pair(w, options, oracle_mod):
    sim = fe.Simulation(w.params, w.positions, device=0, options=options)
    orc = oracle_mod.Oracle(w.params, w.positions, options=options)
    return sim, orc


def sync_fields(sim, orc, names):
    """Copy the oracle's fields into the GPU so the next stage starts from identical input."""
    for n in names:
        sim.upload(n, orc.download(n))


SCENES = {
    "rest_density_4800": scene.rest_density_scene,
    "dam_break_32k": lambda: scene.dam_break_32k(seed=1),
}


@pytest.fixture(scope="module", params=list(SCENES))
def workload(request):
    w = SCENES[request.param]()
    print(f"[gpu parity] scene={w.name} n={w.n}", flush=True)
    return w


@pytest.mark.parametrize("sort", [fe.FE_SORT_RADIX, fe.FE_SORT_BITONIC])
def test_neighbor_search_bit_exact(workload, sort, oracle_mod):
    sim, orc = pair(workload, opts(sort=sort), oracle_mod)
    sim.neighbor_search()
    orc.neighbor_search()
    for name in ("grid_entries", "grid_start", "position_b", "velocity_b"):
        assert_exact(name, sim.download(name), orc.download(name))


def test_stage_by_stage(workload, oracle_mod):
    """Every ComputeStep from identical inputs (the stated bar is 1e-5 relative; the HIP
    kernels evaluate the reference's operation sequence, so they also agree bit for bit)."""
    rng = np.random.default_rng(7)
    sim, orc = pair(workload, opts(), oracle_mod)
    n = workload.n
    sim.neighbor_search()
    orc.neighbor_search()
    sim.density_alpha()
    orc.density_alpha()
    assert_both("densities", sim.download("densities"), orc.download("densities"))
    assert_both("factors", sim.download("factors"), orc.download("factors"))
    sync_fields(sim, orc, ["densities", "factors"])

    v = np.zeros((n, 4), F)
    v[:, :3] = rng.normal(0, 0.5, (n, 3))
    for o in (sim, orc):
        o.upload("velocity_b", v)
        o.viscosity()
    assert_both("viscosity velocity_a", sim.download("velocity_a"), orc.download("velocity_a"))
    sync_fields(sim, orc, ["velocity_a"])

    for o in (sim, orc):
        o.density_source()
    assert_both("density source_terms", sim.download("source_terms"), orc.download("source_terms"))
    assert_exact("pressures zeroed", sim.download("pressures"), orc.download("pressures"))
    sync_fields(sim, orc, ["source_terms"])

    p = rng.uniform(0, 2000, n).astype(F)
    for o in (sim, orc):
        o.upload("pressures", p)
        o.pressure_force()
    assert_both("pressure_accelerations", sim.download("pressure_accelerations"),
                 orc.download("pressure_accelerations"))
    sync_fields(sim, orc, ["pressure_accelerations"])
    for o in (sim, orc):
        o.pressure_update()
    assert_both("pressure_update pressures", sim.download("pressures"), orc.download("pressures"))
    sync_fields(sim, orc, ["pressures"])

    for o in (sim, orc):
        o.pressure_integration()
    assert_both("integrated position_a", sim.download("position_a"), orc.download("position_a"))
    assert_both("integrated velocity_a", sim.download("velocity_a"), orc.download("velocity_a"))
    sync_fields(sim, orc, ["position_a", "velocity_a"])

    for o in (sim, orc):
        o.neighbor_search()
    for name in ("grid_entries", "grid_start", "position_b", "velocity_b"):
        assert_exact("post-integrate " + name, sim.download(name), orc.download(name))
    for o in (sim, orc):
        o.density_alpha()
    sync_fields(sim, orc, ["densities", "factors"])
    for o in (sim, orc):
        o.divergence_source()
    assert_both("divergence source_terms", sim.download("source_terms"), orc.download("source_terms"))
    sync_fields(sim, orc, ["source_terms"])
    for o in (sim, orc):
        o.divergence_integration()
    assert_both("divergence velocity_a", sim.download("velocity_a"), orc.download("velocity_a"))
    assert_both("colors", sim.download("colors"), orc.download("colors"))
    sync_fields(sim, orc, ["velocity_a"])
    for o in (sim, orc):
        o.stats()
    assert_exact("stats", sim.read_stats(), orc.read_stats())


FIELDS_STEP = ("position_a", "velocity_a", "position_b", "velocity_b", "densities", "factors",
               "source_terms", "pressures", "pressure_accelerations", "colors")


@pytest.mark.parametrize("sort", [fe.FE_SORT_RADIX, fe.FE_SORT_BITONIC])
def test_frames_match_oracle(workload, sort, oracle_mod):
    """Whole frames in the reference arithmetic reproduce the oracle bit for bit."""
    sim, orc = pair(workload, opts(sort=sort), oracle_mod)
    for frame in range(3):
        sim.step()
        orc.step()
        assert_exact(f"frame {frame} grid_entries", sim.download("grid_entries"), orc.download("grid_entries"))
        assert_exact(f"frame {frame} grid_start", sim.download("grid_start"), orc.download("grid_start"))
        for name in FIELDS_STEP:
            assert_exact(f"frame {frame} {name}", sim.download(name), orc.download(name))
        assert_exact(f"frame {frame} stats", sim.read_stats(), orc.read_stats())


def rms(x):
    x = np.asarray(x, dtype=np.float64)
    return float(np.sqrt(np.mean(x * x)))


@pytest.mark.parametrize("scene_name", ["rest_density_4800", "dam_break_32k"])
def test_fp64_error_report(scene_name, oracle_mod):
    """FP64 re-evaluation of the density phase (neighbour search through integration, four
    Jacobi sweeps) from one FP32 state: the GPU's per-field error against FP64 is stated and
    equals the FP32 reference arithmetic's (the two agree bit for bit)."""
    w = SCENES[scene_name]()
    warm = oracle_mod.Oracle(w.params, w.positions)
    for _ in range(3):
        warm.step()
    state = {k: warm.download(k) for k in ("position_a", "velocity_a")}
    sim = fe.Simulation(w.params, w.positions, device=0, options=opts())
    o32 = oracle_mod.Oracle(w.params, w.positions)
    o64 = oracle_mod.Oracle(w.params, w.positions, fp64=True)
    for ev in (sim, o32, o64):
        for k, v in state.items():
            ev.upload(k, v)
        ev.neighbor_search()
        ev.density_alpha()
        ev.viscosity()
        ev.density_source()
        for _ in range(4):
            ev.pressure_force()
            ev.pressure_update()
        ev.pressure_integration()
    for name in ("densities", "factors", "source_terms", "pressures", "pressure_accelerations", "velocity_a",
                 "position_a"):
        ref = o64.download(name, as_double=True)
        g = sim.download(name)
        assert_exact(name, g, o32.download(name))
        scale = rms(ref)
        err = np.abs(g.astype(np.float64) - ref)
        rel = err.max() / scale if scale else 0.0
        print(f"[fp64] {scene_name} {name}: max |fp32 - fp64| / rms = {rel:.3e}, rms err / rms = "
              f"{(rms(err) / scale if scale else 0.0):.3e}")
        assert np.isfinite(g).all()


def test_list_fallbacks_match_oracle(oracle_mod):
    """Particles with more than 64 gradient neighbours, and a neighbour list made stale by
    an upload of position_b, both take the bucket walk and still match the oracle."""
    rng = np.random.default_rng(3)
    w = scene.dam_break_32k(seed=4)
    pos = w.positions.copy()
    # a clump of 300 particles inside one half-smoothing-length cube: ~299 pairs each
    c0 = np.array([0.5, 0.3, 0.5], np.float32)
    pos[:300] = c0 + rng.uniform(0, 0.02, (300, 3)).astype(np.float32)
    sim, orc = pair(scene.Workload("clump", pos, w.params), opts(), oracle_mod)
    for o in (sim, orc):
        o.neighbor_search()
        o.density_alpha()
        o.viscosity()
        o.density_source()
        o.pressure_force()
        o.pressure_update()
    for name in ("densities", "factors", "velocity_a", "source_terms", "pressure_accelerations", "pressures"):
        assert_exact("clump " + name, sim.download(name), orc.download(name))
    # stale list: move position_b after density_alpha; the reference gathers use the moved
    # positions, so must the HIP path
    pb = orc.download("position_b")
    pb[:, :3] += rng.normal(0, 0.002, (pb.shape[0], 3)).astype(np.float32)
    for o in (sim, orc):
        o.upload("position_b", pb)
        o.viscosity()
        o.pressure_force()
    for name in ("velocity_a", "pressure_accelerations"):
        assert_exact("stale-list " + name, sim.download(name), orc.download(name))


@pytest.mark.parametrize("scale", [0.9, 1.1])
def test_new_smoothing_radius_between_stages(scale, oracle_mod):
    """A new h between stages (fe_set_params after density_alpha): the listed pairs were
    those within the old h, so the following gathers walk the buckets with the new h, as
    the oracle does, in place of reusing the list."""
    w = scene.dam_break_32k(seed=5)
    sim, orc = pair(w, opts(), oracle_mod)
    vb = np.zeros((w.n, 4), np.float32)
    vb[:, :3] = np.random.default_rng(7).normal(0, 0.05, (w.n, 3)).astype(np.float32)
    for o in (sim, orc):
        o.neighbor_search()
        o.density_alpha()
        o.upload("velocity_b", vb)
    p = fe.FeParams.from_buffer_copy(sim.params)
    p.smoothing_radius = F(p.smoothing_radius) * F(scale)
    for o in (sim, orc):
        o.set_params(p)
        o.viscosity()
        o.density_source()
        o.pressure_force()
        o.pressure_update()
        o.pressure_force()
    for name in ("velocity_a", "source_terms", "pressures", "pressure_accelerations"):
        assert_exact(f"h x {scale} {name}", sim.download(name), orc.download(name))
    assert np.count_nonzero(orc.download("pressure_accelerations")[:, :3]) > 0


def test_converge_mode_iterations_match(oracle_mod):
    w = scene.dam_break_32k(seed=2)
    o = opts(mode=fe.FE_SOLVER_CONVERGE, maxit=100)
    sim, orc = pair(w, o, oracle_mod)
    for frame in range(3):
        sim.step()
        orc.step()
        assert sim.last_iterations() == orc.last_iterations(), frame
        for name in ("position_a", "velocity_a", "pressures"):
            assert_exact(f"converge frame {frame} {name}", sim.download(name), orc.download(name))
    d, v, frames = sim.iteration_totals()
    assert frames == 3 and d >= 3 and v >= 3


def test_converge_two_sided_residual_matches(oracle_mod):
    """FE_RESIDUAL_TWO_SIDED (no reference counterpart): sweep counts, fields and the exit
    reports' inputs equal the oracle's, frame by frame, and the two-sided test needs at least
    as many sweeps as the compression-only one."""
    w = scene.dam_break_32k(seed=2)
    o = opts(mode=fe.FE_SOLVER_CONVERGE, maxit=100)
    o.residual = fe.FE_RESIDUAL_TWO_SIDED
    sim, orc = pair(w, o, oracle_mod)
    one = fe.Simulation(w.params, w.positions, device=0, options=opts(mode=fe.FE_SOLVER_CONVERGE, maxit=100))
    for frame in range(4):
        sim.step()
        orc.step()
        one.step()
        assert sim.last_iterations() == orc.last_iterations(), frame
        for name in ("position_a", "velocity_a", "pressures"):
            assert_exact(f"two-sided frame {frame} {name}", sim.download(name), orc.download(name))
        if frame == 0:
            d2, v2 = sim.last_iterations()
            d1, v1 = one.last_iterations()
            assert d2 >= d1 and v2 >= v1
    one.close()


def test_solve_report_closed_form(oracle_mod):
    """fe_solve_report of a FIXED solve: its inputs are the last sweep's r = s - Ap, which the
    oracle reproduces (one sweep from p = 0: Ap = 0 for the density solve, r = s)."""
    w = scene.dam_break_32k(seed=4)
    o = opts()
    sim, orc = pair(w, o, oracle_mod)
    sim.step()
    orc.step()
    sim.neighbor_search()
    sim.density_alpha()
    sim.viscosity()
    sim.density_source()
    s = sim.download("source_terms").astype(np.float64)
    p = w.params
    sim.solve_report_enable(True)
    sim.set_params(scene.make_params(p.particle_radius, p.particle_mass, p.smoothing_radius, p.target_density,
                                     p.viscosity, p.relax_factor, p.dt, 1, 1, tuple(p.gravity[:3]), tuple(p.box_min[:3]),
                                     tuple(p.box_max[:3])))
    sim.solve_density()
    rep = sim.solve_report(0)
    assert rep["solves"] == 1
    # first sweep from p = 0: a = 0, so Ap = 0 and r = s exactly
    scale = float(np.float32(p.dt)) / float(np.float32(p.target_density))
    r = s * scale
    assert abs(rep["mean_abs"] - np.abs(r).mean()) <= 1e-6 * max(np.abs(r).mean(), 1e-12) + 2 ** -24
    assert abs(rep["mean_expansion"] - np.maximum(r, 0).mean()) <= 1e-6 * np.abs(r).mean() + 2 ** -24
    assert abs(rep["max_abs"] - np.abs(r).max()) <= 2 ** -23 + 1e-6 * np.abs(r).max()
    pres = sim.download("pressures")
    want = (np.maximum(r, 0) * (pres > 0)).mean()
    assert abs(rep["mean_expansion_pressured"] - want) <= 1e-6 * np.abs(r).mean() + 2 ** -24
    sim.close()


def test_host_flags_cfl_and_early_exit(oracle_mod):
    w = scene.rest_density_scene()
    o = opts(use_cfl=1, use_solver_error_threshold=1)
    sim, orc = pair(w, o, oracle_mod)
    for frame in range(4):
        sim.step(w.params.dt)
        orc.step(w.params.dt)
        assert sim.last_iterations() == orc.last_iterations(), frame
        assert sim.params.dt == orc.params.dt, frame
        assert_exact(f"flags frame {frame} stats", sim.read_stats(), orc.read_stats())
        assert_close(f"flags frame {frame} position_a", sim.download("position_a"), orc.download("position_a"))


@pytest.mark.parametrize("log2n", [20, 22])
def test_radix_sort_random_keys_permutation(log2n):
    n = 1 << log2n
    print(f"[gpu parity] radix 2^{log2n}", flush=True)
    rng = np.random.default_rng(log2n)
    keys = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    entries = np.stack([keys, np.arange(n, dtype=np.uint32)], axis=1)
    sim = fe.Simulation(scene.make_params(0.01, 0.008, 0.04, 1000, 0.01, 0.5, 1e-3, 1, 1, (0, -9.81, 0),
                                          (0, 0, 0), (1, 1, 1)), np.zeros((n, 3), F), device=0)
    sim.upload("grid_entries", entries)
    sim.sort()
    out = sim.download("grid_entries")
    perm = np.argsort(keys, kind="stable")
    assert_exact("radix permutation", out[:, 1], perm.astype(np.uint32))
    assert_exact("radix keys", out[:, 0], keys[perm])


def test_bitonic_network_matches_oracle(oracle_mod):
    n = 1 << 16
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 4096, size=n).astype(np.uint32)   # many equal keys: order is the network's
    entries = np.stack([keys, np.arange(n, dtype=np.uint32)], axis=1)
    sim = fe.Simulation(scene.make_params(0.01, 0.008, 0.04, 1000, 0.01, 0.5, 1e-3, 1, 1, (0, -9.81, 0),
                                          (0, 0, 0), (1, 1, 1)), np.zeros((n, 3), F), device=0,
                        options=opts(sort=fe.FE_SORT_BITONIC))
    sim.upload("grid_entries", entries)
    sim.sort()
    assert_exact("bitonic", sim.download("grid_entries"), oracle_mod.bitonic_sort_pairs(entries))


def test_dam_break_1m_neighbor_search_and_density(oracle_mod):
    w = scene.dam_break_1m(seed=0)
    print(f"[gpu parity] {w.name} n={w.n}", flush=True)
    sim, orc = pair(w, opts(), oracle_mod)
    sim.neighbor_search()
    orc.neighbor_search()
    for name in ("grid_entries", "grid_start", "position_b"):
        assert_exact("1M " + name, sim.download(name), orc.download(name))
    sim.density_alpha()
    orc.density_alpha()
    assert_close("1M densities", sim.download("densities"), orc.download("densities"))
    assert_close("1M factors", sim.download("factors"), orc.download("factors"))


def test_dam_break_1m_frames(oracle_mod):
    """The benchmark's own frames: the 1M dam break from rest in converge mode with the radix
    sort, autotuned after the second frame as the benchmark does, bit for bit with the oracle
    (stats and sweep counts every frame, every field after frames 1 and 3)."""
    w = scene.dam_break_1m(seed=0)
    print(f"[gpu parity] {w.name} n={w.n} frames", flush=True)
    o = opts(mode=fe.FE_SOLVER_CONVERGE, density_tol=1e-3, divergence_tol=1e-2)
    with fe.Simulation(w.params, w.positions, device=0, options=o) as sim:
        orc = oracle_mod.Oracle(w.params, w.positions, options=o)
        try:
            for frame in range(4):
                if frame == 2:
                    sim.autotune_now(reps=3)
                sim.step()
                orc.step()
                assert_exact(f"1M frame {frame} stats", sim.read_stats(), orc.read_stats())
                assert sim.last_iterations() == orc.last_iterations(), frame
                if frame in (1, 3):
                    for name in FIELDS_STEP + ("grid_entries", "grid_start"):
                        assert_exact(f"1M frame {frame} {name}", sim.download(name), orc.download(name))
        finally:
            orc.close()


def test_block_4m_neighbor_search_and_density(oracle_mod):
    """The 2^22-particle block (256 x 128 x 128 in a 6 x 3 x 3 m box, T = 2^22): hash, sort,
    grid_start, reorder, densities and factors bit for bit, in both sorts."""
    w = scene.block_4m(seed=0)
    for sort in (fe.FE_SORT_RADIX, fe.FE_SORT_BITONIC):
        print(f"[gpu parity] {w.name} n={w.n} sort={sort}", flush=True)
        orc = oracle_mod.Oracle(w.params, w.positions, options=opts(sort=sort))
        try:
            orc.neighbor_search()
            orc.density_alpha()
            want = {name: orc.download(name) for name in
                    ("grid_entries", "grid_start", "position_b", "velocity_b", "densities", "factors")}
        finally:
            orc.close()
        with fe.Simulation(w.params, w.positions, device=0, options=opts(sort=sort)) as sim:
            sim.neighbor_search()
            sim.density_alpha()
            for name, ref in want.items():
                assert_exact(f"4M sort {sort} {name}", sim.download(name), ref)
