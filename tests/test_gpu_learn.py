"""The learned warm start on the GPU, through the C ABI: the MFMA forward and backward
kernels against the FP64 MLP oracle (oracle/mlp.py, BF16 rounding points emulated), the
Adam step, the feature gather against the C++ oracle, the prediction path, and what the
trained network gains over the reference start (sweeps of the converging solver).

Tolerances. The kernels hold BF16 operands and accumulate in FP32; the oracle rounds to
BF16 at the same points and computes in FP64. A BF16 rounding that lands on the other
side of a tie (FP32 vs FP64 pre-activation) moves one activation by 2^-8 relative, so
outputs agree to ~1e-3 absolute at O(1) values, and gradients to ~1e-2 in relative norm.
"""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

pytestmark = pytest.mark.gpu

F = np.float32
NORM_ID = np.concatenate([np.zeros(12), np.ones(12), [0.0, 1.0]]).astype(F)


@pytest.fixture(scope="module")
def mlp():
    from oracle import mlp as m
    return m


@pytest.fixture(scope="module")
def small_sim():
    w = scene.rest_density_scene()
    sim = fe.Simulation(w.params, w.positions, device=0)
    yield sim
    sim.close()


def conv_opts(warm=fe.FE_WARM_START_REFERENCE):
    o = fe.default_options()
    o.solver_mode = fe.FE_SOLVER_CONVERGE
    o.density_tol = 1e-3
    o.divergence_tol = 1e-2
    o.max_iterations = 100
    o.warm_start = warm
    return o


@pytest.mark.parametrize("n", [1, 17, 1000, 4096])
def test_forward_matches_oracle(small_sim, mlp, n):
    p = mlp.init_params(1).astype(F)
    small_sim.learn_set_weights(fe.FE_NET_DENSITY, p, NORM_ID)
    x = np.random.default_rng(n).normal(size=(n, 12)).astype(F)
    y = small_sim.learn_forward(fe.FE_NET_DENSITY, x)
    ref, _ = mlp.forward(p.astype(np.float64), x, bf16_points=True)
    err = np.abs(y - ref)
    print(f"n={n}: max |gpu - oracle(bf16 points)| = {err.max():.2e}, mean {err.mean():.2e}")
    assert np.isfinite(y).all()
    assert err.max() <= 5e-3 * (1 + np.abs(ref).max())
    assert err.mean() <= 2e-4
    ref64, _ = mlp.forward(p.astype(np.float64), x, bf16_points=False)
    assert np.abs(y - ref64).max() <= 5e-2 * (1 + np.abs(ref64).max())


@pytest.mark.parametrize("n", [32, 1000, 4096])
def test_gradient_matches_oracle(small_sim, mlp, n):
    p = mlp.init_params(2).astype(F)
    small_sim.learn_set_weights(fe.FE_NET_DIVERGENCE, p, NORM_ID)
    rng = np.random.default_rng(100 + n)
    x = rng.normal(size=(n, 12)).astype(F)
    t = rng.normal(size=n).astype(F)
    g, sse = small_sim.learn_gradient(fe.FE_NET_DIVERGENCE, x, t)
    sse_ref, g_ref = mlp.loss_grad(p.astype(np.float64), x, t, bf16_points=True)
    assert np.isfinite(g).all()
    assert abs(sse - sse_ref) <= 1e-3 * sse_ref
    offs = [0, 768, 832, 4928, 4992, 9088, 9152, 9216, 9217]
    names = ["W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4"]
    for name, a, b in zip(names, offs, offs[1:]):
        rel = np.linalg.norm(g[a:b] - g_ref[a:b]) / (np.linalg.norm(g_ref[a:b]) + 1e-30)
        print(f"n={n} {name}: relative gradient error {rel:.2e}")
        assert rel < 1e-2, name
    _, g64 = mlp.loss_grad(p.astype(np.float64), x, t, bf16_points=False)
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 5e-2


def collect_rest_scene(per_solve=256, capacity=4096, frames=2):
    w = scene.rest_density_scene()
    sim = fe.Simulation(w.params, w.positions, device=0, options=conv_opts())
    sim.learn_collect(per_solve, capacity)
    sim.run(frames)
    sim.sync()
    return sim


def test_training_step_matches_oracle_adam(mlp):
    sim = collect_rest_scene()
    try:
        net = fe.FE_NET_DENSITY
        x, t = sim.learn_dataset(net)
        assert len(t) == 512
        sim.learn_train(net, steps=0, seed=7)
        p0, norm = sim.learn_weights(net)
        # normalisation in double from the dataset, as documented
        xd = x.astype(np.float64)
        mean = xd.mean(axis=0)
        sd = xd.std(axis=0)
        sd = np.where(sd > 1e-12 * (1 + np.abs(mean)), sd, 1.0)
        assert np.allclose(norm[:12], mean, rtol=1e-6, atol=1e-30)
        assert np.allclose(norm[12:24], sd, rtol=1e-6)
        assert np.isclose(norm[24], t.astype(np.float64).mean(), rtol=1e-6, atol=1e-30)
        xs = ((x - norm[:12]) / norm[12:24]).astype(F)
        ts = ((t - norm[24]) / norm[25]).astype(F)
        steps, lr = 3, 1e-3
        losses = sim.learn_train(net, steps=steps, batch=65536, lr=lr, seed=7)
        p3, _ = sim.learn_weights(net)
        p = p0.astype(np.float64)
        m1 = np.zeros_like(p)
        m2 = np.zeros_like(p)
        for k in range(steps):
            sse, g = mlp.loss_grad(p, xs, ts, bf16_points=True)
            assert abs(losses[k] - sse / len(ts)) <= 2e-3 * sse / len(ts), k
            lrk = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * k / (steps - 1))))
            p, m1, m2 = mlp.adam_step(p, m1, m2, g, k + 1, lrk)
        close = np.abs(p3 - p) <= 0.05 * lr * steps
        print(f"Adam: {close.mean() * 100:.2f}% of parameters within 5% of one step")
        assert close.mean() >= 0.99
        # deterministic: the same seed gives the same weights, bit for bit
        sim.learn_train(net, steps=steps, batch=65536, lr=lr, seed=7)
        p3b, _ = sim.learn_weights(net)
        assert np.array_equal(p3, p3b)
    finally:
        sim.close()


def stages_to(x, point, source=True):
    """Run the reference stages of one frame on a Simulation or an Oracle up to the source
    term of the solve `point` ("density" or "divergence"), that stage included if `source`."""
    x.neighbor_search()
    x.density_alpha()
    x.viscosity()
    if point == "density":
        if source:
            x.density_source()
        return
    x.density_source()
    x.solve_density()
    x.pressure_integration()
    x.neighbor_search()
    x.density_alpha()
    if source:
        x.divergence_source()


def finish_frame(x, point):
    if point == "density":
        x.solve_density()
        x.pressure_integration()
        x.neighbor_search()
        x.density_alpha()
        x.divergence_source()
    x.solve_divergence()
    x.divergence_integration()
    x.stats()


@pytest.mark.parametrize("net,point", [(fe.FE_NET_DENSITY, "density"), (fe.FE_NET_DIVERGENCE, "divergence")])
def test_features_and_previous_pressure_match_oracle(oracle_mod, net, point):
    w = scene.dam_break_32k(seed=2)
    o = fe.default_options()
    sim = fe.Simulation(w.params, w.positions, device=0, options=o)
    orc = oracle_mod.Oracle(w.params, w.positions, options=o)
    try:
        sim.learn_collect(0)       # keeps the pressure history without sampling
        for _ in range(3):
            sim.step()
            orc.step()
        # one frame by stages, keeping the pressures each solve ends with
        stages_to(sim, "density")
        sim.solve_density()
        p_dens = sim.download("pressures")
        sim.pressure_integration()
        sim.neighbor_search()
        sim.density_alpha()
        sim.divergence_source()
        sim.solve_divergence()
        p_div = sim.download("pressures")
        sim.divergence_integration()
        sim.stats()
        stages_to(orc, "divergence")
        finish_frame(orc, "divergence")
        # the next frame up to the solve under test; on the GPU the source-term stage is the
        # one that gathers the features
        stages_to(sim, point, source=False)
        f, pp = sim.learn_features(net)
        stages_to(orc, point)
        for name in ("densities", "factors", "source_terms", "velocity_a", "position_b", "pressures"):
            assert np.array_equal(sim.download(name), orc.download(name)), name
        vals = sim.download("grid_entries")[: w.n, 1]
        want_pp = (p_dens if net == fe.FE_NET_DENSITY else p_div)[vals]
        assert np.array_equal(pp, want_pp)
        ref = orc.features(pp)
        assert np.array_equal(f[:, 3], ref[:, 3])            # pair counts, exact
        for k in range(12):
            a, b = f[:, k].astype(np.float64), ref[:, k].astype(np.float64)
            rms = np.sqrt(np.mean(b * b)) + 1e-30
            err = np.abs(a - b) / np.maximum(np.abs(b), rms)
            assert err.max() <= 1e-5, (k, err.max())
        assert (pp > 0).any()
    finally:
        sim.close()


def test_prediction_matches_oracle_forward(mlp):
    w = scene.dam_break_32k(seed=3)
    sim = fe.Simulation(w.params, w.positions, device=0)
    try:
        sim.learn_collect(0)
        sim.run(3)
        stages_to(sim, "density", source=False)
        f, _ = sim.learn_features(fe.FE_NET_DENSITY)
        fd = f.astype(np.float64)
        norm = np.concatenate([fd.mean(0), np.maximum(fd.std(0), 1e-6), [150.0, 400.0]]).astype(F)
        p = mlp.init_params(4).astype(F)
        sim.learn_set_weights(fe.FE_NET_DENSITY, p, norm)
        sim.learn_predict(fe.FE_NET_DENSITY)
        pres = sim.download("pressures")
        xs = ((f - norm[:12]) / norm[12:24]).astype(F)
        y, _ = mlp.forward(p.astype(np.float64), xs, bf16_points=True)
        want = np.maximum(y * norm[25] + norm[24], 0.0)
        err = np.abs(pres - want)
        print(f"prediction: max err {err.max():.3e} (target std 400)")
        assert err.max() <= 5e-3 * 400.0 * (1 + np.abs(y).max())
        assert (pres >= 0).all()
        # the pressure solve runs from the prediction (its first sweep sees it)
        it = sim.solve_density()
        assert it >= 1
    finally:
        sim.close()


def test_learned_warm_start_needs_a_network():
    w = scene.rest_density_scene()
    sim = fe.Simulation(w.params, w.positions, device=0)
    try:
        with pytest.raises(fe.FluidEngineError):
            sim.set_options(conv_opts(fe.FE_WARM_START_LEARNED))
    finally:
        sim.close()


TRAIN_FRAMES = 100
EVAL_FRAMES = 200


def test_learned_warm_start_reduces_sweeps():
    """DB-32K, converging solver (0.1 % / 1 %, at most 100 sweeps): frames 0-99 collect
    2048 samples per solve with the reference start, both networks are trained on the GPU,
    then frames 100-299 run with the learned start; a second context runs all 300 frames
    with the reference start. Compares the sweeps of frames 100-299."""
    w = scene.dam_break_32k(seed=0)
    ref = fe.Simulation(w.params, w.positions, device=0, options=conv_opts())
    lrn = fe.Simulation(w.params, w.positions, device=0, options=conv_opts())
    try:
        lrn.learn_collect(2048, 1 << 18)
        for _ in range(TRAIN_FRAMES):
            ref.step()
            lrn.step()
        assert np.array_equal(ref.download("position_a"), lrn.download("position_a"))
        lrn.learn_collect(0)
        assert lrn.learn_dataset_size(0) == TRAIN_FRAMES * 2048
        l0 = lrn.learn_train(fe.FE_NET_DENSITY, steps=2000, batch=4096, lr=2e-3, seed=1)
        l1 = lrn.learn_train(fe.FE_NET_DIVERGENCE, steps=2000, batch=4096, lr=2e-3, seed=2)
        print(f"density net loss {l0[:20].mean():.4f} -> {l0[-50:].mean():.4f}; "
              f"divergence net loss {l1[:20].mean():.4f} -> {l1[-50:].mean():.4f}")
        assert l0[-50:].mean() < 0.5 * l0[:20].mean()
        lrn.set_options(conv_opts(fe.FE_WARM_START_LEARNED))
        ref.reset_iteration_totals()
        lrn.reset_iteration_totals()
        for _ in range(EVAL_FRAMES):
            ref.step()
            lrn.step()
        rd, rv, _ = ref.iteration_totals()
        ld, lv, _ = lrn.iteration_totals()
        gain = (rd + rv) / max(ld + lv, 1)
        print(f"sweeps over {EVAL_FRAMES} frames: reference {rd} + {rv} = {rd + rv}, "
              f"learned {ld} + {lv} = {ld + lv}, gain {gain:.2f}x "
              f"(density {rd / max(ld, 1):.2f}x)")
        assert gain > 1.0
        # the learned run is still a converged fluid: bounded density error, finite state
        assert np.isfinite(lrn.download("position_a")).all()
        err_l = lrn.read_stats()[1] / w.n
        err_r = ref.read_stats()[1] / w.n
        print(f"mean |rho - rho0|: reference {err_r:.2f}, learned {err_l:.2f} kg/m^3")
        assert err_l <= 1.5 * err_r + 5.0
    finally:
        ref.close()
        lrn.close()


def test_learned_warm_start_halves_sweeps_1m():
    """The 1,048,576-particle dam break: samples from frames 0-99 (2048 per solve, 204,800 per
    network), both networks trained on the GPU, then frames 100-499 with the learned start
    against a second context with the reference start. Both stop on the two-sided residual
    (FE_RESIDUAL_TWO_SIDED): a pressured particle counts |s - Ap|, so a start that over-predicts
    pressure is not taken as converged. The total Jacobi sweeps to the same 0.1 % / 1 %
    thresholds must drop at least 2x, and the exit residuals (the next 50 frames, reported)
    must be no worse than the reference start's."""
    w = scene.dam_break_1m(seed=0)
    ref = fe.Simulation(w.params, w.positions, device=0, options=conv_opts())
    lrn = fe.Simulation(w.params, w.positions, device=0, options=conv_opts())
    try:
        lrn.learn_collect(2048, 1 << 18)
        ref.run(100)
        lrn.run(100)
        lrn.learn_collect(0)
        lrn.learn_train(fe.FE_NET_DENSITY, steps=2000, batch=4096, lr=2e-3, seed=1)
        lrn.learn_train(fe.FE_NET_DIVERGENCE, steps=2000, batch=4096, lr=2e-3, seed=2)
        ro = conv_opts()
        ro.residual = fe.FE_RESIDUAL_TWO_SIDED
        lo = conv_opts(fe.FE_WARM_START_LEARNED)
        lo.residual = fe.FE_RESIDUAL_TWO_SIDED
        ref.set_options(ro)
        lrn.set_options(lo)
        ref.reset_iteration_totals()
        lrn.reset_iteration_totals()
        ref.run(400)
        lrn.run(400)
        rd, rv, _ = ref.iteration_totals()
        ld, lv, _ = lrn.iteration_totals()
        gain = (rd + rv) / max(ld + lv, 1)
        print(f"1M, frames 100-499, two-sided residual: reference {rd} + {rv} sweeps, learned {ld} + {lv}, "
              f"{gain:.2f}x fewer")
        assert gain >= 2.0
        for sim in (ref, lrn):
            sim.solve_report_enable(True)
            sim.run(50)
        for kind, tol in ((0, 1e-3), (1, 1e-2)):
            r, l = ref.solve_report(kind), lrn.solve_report(kind)
            print(f"exit residuals, solve {kind}: reference {r}, learned {l}")
            assert r["solves"] == l["solves"] == 50
            # the pressured particles' expansion part is within the tolerance for both starts
            assert l["mean_expansion_pressured"] <= tol and r["mean_expansion_pressured"] <= tol
            # the learned start leaves no larger two-sided residual than the reference start
            assert l["mean_abs"] <= 1.05 * r["mean_abs"]
        err_l = lrn.read_stats()[1] / w.n
        err_r = ref.read_stats()[1] / w.n
        assert err_l <= 1.5 * err_r + 5.0
        assert np.isfinite(lrn.download("velocity_a")).all()
    finally:
        ref.close()
        lrn.close()
