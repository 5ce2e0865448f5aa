"""Closed-form pins of the oracle's velocity-dependent terms (no GPU).

Each input is chosen so that a plausible mistake in the restatement changes the answer by
far more than FP32 rounding: a flipped sign of v_j - v_i or of a * dt, the neighbour's
density m / rho_j (viscosity.comp:55) replaced by the particle's own, a transposed
component in a dot product.

Lattice cases use the 12^3 zero-jitter block with h = 2 s: an interior particle has 6
neighbours at s, 12 at sqrt(2) s, 8 at sqrt(3) s and 6 at 2 s (q = 1, W = grad W = 0).
"""
import math

import numpy as np
import pytest

from fluid_engine import scene
from fluid_engine._ffi import FeOptions

F = np.float32

SHELLS = [(6, 0.5), (12, math.sqrt(2) / 2), (8, math.sqrt(3) / 2)]


def sigma(h):
    return 8.0 / (math.pi * h ** 3)


def w64(q, h):
    """Cubic spline of common.glsl:38-54 in float64."""
    s = sigma(h)
    if q >= 1.0:
        return 0.0
    return s * (6 * q ** 3 - 6 * q ** 2 + 1) if q <= 0.5 else 2 * s * (1 - q) ** 3


def gf64(q, h):
    """Scalar factor of common.glsl:56-73: grad W = gf / (h r) * r_vec."""
    s = sigma(h)
    if q >= 1.0:
        return 0.0
    return s * (18 * q * q - 12 * q) if q <= 0.5 else -6 * s * (1 - q) ** 2


def opts(sort=1):
    o = FeOptions()
    o.sort_algorithm = sort
    o.solver_mode = 0
    o.max_iterations = 100
    o.density_tol, o.divergence_tol = 1e-3, 1e-2
    return o


def with_params(w, **kw):
    p = w.params
    vals = dict(radius=p.particle_radius, mass=p.particle_mass, h=p.smoothing_radius, rho0=p.target_density,
                nu=p.viscosity, omega=p.relax_factor, dt=p.dt)
    vals.update(kw)
    return scene.make_params(vals["radius"], vals["mass"], vals["h"], vals["rho0"], vals["nu"], vals["omega"],
                             vals["dt"], 1, 1, (0.0, -9.81, 0.0), (-10.0, -10.0, -10.0), (10.0, 10.0, 10.0))


@pytest.fixture(scope="module")
def lattice(oracle_mod):
    """The 12^3 lattice after neighbour search and density: (workload, oracle, sorted
    positions, the sorted index of the interior particle (6, 6, 6), its neighbours)."""
    w = scene.dam_break((12, 12, 12), (1.0, 1.0, 1.0), jitter_amp=0.0)
    prm = with_params(w, nu=1.0, dt=2e-3)
    o = oracle_mod.Oracle(prm, w.positions, options=opts())
    o.neighbor_search()
    o.density_alpha()
    e = o.download("grid_entries")
    pb = o.download("position_b")[:, :3].astype(np.float64)
    k = int(np.nonzero(e[: w.n, 1] == (6 * 12 + 6) * 12 + 6)[0][0])
    d = np.linalg.norm(pb - pb[k], axis=1)
    nb = np.nonzero((d < 0.0401) & (d > 0))[0]
    return w, prm, o, pb, k, nb


def lattice_rho(s):
    h = 2 * s
    m = 1000.0 * s ** 3
    return m * (w64(0.0, h) + sum(c * w64(q, h) for c, q in SHELLS))


def test_lattice_neighbour_densities_are_interior(lattice):
    """Precondition of the viscosity closed form: every neighbour of the chosen particle
    has the interior density (no neighbour sees a duplicated bucket or the block's edge)."""
    w, prm, o, pb, k, nb = lattice
    assert len(nb) == 32      # 6 + 12 + 8 + 6
    rho = o.download("densities", as_double=True)
    want = lattice_rho(float(F(0.02)))
    assert np.abs(rho[nb] - want).max() / want < 5e-6
    assert abs(rho[k] - want) / want < 5e-6


def test_viscosity_quadratic_field_closed_form(lattice):
    """velocity_b = (A (x - x_k)^2, 0, 0): v_j - v_i = A d_x^2 >= 0 for every neighbour, so
    the XSPH sum is A (m / rho) s^2 (2 W(s) + 8 W(sqrt2 s) + 8 W(sqrt3 s)) > 0. A flipped
    v_j - v_i makes it negative; dropping the kernel weight or the m / rho factor moves it by
    orders of magnitude."""
    w, prm, o, pb, k, nb = lattice
    A = 50.0
    v = np.zeros((w.n, 4), F)
    v[:, 0] = (A * (pb[:, 0] - pb[k, 0]) ** 2).astype(F)
    o.upload("velocity_b", v)
    o.viscosity()
    va = o.download("velocity_a").astype(np.float64)
    s = float(F(0.02))
    h = 2 * s
    m = float(F(prm.particle_mass))
    rho = lattice_rho(s)
    shells = 2 * w64(0.5, h) + 8 * w64(math.sqrt(2) / 2, h) + 8 * w64(math.sqrt(3) / 2, h)
    want = float(F(prm.viscosity)) * A * (m / rho) * s * s * shells
    assert want > 0
    assert abs(va[k, 0] - want) <= 2e-5 * want, (va[k, 0], want)
    # no cross-talk into y and z (gravity only)
    assert va[k, 1] == F(F(-9.81) * F(prm.dt))
    assert va[k, 2] == 0


def test_source_terms_linear_field_closed_form(lattice):
    """v = C (x - x_k) with a general matrix C. Then (v_i - v_j) . grad W_ij =
    (r^T C r) gf(q) / (h r); over the cubic shells only trace(C) survives, so
    div_i = m tr(C)/3 sum_shells n gf(q) q (negative for tr(C) > 0: expansion).
    density_source_term.comp:57-58,69: s = (rho0 - rho_i)/dt - div_i;
    divergence_source_term.comp:53-54,61: s = -div_i.
    A flipped v_i - v_j or grad W sign flips div_i; a transposed component picks up C's
    off-diagonal entries."""
    w, prm, o, pb, k, nb = lattice
    C = np.array([[1.0, 0.7, 0.0], [0.3, -0.4, 0.5], [0.2, 0.0, 0.9]])
    v = np.zeros((w.n, 4), F)
    v[:, :3] = ((pb - pb[k]) @ C.T).astype(F)
    s = float(F(0.02))
    h = 2 * s
    m = float(F(prm.particle_mass))
    div = m * np.trace(C) / 3.0 * sum(c * gf64(q, h) * q for c, q in SHELLS)
    assert div < -1000.0
    rho = o.download("densities", as_double=True)
    dt = float(F(prm.dt))

    o.upload("velocity_a", v)
    o.density_source()
    src = o.download("source_terms").astype(np.float64)
    want = (1000.0 - rho[k]) / dt - div
    assert abs(src[k] - want) <= 2e-5 * abs(div), (src[k], want)
    assert (o.download("pressures") == 0).all()

    o.upload("velocity_a", v)
    o.divergence_source()
    src = o.download("source_terms").astype(np.float64)
    assert abs(src[k] - (-div)) <= 2e-5 * abs(div), (src[k], -div)


def cluster_with_padding(oracle_mod, cluster, h, log2t=10, tries=200):
    """`cluster` (float32 [c, 3]) shifted so that the 27 neighbour cells of every cluster
    particle hash to distinct slots of a 2^log2t table, padded with isolated particles (more
    than h apart) to fill the table: each pair of the cluster is then visited exactly once."""
    T = 1 << log2t
    c = cluster.shape[0]
    g = int(math.ceil((T - c) ** (1.0 / 3.0)))
    ii = np.stack(np.meshgrid(np.arange(g), np.arange(g), np.arange(g), indexing="ij"), -1).reshape(-1, 3)
    far = (F(5.0) + ii[: T - c].astype(F) * F(0.25)).astype(F)
    hf = F(h)
    for t in range(tries):
        pos = (cluster + F(0.013) * F(t)).astype(F)
        cells = np.floor(pos * (F(1.0) / hf)).astype(np.int64)
        ok = True
        for cc in cells:
            slots = {oracle_mod.cell_hash(int(cc[0] + dx), int(cc[1] + dy), int(cc[2] + dz), T)
                     for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)}
            if len(slots) != 27:
                ok = False
                break
        if ok:
            return np.concatenate([pos, far]), t
    raise AssertionError("no duplicate-free placement found")


def test_viscosity_uses_neighbour_density(oracle_mod):
    """Three particles on a line at 0, a, a + b (all within h): their densities differ by
    several per cent, so m / rho_j (viscosity.comp:55) and m / rho_i give different sums.
    Closed form from the float32 positions, evaluated in float64."""
    h = 0.04
    a, b = 0.011, 0.021
    cl = np.array([[0.0, 0.0, 0.0], [a, 0.0, 0.0], [a + b, 0.0, 0.0]], F) + F(0.5)
    pos, _ = cluster_with_padding(oracle_mod, cl, h)
    n = pos.shape[0]
    m = 0.008
    prm = scene.make_params(0.01, m, h, 1000.0, 0.8, 0.5, 1e-3, 1, 1, (0.0, -9.81, 0.0),
                            (-10.0, -10.0, -10.0), (10.0, 10.0, 10.0))
    o = oracle_mod.Oracle(prm, pos, options=opts())
    o.neighbor_search()
    o.density_alpha()
    e = o.download("grid_entries")
    order = e[:n, 1]
    where = {int(order[i]): i for i in range(n)}
    x = pos[:3].astype(np.float64)
    mf, hf = float(F(m)), float(F(h))
    dist = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    rho = np.array([mf * sum(w64(dist[i, j] / hf, hf) for j in range(3)) for i in range(3)])
    assert rho.max() / rho.min() > 1.05
    got_rho = o.download("densities", as_double=True)
    for i in range(3):
        assert abs(got_rho[where[i]] - rho[i]) / rho[i] < 2e-6
    u = np.array([[0.3, -0.2, 0.1], [-0.5, 0.4, 0.25], [0.9, 0.05, -0.6]])
    v = np.zeros((n, 4), F)
    for i in range(3):
        v[where[i], :3] = u[i]
    o.upload("velocity_b", v)
    o.viscosity()
    va = o.download("velocity_a").astype(np.float64)
    nu, dt = float(F(0.8)), float(F(1e-3))
    g = np.array([0.0, float(F(-9.81)), 0.0])
    for i in range(3):
        u32 = u.astype(F).astype(np.float64)
        visc = sum((mf / rho[j]) * (u32[j] - u32[i]) * w64(dist[i, j] / hf, hf) for j in range(3) if j != i)
        own = sum((mf / rho[i]) * (u32[j] - u32[i]) * w64(dist[i, j] / hf, hf) for j in range(3) if j != i)
        want = u32[i] + nu * visc + g * dt
        wrong = u32[i] + nu * own + g * dt
        got = va[where[i], :3]
        tol = 1e-6 + 1e-5 * np.abs(nu * visc)
        assert (np.abs(got - want) <= tol).all(), (i, got, want)
        # the own-density variant is distinguishable at this tolerance
        assert (np.abs(wrong - want) > 5 * tol).any()


def test_pressure_integration_nonzero_acceleration_closed_form(oracle_mod):
    """pressure_integration.comp:29-30: v' = v + a dt, x' = x_b + v' dt, away from the walls.
    With |a dt| comparable to |v|, a flipped a * dt or an x' from the old velocity moves the
    result by more than 0.1 m/s."""
    pos = np.array([[0.5, 0.5, 0.5], [0.3, 0.6, 0.4], [0.7, 0.2, 0.55]], F)
    dt = 0.01
    prm = scene.make_params(0.01, 0.008, 0.04, 1000.0, 0.01, 0.5, dt, 1, 1, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    o = oracle_mod.Oracle(prm, pos)
    o.upload("position_b", np.concatenate([pos, np.ones((3, 1), F)], axis=1))
    v = np.array([[0.2, 0.1, -0.3, 0], [-0.4, 0.0, 0.25, 0], [0.05, 0.3, 0.0, 0]], F)
    a = np.array([[30.0, -50.0, 70.0, 0], [-12.0, 25.0, 40.0, 0], [0.0, -80.0, 15.0, 0]], F)
    o.upload("velocity_a", v)
    o.upload("pressure_accelerations", a)
    o.pressure_integration()
    x = o.download("position_a").astype(np.float64)
    u = o.download("velocity_a").astype(np.float64)
    dtf = float(F(dt))
    want_v = v[:, :3].astype(np.float64) + a[:, :3].astype(np.float64) * dtf
    want_x = pos.astype(np.float64) + want_v * dtf
    assert np.abs(u[:, :3] - want_v).max() < 1e-6
    assert np.abs(x[:, :3] - want_x).max() < 1e-7
    assert (x[:, 3] == 1).all() and (u[:, 3] == 0).all()


def test_divergence_integration_nonzero_acceleration_closed_form(oracle_mod):
    """divergence_integration.comp: velocity_a += a dt, with the colour ramp read at
    |v'| / 4 (a flipped a * dt changes both the velocity and the colour band)."""
    pos = np.zeros((2, 3), F)
    dt = 0.02
    prm = scene.make_params(0.01, 0.008, 0.04, 1000.0, 0.01, 0.5, dt, 1, 1, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    o = oracle_mod.Oracle(prm, pos)
    v = np.array([[1.0, 0.0, 0.0, 0], [0.0, 0.5, -0.5, 0]], F)
    a = np.array([[50.0, 0.0, 0.0, 0], [10.0, -25.0, 25.0, 0]], F)
    o.upload("velocity_a", v)
    o.upload("pressure_accelerations", a)
    o.divergence_integration()
    u = o.download("velocity_a").astype(np.float64)
    c = o.download("colors").astype(np.float64)
    want = v[:, :3].astype(np.float64) + a[:, :3].astype(np.float64) * float(F(dt))
    assert np.abs(u[:, :3] - want).max() < 1e-6
    # particle 0: |v'| = 2 -> t = 0.5, exactly the water colour
    assert np.allclose(c[0], [0.2, 0.5, 1.0, 1.0], atol=1e-6)
    # particle 1: v' = (0.2, 0, 0): t = 0.05, 10 % of the way from deep to water
    t2 = 2 * 0.05
    assert np.allclose(c[1, :3], [0.05 + t2 * 0.15, 0.1 + t2 * 0.4, 0.5 + t2 * 0.5], atol=1e-6)
