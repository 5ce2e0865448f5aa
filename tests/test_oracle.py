"""Pins of the CPU oracle (no GPU): the reference's golden capture, closed forms of the
SPH kernel on a lattice, the reference's own sort tests, invariants of every stage.

Each test fixes something a plausible mistake in the restatement would break (a dropped
neighbour shell, a wrong sign, a transposed index, an unstable sort, the cell rounding).
"""
import math

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from fluid_engine import scene
from fluid_engine._ffi import FeOptions

F = np.float32
SENT = 0xFFFFFFFF


def opts(sort=1, mode=0, tol=(1e-3, 1e-2), maxit=100):
    o = FeOptions()
    o.sort_algorithm = sort
    o.solver_mode = mode
    o.density_error_pct = 0.1
    o.divergence_error_pct = 0.5
    o.density_tol, o.divergence_tol = tol
    o.max_iterations = maxit
    return o


# ---------------------------------------------------------------------------------------
# Golden capture: scripts/rest_density.csv (rest_density_diagnostic,
# src/renderer/pipelines/convergence_benchmark.rs:420-542), 4800 densities in the order of
# the bitonic-sorted buffer.

def test_rest_density_golden_in_place(oracle_mod):
    w = scene.rest_density_scene()
    assert w.n == 4800
    ref = np.loadtxt("tests/golden/rest_density.csv", skiprows=1, dtype=np.float64)
    o = oracle_mod.Oracle(w.params, w.positions, options=opts(sort=0))
    o.neighbor_search()
    o.density_alpha()
    d = o.download("densities").astype(np.float64)
    # every row, in place: this fixes the bitonic network, the hash and the cell rounding
    rel = np.abs(d - ref) / np.abs(ref)
    assert rel.max() < 1e-5, rel.max()
    assert np.abs(d - ref).max() < 1e-3


def test_rest_density_ieee_cell_rounding_would_fail(oracle_mod):
    """The cell coordinate floor(pos * (1/h)) is what reproduces the capture's bucket order;
    floor(pos / h) (IEEE quotient) assigns 240 face-aligned lattice particles elsewhere."""
    w = scene.rest_density_scene()
    h = F(w.params.smoothing_radius)
    a = np.floor(w.positions * (F(1.0) / h))
    b = np.floor(w.positions / h)
    assert (a != b).any(axis=1).sum() == 240


def test_rest_density_statistics(oracle_mod):
    """Mean 945.91, max 999.9745 and mean |rho - rho0| 54.09 of the capture."""
    w = scene.rest_density_scene()
    o = oracle_mod.Oracle(w.params, w.positions, options=opts(sort=0))
    o.neighbor_search()
    o.density_alpha()
    d = o.download("densities").astype(np.float64)
    assert abs(d.mean() - 945.91) < 0.01
    assert abs(d.max() - 999.9745) < 1e-3
    assert abs(np.abs(d - 1000.0).mean() - 54.09) < 0.01


# ---------------------------------------------------------------------------------------
# Lattice closed forms (h = 2 s): interior neighbours at |d| = s (6), sqrt2 s (12),
# sqrt3 s (8); the 6 at 2s sit at q = 1 and contribute W = 0.

def lattice_closed_forms(rho0, s):
    h = 2.0 * s
    m = rho0 * s ** 3
    sig = 8.0 / (math.pi * h ** 3)

    def w(q):
        return sig * (6 * q ** 3 - 6 * q ** 2 + 1) if q <= 0.5 else 2 * sig * (1 - q) ** 3

    def gf(q):
        return sig * (18 * q * q - 12 * q) if q <= 0.5 else -6 * sig * (1 - q) ** 2

    shells = [(6, 0.5), (12, math.sqrt(2) / 2), (8, math.sqrt(3) / 2)]
    rho = m * (w(0.0) + sum(c * w(q) for c, q in shells))
    # |m grad W|^2 = (m gf / h)^2 for every neighbour; sum of m grad W is 0 by symmetry
    alpha = sum(c * (m * gf(q) / h) ** 2 for c, q in shells)
    return rho, alpha


def test_lattice_density_closed_form_value():
    rho, _ = lattice_closed_forms(1000.0, 0.02)
    assert abs(rho - 999.9725) < 1e-4


@pytest.mark.parametrize("fp64", [False, True])
def test_lattice_interior_density_and_alpha(oracle_mod, fp64):
    w = scene.dam_break((12, 12, 12), (1.0, 1.0, 1.0), jitter_amp=0.0)
    o = oracle_mod.Oracle(w.params, w.positions, fp64=fp64)
    o.neighbor_search()
    o.density_alpha()
    e = o.download("grid_entries")
    d = o.download("densities", as_double=True)
    a = o.download("factors", as_double=True)
    # particle (6, 6, 6) of the block is interior
    orig = (6 * 12 + 6) * 12 + 6
    k = int(np.nonzero(e[: w.n, 1] == orig)[0][0])
    rho, alpha = lattice_closed_forms(1000.0, float(F(0.02)))
    # positions are FP32 lattice points (relative rounding ~1e-7 of r), which bounds how
    # close even the FP64 evaluation can get; the smallest shell is 1.2 % of rho
    assert abs(d[k] - rho) / rho < 5e-6
    assert abs(a[k] - alpha) / alpha < 5e-5


def test_fp32_vs_fp64_oracle_agree(oracle_mod):
    w = scene.rest_density_scene()
    o32 = oracle_mod.Oracle(w.params, w.positions)
    o64 = oracle_mod.Oracle(w.params, w.positions, fp64=True)
    for o in (o32, o64):
        o.neighbor_search()
        o.density_alpha()
    d32 = o32.download("densities").astype(np.float64)
    d64 = o64.download("densities", as_double=True)
    assert np.abs(d32 - d64).max() / 1000.0 < 2e-6


# ---------------------------------------------------------------------------------------
# Hash (common.glsl:30-35) and kernels (common.glsl:38-73).

def test_cell_hash_closed_form(oracle_mod):
    T = 1 << 15
    for c in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (3, -7, 11), (-100, -200, -300)]:
        u = [x & 0xFFFFFFFF for x in c]
        want = (((u[0] * 73856093) & 0xFFFFFFFF) ^ ((u[1] * 19349663) & 0xFFFFFFFF)
                ^ ((u[2] * 83492791) & 0xFFFFFFFF)) % T
        assert oracle_mod.cell_hash(*c, T) == want
    assert oracle_mod.cell_hash(1, 0, 0, T) == 73856093 % T
    assert oracle_mod.cell_hash(-1, 0, 0, 1 << 20) == ((0xFFFFFFFF * 73856093) & 0xFFFFFFFF) % (1 << 20)


def test_kernel_w_values_and_normalisation(oracle_mod):
    h = 0.04
    sig = 8.0 / (math.pi * h ** 3)
    assert abs(oracle_mod.kernel_w64(0.0, h) - sig) < 1e-9 * sig
    assert abs(oracle_mod.kernel_w64(0.5 * h, h) - 0.25 * sig) < 1e-9 * sig
    assert oracle_mod.kernel_w64(h, h) == 0.0
    # continuity at q = 1/2: both branches give sigma / 4
    assert abs(oracle_mod.kernel_w64(0.5 * h * (1 - 1e-12), h) - oracle_mod.kernel_w64(0.5 * h * (1 + 1e-12), h)) < 1e-6 * sig
    # integral of W over R^3 is 1
    r = (np.arange(20000) + 0.5) * (h / 20000)
    integral = sum(4 * math.pi * x * x * oracle_mod.kernel_w64(float(x), h) for x in r) * (h / 20000)
    assert abs(integral - 1.0) < 1e-6


def test_kernel_grad_is_derivative_of_w(oracle_mod):
    h = 0.04
    for q in (0.1, 0.3, 0.49, 0.6, 0.8, 0.95):
        r = q * h
        g = oracle_mod.kernel_grad([r, 0.0, 0.0], r, h)
        eps = 1e-4 * h
        dw = (oracle_mod.kernel_w64(r + eps, h) - oracle_mod.kernel_w64(r - eps, h)) / (2 * eps)
        assert abs(g[0] - dw) < 2e-3 * abs(dw) + 1e-3
        assert g[1] == 0 and g[2] == 0
    # points from j to i: for r_vec = (r, 0, 0) the gradient is negative (W decreases)
    assert oracle_mod.kernel_grad([0.01, 0, 0], 0.01, h)[0] < 0
    assert list(oracle_mod.kernel_grad([0, 0, 0], 0.0, h)) == [0, 0, 0]


# ---------------------------------------------------------------------------------------
# Sort: the reference's CPU sort tests (src/renderer/pipelines/sorter.rs:585-646), on the
# literal radix restatement and on the bitonic network, plus the stable-permutation closed
# form of the radix result.

def padded(hashes):
    n = len(hashes)
    t = 1
    while t < n:
        t <<= 1
    p = np.full((t, 2), SENT, dtype=np.uint32)
    p[:n, 0] = hashes
    p[:n, 1] = np.arange(n)
    return p


SORTS = ["radix_sort_pairs", "bitonic_sort_pairs"]


@pytest.mark.parametrize("fn", SORTS)
def test_sort_small_known(oracle_mod, fn):
    out = getattr(oracle_mod, fn)(padded([3, 1, 4, 1, 5, 9, 2, 6]))
    assert (np.diff(out[:, 0].astype(np.int64)) >= 0).all()


@pytest.mark.parametrize("fn", SORTS)
def test_sort_matches_stdlib(oracle_mod, fn):
    i = np.arange(1000, dtype=np.uint64)
    h = ((i * np.uint64(2654435761) + (i >> np.uint64(3))) & np.uint64(0xFFFFFFFF)) % np.uint64(131072)
    p = padded(h.astype(np.uint32))
    out = getattr(oracle_mod, fn)(p)
    assert (out[:, 0] == np.sort(p[:, 0])).all()


@pytest.mark.parametrize("fn", SORTS)
def test_sort_all_same_hash(oracle_mod, fn):
    out = getattr(oracle_mod, fn)(padded([42] * 300))
    assert (np.diff(out[:, 0].astype(np.int64)) >= 0).all()


@pytest.mark.parametrize("fn", SORTS)
def test_sort_reverse_order(oracle_mod, fn):
    out = getattr(oracle_mod, fn)(padded(list(range(511, -1, -1))))
    assert (np.diff(out[:, 0].astype(np.int64)) >= 0).all()


@pytest.mark.parametrize("fn", SORTS)
def test_sentinels_land_at_end(oracle_mod, fn):
    out = getattr(oracle_mod, fn)(padded(list(range(100))))
    assert int(np.argmax(out[:, 0] == SENT)) == 100
    assert (out[100:, 0] == SENT).all()


def test_radix_all_same_hash_keeps_index_order(oracle_mod):
    out = oracle_mod.radix_sort_pairs(padded([42] * 300))
    assert (out[:300, 1] == np.arange(300)).all()


@settings(max_examples=40, deadline=None)
@given(st.integers(1, 12), st.integers(1, 1 << 32), st.integers(0, 2 ** 31))
def test_radix_is_stable_sort(oracle_mod, log2n, keyspace, seed):
    rng = np.random.default_rng(seed)
    n = 1 << log2n
    keys = rng.integers(0, keyspace, size=n, dtype=np.uint64).astype(np.uint32)
    p = np.stack([keys, np.arange(n, dtype=np.uint32)], axis=1)
    out = oracle_mod.radix_sort_pairs(p)
    perm = np.argsort(keys, kind="stable")
    assert (out[:, 1] == perm).all()
    assert (out[:, 0] == keys[perm]).all()


@settings(max_examples=30, deadline=None)
@given(st.integers(1, 11), st.integers(1, 64), st.integers(0, 2 ** 31))
def test_bitonic_sorts_and_permutes(oracle_mod, log2n, keyspace, seed):
    rng = np.random.default_rng(seed)
    n = 1 << log2n
    keys = rng.integers(0, keyspace, size=n).astype(np.uint32)
    p = np.stack([keys, np.arange(n, dtype=np.uint32)], axis=1)
    out = oracle_mod.bitonic_sort_pairs(p)
    assert (np.diff(out[:, 0].astype(np.int64)) >= 0).all()
    assert (np.sort(out[:, 1]) == np.arange(n)).all()
    assert (keys[out[:, 1]] == out[:, 0]).all()


# ---------------------------------------------------------------------------------------
# Neighbour-search stages and solver invariants on a jittered block.

@pytest.fixture(scope="module")
def small_block():
    return scene.dam_break((10, 9, 8), (1.0, 1.0, 1.0), seed=3)


def test_offsets_and_reorder_closed_form(oracle_mod, small_block):
    w = small_block
    o = oracle_mod.Oracle(w.params, w.positions)
    o.neighbor_search()
    e = o.download("grid_entries")
    gs = o.download("grid_start")
    keys = e[:, 0]
    # hashes are those of the cells of position_a
    h = F(w.params.smoothing_radius)
    c = np.floor(w.positions * (F(1.0) / h)).astype(np.int64).astype(np.uint32)
    with np.errstate(over="ignore"):
        hs = ((c[:, 0] * np.uint32(73856093)) ^ (c[:, 1] * np.uint32(19349663))
              ^ (c[:, 2] * np.uint32(83492791))) % np.uint32(o.table_size)
    assert (keys[: w.n] == hs[e[: w.n, 1]]).all()
    assert (np.diff(keys.astype(np.int64)) >= 0).all()
    for k in range(o.table_size):
        idx = np.nonzero(keys == k)[0]
        assert gs[k] == (idx[0] if len(idx) else SENT)
    pb = o.download("position_b")
    assert (pb[:, :3] == w.positions[e[: w.n, 1]]).all()


def test_uniform_velocity_invariants(oracle_mod, small_block):
    w = small_block
    o = oracle_mod.Oracle(w.params, w.positions)
    o.neighbor_search()
    o.density_alpha()
    v = np.zeros((w.n, 4), F)
    v[:, :3] = [0.3, -1.25, 0.75]
    o.upload("velocity_b", v)
    o.viscosity()
    va = o.download("velocity_a")
    p = w.params
    want = v[:, :3] + np.array(p.gravity[:3], F) * F(p.dt)
    assert (va[:, :3] == want).all()          # XSPH of a uniform field is zero
    assert (va[:, 3] == 0).all()
    o.upload("velocity_a", v)
    o.density_source()
    s = o.download("source_terms")
    rho = o.download("densities")
    assert (s == (F(p.target_density) - rho) / F(p.dt)).all()   # rigid translation: no divergence
    assert (o.download("pressures") == 0).all()


def _hashes(c, T):
    c = np.asarray(c, dtype=np.int64).astype(np.uint32)
    with np.errstate(over="ignore"):
        return ((c[..., 0] * np.uint32(73856093)) ^ (c[..., 1] * np.uint32(19349663))
                ^ (c[..., 2] * np.uint32(83492791))) % np.uint32(T)


def momentum_scene():
    """An 8^3 jittered cluster plus 65024 particles parked on one far point, so that the
    table has 2^16 slots and the cluster's 6^3 neighbour cells all hash to distinct slots:
    no bucket is visited twice and pair forces must cancel."""
    T = 1 << 16
    for shift in range(50):
        w = scene.dam_break((8, 8, 8), (1.0, 1.0, 1.0), seed=11)
        pos = w.positions + F(0.04) * F(shift)
        h = F(w.params.smoothing_radius)
        c = np.floor(pos * (F(1.0) / h)).astype(np.int64)
        lo = c.min(axis=0) - 1
        hi = c.max(axis=0) + 1
        cells = np.stack(np.meshgrid(*[np.arange(a, b + 1) for a, b in zip(lo, hi)], indexing="ij"), -1).reshape(-1, 3)
        far = np.full((T - pos.shape[0], 3), 50.0, F)
        far_slot = _hashes(np.floor(far[:1] * (F(1.0) / h)), T)
        hs = _hashes(cells, T)
        if len(np.unique(hs)) == len(hs) and far_slot[0] not in hs:
            return w.params, np.concatenate([pos, far]), pos.shape[0]
    raise AssertionError("no collision-free placement found")


def test_pressure_force_zero_pressure_and_momentum(oracle_mod):
    params, pos, k = momentum_scene()
    n = pos.shape[0]
    o = oracle_mod.Oracle(params, pos)
    o.neighbor_search()
    o.density_alpha()
    o.upload("pressures", np.zeros(n, F))
    o.pressure_force()
    assert (o.download("pressure_accelerations") == 0).all()
    rng = np.random.default_rng(0)
    o.upload("pressures", rng.uniform(0, 5000, n).astype(F))
    o.pressure_force()
    e = o.download("grid_entries")
    cluster = e[:n, 1] < k
    a = o.download("pressure_accelerations")[:, :3].astype(np.float64)[cluster]
    # equal masses: sum of accelerations of a closed set of pairs is zero up to rounding
    assert np.abs(a.sum(axis=0)).max() < 1e-5 * np.abs(a).sum(axis=0).max()
    # dropping the i-term or the j-term of (p_i/rho_i^2 + p_j/rho_j^2) breaks the balance:
    # with a single pressurised particle the force pushes every neighbour outward
    pb = o.download("position_b")
    sel = np.nonzero(cluster)[0]
    kk = sel[int(np.argmin(np.linalg.norm(pb[sel, :3] - pb[sel, :3].mean(axis=0), axis=1)))]
    p = np.zeros(n, F)
    p[kk] = 1000.0
    o.upload("pressures", p)
    o.pressure_force()
    a = o.download("pressure_accelerations")[:, :3]
    d = pb[:, :3] - pb[kk, :3]
    near = (np.linalg.norm(d, axis=1) < params.smoothing_radius) & (np.arange(n) != kk)
    assert near.sum() > 20
    assert ((a[near] * d[near]).sum(axis=1) > 0).all()
    assert abs(a[near].astype(np.float64).sum(axis=0) + a[kk]).max() < 1e-4 * np.abs(a[kk]).max()


def test_pressure_update_diagonal_is_alpha(oracle_mod, small_block):
    """With p = p0 at one particle, src = 0 and omega = 1, one Jacobi sweep gives
    Ap_i = a_ii p0 exactly in exact arithmetic, hence p_i -> 0: this ties alpha to the
    pressure_force / pressure_update pair."""
    w = small_block
    prm = scene.make_params(w.params.particle_radius, w.params.particle_mass, w.params.smoothing_radius,
                            1000.0, 0.01, 1.0, 1e-3, 1, 1, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    o = oracle_mod.Oracle(prm, w.positions)
    o.neighbor_search()
    o.density_alpha()
    pb = o.download("position_b")
    k = int(np.argmin(np.linalg.norm(pb[:, :3] - pb[:, :3].mean(axis=0), axis=1)))
    p = np.zeros(w.n, F)
    p[k] = 2500.0
    o.upload("pressures", p)
    o.upload("source_terms", np.zeros(w.n, F))
    o.pressure_force()
    o.pressure_update()
    assert abs(float(o.download("pressures")[k])) < 1e-4 * 2500.0


def test_two_sided_residual_counts_pressured_expansion(oracle_mod, small_block):
    """FE_RESIDUAL_TWO_SIDED (no reference counterpart): one particle with p_k = 2500 and
    s = 0 everywhere. Its residual r_k = -Ap_k = -a_kk p_k is an expansion (r_k > 0), which the
    compression-only residual ignores and the two-sided one counts as |r_k| dt / rho0; the
    unpressured particles count the same in both. a_kk = -(dt / rho_k^2) alpha_k."""
    w = small_block
    dt = 1e-3
    prm = scene.make_params(w.params.particle_radius, w.params.particle_mass, w.params.smoothing_radius,
                            1000.0, 0.01, 1.0, dt, 1, 1, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    qs = []
    for resid in (0, 1):
        o = oracle_mod.Oracle(prm, w.positions, options=opts())
        op = opts()
        op.residual = resid
        o.set_options(op)
        o.neighbor_search()
        o.density_alpha()
        pb = o.download("position_b")
        k = int(np.argmin(np.linalg.norm(pb[:, :3] - pb[:, :3].mean(axis=0), axis=1)))
        p = np.zeros(w.n, F)
        p[k] = 2500.0
        o.upload("pressures", p)
        o.upload("source_terms", np.zeros(w.n, F))
        o.pressure_force()
        qs.append(int(o.pressure_update()))
        rho = float(o.download("densities")[k])
        alpha = float(o.download("factors")[k])
    a_kk = -(dt / (rho * rho)) * alpha
    want = -a_kk * 2500.0 * dt / 1000.0 * 2 ** 32
    extra = qs[1] - qs[0]
    assert want > 0
    assert abs(extra - want) <= 1e-4 * want, (extra, want)


def test_integration_collision_closed_form(oracle_mod):
    pos = np.array([[0.5, 0.5, 0.5], [0.5, 0.005, 0.5], [0.999, 0.5, 0.5], [0.5, 0.5, -0.2]], F)
    prm = scene.make_params(0.01, 0.008, 0.04, 1000.0, 0.01, 0.5, 0.5, 1, 1, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    o = oracle_mod.Oracle(prm, pos)
    o.upload("position_b", np.concatenate([pos, np.ones((4, 1), F)], axis=1))
    v = np.array([[0.2, 0.0, 0.0, 0], [0.2, -0.1, 0.4, 0], [0.1, 0.2, 0.3, 0], [0.2, 0.2, 0.0, 0]], F)
    o.upload("velocity_a", v)
    o.upload("pressure_accelerations", np.zeros((4, 4), F))
    o.pressure_integration()
    x = o.download("position_a")
    u = o.download("velocity_a")
    # free particle: x' = x + v dt
    assert np.allclose(x[0, :3], [0.6, 0.5, 0.5]) and np.allclose(u[0, :3], [0.2, 0, 0])
    # floor (y checked first): y -> min + r + 0.001, v_y * -0.5, tangential * 0.95
    assert np.isclose(x[1, 1], 0.011) and np.allclose(u[1, :3], [0.19, 0.05, 0.38])
    # x wall at max: x -> max - r - 0.001
    assert np.isclose(x[2, 0], 0.989) and np.allclose(u[2, :3], [-0.05, 0.19, 0.285])
    # z below min
    assert np.isclose(x[3, 2], 0.011) and np.allclose(u[3, :3], [0.19, 0.19, 0.0])
    assert (x[:, 3] == 1).all() and (u[:, 3] == 0).all()


def test_divergence_integration_colours(oracle_mod):
    pos = np.zeros((3, 3), F)
    prm = scene.make_params(0.01, 0.008, 0.04, 1000.0, 0.01, 0.5, 1.0, 1, 1, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    o = oracle_mod.Oracle(prm, pos)
    o.upload("velocity_a", np.array([[0, 0, 0, 0], [2.0, 0, 0, 0], [0, 5.0, 0, 0]], F))
    o.upload("pressure_accelerations", np.zeros((3, 4), F))
    o.divergence_integration()
    c = o.download("colors")
    assert np.allclose(c[0], [0.05, 0.1, 0.5, 1])     # at rest: deep colour
    assert np.allclose(c[1], [0.2, 0.5, 1.0, 1])      # speed 2 -> t = 0.5: water colour
    assert np.allclose(c[2], [1, 1, 1, 1])            # speed >= 4: foam


def test_stats_integer_closed_form(oracle_mod):
    n = 5
    prm = scene.make_params(0.01, 0.008, 0.04, 1000.0, 0.01, 0.5, 1e-3, 1, 1, (0, -9.81, 0), (0, 0, 0), (1, 1, 1))
    o = oracle_mod.Oracle(prm, np.zeros((n, 3), F))
    v = np.zeros((n, 4), F)
    v[2, :3] = [3.0, 4.0, 0.0]
    o.upload("velocity_a", v)
    o.upload("densities", np.array([1003.7, 990.2, 1000.0, 5e7, 1000.5], F))
    o.upload("source_terms", np.array([12.34, -0.05, 0.0, -1e9, 999.99], F))
    o.stats()
    s = o.read_stats()
    assert s[0] == np.float32(5.0).view(np.uint32)
    assert s[1] == 3 + 9 + 0 + 10000 + 0
    assert s[2] == 123 + 0 + 0 + 10000 + 9999


def test_generator_counts_follow_f32_floor():
    # Q10: f32 floor(w / s) for the workloads' extents
    for extent, want in ((0.64, 32), (2.56, 128), (1.28, 64), (0.4, 20)):
        assert int(np.floor(F(extent) / F(0.02))) == want


def test_residual_quantisation(oracle_mod):
    assert oracle_mod.residual_q(5.0, 1e-3, 1000.0) == 0      # under-dense: no compression error
    q = oracle_mod.residual_q(-1000.0, 1e-3, 1000.0)          # 1e-3 relative error
    assert abs(q / 2 ** 32 - 1e-3) < 1e-9
    assert oracle_mod.residual_q(-1e30, 1e-3, 1000.0) == 256 * 2 ** 32


# ---------------------------------------------------------------------------------------
# Warm-start features (no reference counterpart) on the lattice: closed forms for an interior
# particle with h = 2 s (33 pairs: shells of 1, 6, 12, 8 and the 6 at exactly r = h).

def test_features_lattice_closed_forms(oracle_mod):
    w = scene.dam_break((12, 12, 12), (1.0, 1.0, 1.0), jitter_amp=0.0)
    o = oracle_mod.Oracle(w.params, w.positions)
    o.neighbor_search()
    o.density_alpha()
    v = np.zeros((w.n, 4), F)
    v[:, :3] = [0.5, -2.0, 1.0]             # uniform velocity: no relative motion
    o.upload("velocity_a", v)
    prev = np.full(w.n, 2.5, F)
    prev[0] = 7.0
    f = o.features(prev).astype(np.float64)
    e = o.download("grid_entries")
    rho = o.download("densities", as_double=True)
    m = float(F(w.params.particle_mass))
    orig = (6 * 12 + 6) * 12 + 6            # an interior particle
    k = int(np.nonzero(e[: w.n, 1] == orig)[0][0])
    rho_cf, _ = lattice_closed_forms(1000.0, float(F(0.02)))
    # the 6 neighbours at exactly r = h pass r^2 <= h^2 or not by FP32 rounding of r^2 (they
    # contribute W = 0); each one missing removes |d|^2 / 4 = 1 from sum |r|^2 / h^2 and at
    # most 1 from |sum r| / h
    missing = 33 - f[k, 3]
    assert 0 <= missing <= 6
    assert abs(f[k, 5] * m - rho_cf) / rho_cf < 5e-6          # sum W = rho / m
    assert abs(f[k, 6] - 2.5 * f[k, 5]) <= 1e-6 * f[k, 6]      # p_prev uniform 2.5
    assert f[k, 4] == 2.5
    assert abs(f[k, 7]) <= missing + 1e-5
    assert abs(f[k, 8] - (19.5 - missing)) < 1e-4             # sum |d|^2 / 4 = 78 / 4
    assert f[k, 9] == 0 and f[k, 10] == 0
    assert f[k, 1] == 0 and f[k, 11] == 0                      # source terms are still 0
    assert abs(f[k, 0] - (rho[k] - 1000.0) / 1000.0) < 1e-7
    # the particle whose previous pressure is 7.0 has it as feature 4
    k0 = int(np.nonzero(e[: w.n, 1] == 0)[0][0])
    assert (f[:, 3] >= 8).all()                                # corners: 1 + 3 + 3 + 1 at least
    assert f[k0, 4] == prev[k0]
