"""The consistent divergence order (FE_DIVERGENCE_CONSISTENT, no reference counterpart) in the
oracle, pinned on the CPU: where the post-integrate sort leaves the order unchanged (a scene at
rest, no particle changes cell) it equals the reference order bit for bit; where the order
changes, the divergence correction lands on the right particles, so the compression rate
recomputed from the corrected velocities falls well below its value before the solve (the
source is linear in the velocities) and below what the reference order leaves.
"""
import numpy as np
import pytest

import fluid_engine as fe
from fluid_engine import scene

F = np.float32


def opts(order, **kw):
    o = fe.default_options()
    o.sort_algorithm = fe.FE_SORT_RADIX
    o.divergence_order = order
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def test_identity_order_equals_reference(oracle_mod):
    """The rest-density lattice shifted by a quarter cell in y: its layers at y = 0.52 + 0.04 k
    would sit on cell faces (h = 0.08) and the first frame's fall would move them across; shifted,
    no particle changes cell in a frame and the post-integrate sort keeps the order."""
    w = scene.rest_density_scene()
    pos = w.positions.copy()
    pos[:, 1] += F(0.02)
    w = scene.Workload(w.name, pos, w.params)
    a = oracle_mod.Oracle(w.params, w.positions, options=opts(fe.FE_DIVERGENCE_REFERENCE))
    b = oracle_mod.Oracle(w.params, w.positions, options=opts(fe.FE_DIVERGENCE_CONSISTENT))
    try:
        a.step()
        b.step()
        # the post-integrate sort kept the order: its permutation is the identity
        entries = b.download("grid_entries")
        assert np.array_equal(entries[: w.n, 1], np.arange(w.n, dtype=np.uint32))
        for name in ("position_a", "velocity_a", "colors", "densities", "pressures", "source_terms",
                     "pressure_accelerations"):
            assert np.array_equal(a.download(name).view(np.uint32), b.download(name).view(np.uint32)), name
        # the consistent order also carries the correction into velocity_b
        assert np.array_equal(b.download("velocity_b").view(np.uint32), b.download("velocity_a").view(np.uint32))
    finally:
        a.close()
        b.close()


def compression(src):
    """Mean compression rate max(0, D rho / Dt) = max(0, -s): what the divergence solve, whose
    pressures are clamped at zero, drives to zero."""
    return float(np.mean(np.maximum(0.0, -np.asarray(src, np.float64))))


def divergence_residuals(oracle_mod, w, state, order):
    """Run the divergence phase of one substep by stages from `state` (position_a, velocity_a of a
    moving fluid) and return (compression before the solve, compression recomputed after the
    correction, fraction of particles whose sorted index moved)."""
    o = oracle_mod.Oracle(w.params, w.positions, options=opts(order))
    try:
        o.upload("position_a", state[0])
        o.upload("velocity_a", state[1])
        o.neighbor_search()
        o.density_alpha()
        o.viscosity()
        o.density_source()
        o.solve_density()
        o.pressure_integration()
        o.neighbor_search()
        moved = float(np.mean(o.download("grid_entries")[: w.n, 1] != np.arange(w.n)))
        o.density_alpha()
        o.divergence_source()
        before = compression(o.download("source_terms"))
        o.solve_divergence()
        o.divergence_integration()
        # measure in the consistent order, from the corrected velocities
        m = oracle_mod.Oracle(w.params, w.positions, options=opts(fe.FE_DIVERGENCE_CONSISTENT))
        try:
            m.upload("position_a", o.download("position_a"))
            m.upload("velocity_a", o.download("velocity_a"))
            m.neighbor_search()
            m.density_alpha()
            m.divergence_source()
            after = compression(m.download("source_terms"))
        finally:
            m.close()
        return before, after, moved
    finally:
        o.close()


@pytest.fixture(scope="module")
def moving_state(oracle_mod):
    """The 32k dam break after 40 frames (the fluid is moving and particles change cells)."""
    w = scene.dam_break_32k(seed=3)
    o = oracle_mod.Oracle(w.params, w.positions, options=opts(fe.FE_DIVERGENCE_REFERENCE))
    try:
        for _ in range(40):
            o.step()
        return w, (o.download("position_a"), o.download("velocity_a"))
    finally:
        o.close()


def test_consistent_correction_reduces_divergence(oracle_mod, moving_state):
    w, state = moving_state
    b_ref, a_ref, moved = divergence_residuals(oracle_mod, w, state, fe.FE_DIVERGENCE_REFERENCE)
    b_con, a_con, moved_c = divergence_residuals(oracle_mod, w, state, fe.FE_DIVERGENCE_CONSISTENT)
    print(f"moved {moved:.3f}: reference {b_ref:.4g} -> {a_ref:.4g}, consistent {b_con:.4g} -> {a_con:.4g}")
    assert moved == moved_c and moved > 0.01
    # the four Jacobi sweeps remove most of the compression once the correction hits the right particles
    assert a_con < 0.5 * b_con
    assert a_con < a_ref
