"""The density splat of the render hand-off (shaders/compute/splat_density.comp:11-51) in the
oracle, pinned on the CPU against what the mathematics fixes: a lone particle's texture is
the stencil floor(200 (1 - d^2/64)^3) (exact rational arithmetic) on the texels with
d^2 < 64 around its centre texel, clipped to the image, with x fastest; a particle on the far
face (uvw = 1) is centred one texel past the image; particles outside the box or with a NaN
coordinate add nothing; and the texture is additive over particle sets and independent of
their order (unsigned sums modulo 2^32).
"""
import numpy as np
import pytest

from fluid_engine import scene

F = np.float32


def exact_weight(d2):
    """floor(200 (1 - d2/64)^3) for d2 < 64, else 0, in integers."""
    d2 = np.asarray(d2, dtype=np.int64)
    w = (200 * (64 - d2) ** 3) // (64 ** 3)
    return np.where(d2 < 64, w, 0).astype(np.uint32)


def brute(centres, res):
    """Sum of the clipped stencils around integer centres; res = (rx, ry, rz)."""
    rx, ry, rz = res
    z, y, x = np.meshgrid(np.arange(rz), np.arange(ry), np.arange(rx), indexing="ij")
    out = np.zeros((rz, ry, rx), np.uint64)
    for cx, cy, cz in centres:
        d2 = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2
        out += exact_weight(d2)
    return (out % (1 << 32)).astype(np.uint32)


def oracle_splat(oracle_mod, positions, res=(40, 36, 32), box_min=(0.0, 0.0, 0.0), box_max=(4.0, 3.6, 3.2)):
    p = scene.make_params(0.01, 0.001, 0.04, 1000.0, 0.01, 0.5, 1e-3, 4, 4, (0.0, -9.81, 0.0), box_min, box_max,
                          grid_res=res)
    o = oracle_mod.Oracle(p, np.asarray(positions, np.float32).reshape(-1, 3))
    try:
        return o.splat_density()
    finally:
        o.close()


def test_weight_table():
    """The stencil weights, in exact arithmetic, at a few closed forms."""
    assert exact_weight([0, 16, 32, 48, 53, 54, 63, 64]).tolist() == [200, 84, 25, 3, 1, 0, 0, 0]


@pytest.mark.parametrize("centre", [(20, 18, 16), (2, 33, 5), (39, 0, 31)])
def test_lone_particle_is_the_stencil(oracle_mod, centre):
    """Box 4 x 3.6 x 3.2 m over 40 x 36 x 32 texels of 0.1 m: a particle at the middle of
    texel c has its centre texel c; the texture is the clipped stencil around c (axis order
    and clipping are both visible with the off-centre cases)."""
    pos = [((centre[0] + 0.5) * 0.1, (centre[1] + 0.5) * 0.1, (centre[2] + 0.5) * 0.1)]
    got = oracle_splat(oracle_mod, pos)
    assert got.shape == (32, 36, 40)
    want = brute([centre], (40, 36, 32))
    assert np.array_equal(got, want)
    assert got[centre[2], centre[1], centre[0]] == 200


def test_far_face_centres_past_the_image(oracle_mod):
    """uvw = 1 exactly (the particle on box_max) gives centre texel = grid_res, one past the
    last texel; its stencil reaches back into the image."""
    got = oracle_splat(oracle_mod, [(4.0, 3.6, 3.2)])
    want = brute([(40, 36, 32)], (40, 36, 32))
    assert np.array_equal(got, want)
    assert got.sum() > 0 and got[31, 35, 39] == exact_weight(3)


def test_outside_and_nan_add_nothing(oracle_mod):
    pos = [(-0.001, 1.0, 1.0), (1.0, 3.6001, 1.0), (1.0, 1.0, 3.25), (np.nan, 1.0, 1.0), (1.0, np.inf, 1.0)]
    got = oracle_splat(oracle_mod, pos)
    assert not got.any()


def test_additive_and_order_free(oracle_mod):
    rng = np.random.default_rng(11)
    a = rng.uniform([0, 0, 0], [4.0, 3.6, 3.2], (300, 3)).astype(F)
    b = rng.uniform([0, 0, 0], [4.0, 3.6, 3.2], (200, 3)).astype(F)
    ta = oracle_splat(oracle_mod, a)
    tb = oracle_splat(oracle_mod, b)
    both = np.concatenate([a, b])
    tab = oracle_splat(oracle_mod, both)
    assert np.array_equal(tab, ta + tb)
    perm = rng.permutation(len(both))
    assert np.array_equal(oracle_splat(oracle_mod, both[perm]), tab)


def test_random_particles_match_brute_force(oracle_mod):
    """Random positions, centres computed independently (uvw * res truncated, with uvw =
    (p - min) * (1 / size) in float) and summed by brute force."""
    rng = np.random.default_rng(2)
    res = (23, 17, 29)
    lo = np.array([-1.5, 0.25, -0.75], F)
    hi = np.array([0.8, 2.0, 1.0], F)
    pos = rng.uniform(lo, hi, (60, 3)).astype(F)
    inv = (F(1.0) / (hi - lo)).astype(F)
    uvw = ((pos - lo) * inv).astype(F)
    centres = (uvw * np.array(res, F)).astype(np.int32)
    got = oracle_splat(oracle_mod, pos, res=res, box_min=tuple(lo), box_max=tuple(hi))
    assert np.array_equal(got, brute(centres.tolist(), res))


def test_bad_grid_res(oracle_mod):
    p = scene.make_params(0.01, 0.001, 0.04, 1000.0, 0.01, 0.5, 1e-3, 4, 4, (0.0, -9.81, 0.0), (0, 0, 0), (1, 1, 1),
                          grid_res=(0, 8, 8))
    o = oracle_mod.Oracle(p, np.zeros((1, 3), F))
    try:
        with pytest.raises(ValueError):
            o.splat_density()
    finally:
        o.close()
