"""Pins the FP64 warm-start MLP oracle (oracle/mlp.py) before it checks the HIP kernels.

The network has no reference counterpart; its forward pass is pinned by closed forms and
its backward pass by central finite differences of the FP64 loss, so a dropped term, a
transposed operand or a wrong GELU derivative fails here.
"""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
from oracle import mlp  # noqa: E402


def test_parameter_count_matches_header():
    assert mlp.N_PARAMS == 9217 == 12 * 64 + 64 + 2 * (64 * 64 + 64) + 64 + 1
    p = mlp.init_params(3)
    assert np.array_equal(mlp.pack(*mlp.unpack(p)), p)


def test_bf16_round_to_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7
    vals = np.array([one, one + ulp / 2, one + 1.5 * ulp, one + 0.75 * ulp, -one - ulp / 2, 3.0e38, 0.0])
    want = np.array([one, one, one + 2 * ulp, one + ulp, -one, 2.9908e38, 0.0])
    got = mlp.bf16(vals)
    assert np.array_equal(got[:5], want[:5])
    assert abs(got[5] / 3.0e38 - 1) < 2 ** -8
    assert got[6] == 0.0
    # idempotent and within half an ulp
    x = np.random.default_rng(0).normal(size=1000) * 10
    b = mlp.bf16(x)
    assert np.array_equal(mlp.bf16(b), b)
    assert np.all(np.abs(b - x) <= np.abs(x) * 2.0 ** -8 * 1.0000001)


def test_gelu_closed_form():
    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    want = [0.5 * v * (1 + math.erf(v / math.sqrt(2))) for v in z]
    assert np.allclose(mlp.gelu(z), want, rtol=0, atol=1e-15)
    eps = 1e-6
    fd = (mlp.gelu(z + eps) - mlp.gelu(z - eps)) / (2 * eps)
    assert np.allclose(mlp.gelu_grad(z), fd, rtol=0, atol=1e-8)


def test_forward_single_active_unit():
    # only W1[0,0] = a, W2[0,0] = W3[0,0] = w4[0] = 1: y = gelu(gelu(gelu(a x0))) + b4
    W1 = np.zeros((12, 64)); W1[0, 0] = 0.7
    W2 = np.zeros((64, 64)); W2[0, 0] = 1.0
    W3 = np.zeros((64, 64)); W3[0, 0] = 1.0
    w4 = np.zeros(64); w4[0] = 1.0
    z = np.zeros(64)
    p = mlp.pack(W1, z, W2, z, W3, z, w4, 0.25)
    x = np.zeros((3, 12)); x[:, 0] = [-1.0, 0.5, 2.0]
    g = lambda v: 0.5 * v * (1 + math.erf(v / math.sqrt(2)))  # noqa: E731
    want = [g(g(g(0.7 * v))) + 0.25 for v in x[:, 0]]
    y, _ = mlp.forward(p, x)
    assert np.allclose(y, want, rtol=0, atol=1e-14)


def test_zero_weights_gradient_closed_form():
    p = np.zeros(mlp.N_PARAMS)
    p[-1] = 0.3
    x = np.random.default_rng(1).normal(size=(5, 12))
    t = np.array([0.0, 1.0, -1.0, 0.5, 2.0])
    sse, g = mlp.loss_grad(p, x, t)
    assert math.isclose(sse, float(((0.3 - t) ** 2).sum()), rel_tol=1e-14)
    assert math.isclose(g[-1], 2.0 * float((0.3 - t).mean()), rel_tol=1e-14)
    # gelu(0) = 0, so every activation and every other gradient is zero except db3..db1 = 0 too
    assert np.count_nonzero(g[:-1]) == 0


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_central_differences(seed):
    rng = np.random.default_rng(seed)
    p = mlp.init_params(seed, scale=1.5)
    x = rng.normal(size=(7, 12))
    t = rng.normal(size=7)
    _, g = mlp.loss_grad(p, x, t)
    # every parameter block, a sample of entries each
    offs = [0, 768, 832, 4928, 4992, 9088, 9152, 9216]
    idx = []
    for a, b in zip(offs, offs[1:] + [mlp.N_PARAMS]):
        idx += list(rng.choice(np.arange(a, b), size=min(12, b - a), replace=False))
    eps = 1e-6
    for k in idx:
        pp = p.copy(); pp[k] += eps
        pm = p.copy(); pm[k] -= eps
        fd = (mlp.loss_grad(pp, x, t)[0] - mlp.loss_grad(pm, x, t)[0]) / 7 / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * (1 + abs(g[k])), (k, fd, g[k])


def test_bf16_points_gradient_close_to_fp64():
    rng = np.random.default_rng(5)
    p = mlp.init_params(5)
    x = rng.normal(size=(256, 12))
    t = rng.normal(size=256)
    _, g64 = mlp.loss_grad(p, x, t)
    _, g16 = mlp.loss_grad(p, x, t, bf16_points=True)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 2e-2
    assert not np.array_equal(g16, g64)


def test_adam_first_step_moves_by_lr_sign():
    g = np.array([3.0, -0.5, 1e-3, 0.0])
    p, m1, m2 = mlp.adam_step(np.zeros(4), np.zeros(4), np.zeros(4), g, 1, 0.01)
    assert np.allclose(p[:3], -0.01 * np.sign(g[:3]), rtol=1e-4)
    assert p[3] == 0.0
