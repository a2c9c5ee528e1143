"""Pins of the surrogate network's CPU restatement (oracle/nn_oracle.py):
its convolution against a direct per-voxel loop, pooling / upsampling
against their adjoint identities, and its analytic gradient of the whole
network against central finite differences of its own forward pass."""
import numpy as np
import pytest

import nn_oracle as nno

CFG = dict(boundary=38.0, rate_scale=0.1, out_scale=1.0)


def test_conv3_matches_direct_loop():
    rng = np.random.default_rng(1)
    S, cin, cout = 4, 2, 3
    x = rng.normal(size=(S, S, S, cin))
    W = rng.normal(size=(cout, cin, 27))
    b = rng.normal(size=cout)
    want = np.zeros((S, S, S, cout))
    for z in range(S):
        for y in range(S):
            for xx in range(S):
                for o in range(cout):
                    s = b[o]
                    for dz in range(3):
                        for dy in range(3):
                            for dx in range(3):
                                zi, yi, xi = z + dz - 1, y + dy - 1, xx + dx - 1
                                if 0 <= zi < S and 0 <= yi < S and 0 <= xi < S:
                                    s += W[o, :, (dz * 3 + dy) * 3 + dx] @ x[zi, yi, xi]
                    want[z, y, xx, o] = s
    assert np.allclose(nno.conv3(x, W, b, False), want, atol=1e-12)
    assert np.allclose(nno.conv3(x, W, b, True), np.maximum(want, 0), atol=1e-12)


def test_conv3_back_is_the_adjoint():
    rng = np.random.default_rng(2)
    S, cin, cout = 5, 3, 4
    x = rng.normal(size=(S, S, S, cin))
    W = rng.normal(size=(cout, cin, 27))
    dA = rng.normal(size=(S, S, S, cout))
    dx, dW, db = nno.conv3_back(x, W, dA)
    zero = np.zeros(cout)
    # <conv(x), dA> is linear in x and in W
    assert np.isclose((nno.conv3(x, W, zero, False) * dA).sum(), (x * dx).sum())
    assert np.isclose((nno.conv3(x, W, zero, False) * dA).sum(), (W * dW).sum())
    assert np.allclose(db, dA.sum((0, 1, 2)))


def test_pool_and_up_adjoints():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(8, 8, 8, 3))
    y = rng.normal(size=(4, 4, 4, 3))
    assert np.isclose((nno.pool(x) * y).sum(), (x * nno.pool_back(y)).sum())
    assert np.isclose((nno.up(y) * x).sum(), (y * nno.up_back(x)).sum())
    h = rng.normal(size=(4, 4, 4, 8))
    assert np.array_equal(nno.fine_to_head(nno.head_to_fine(h)), h)
    f = nno.head_to_fine(h)
    assert f[1, 0, 1] == h[0, 0, 0, 5] and f[0, 1, 0] == h[0, 0, 0, 2]


def test_bf16_rounding():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1e-30])
    r = nno.bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == -2.5
    assert abs(r[4] - 1e-30) < 1e-32


@pytest.fixture(scope="module")
def small_problem():
    rng = np.random.default_rng(4)
    N = 32
    p = 38.0 - 6.0 * rng.random((N, N, N))
    rate = -25.0 * (rng.random((N, N, N)) < 0.1)
    target = p - 2.0 * rng.random((N, N, N))
    params = nno.init_params(5)
    w_off, b_off, _ = nno.param_layout()
    for l in range(len(nno.LAYERS)):  # non-zero biases so every bias gradient path is exercised
        params[b_off[l]:b_off[l] + nno.LAYERS[l][1]] = 0.05 * rng.normal(size=nno.LAYERS[l][1])
    return p, rate, target, params


def test_network_gradient_matches_finite_differences(small_problem):
    p, rate, target, params = small_problem
    loss, g = nno.loss_and_grad(params, p, rate, target, CFG)
    assert loss > 0 and abs(loss - nno.loss(params, p, rate, target, CFG)) <= 1e-12 * loss
    w_off, b_off, total = nno.param_layout()
    rng = np.random.default_rng(6)
    idx = []
    for l in range(len(nno.LAYERS)):  # weights and a bias of every layer
        n = nno.LAYERS[l][0] * nno.LAYERS[l][1] * 27
        idx += list(w_off[l] + rng.choice(n, 2, replace=False)) + [b_off[l]]
    for i in idx:
        # small steps keep ReLU kinks out of the difference quotient
        h = 1e-7 * max(1.0, abs(params[i]))
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        lp = nno.loss(pp, p, rate, target, CFG)
        lm = nno.loss(pm, p, rate, target, CFG)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1e-3, abs(fd)) + 1e-8, (i, fd, g[i])
