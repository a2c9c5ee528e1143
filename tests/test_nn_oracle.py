"""Pins of the surrogate network's CPU restatement (oracle/nn_oracle.py):
its convolution against a direct per-voxel loop, pooling / upsampling
against their adjoint identities, its analytic gradient of the whole
network against central finite differences of its own forward pass, and the
two facts its input and output rest on: an LOD substep with a fixed uptake
map is affine in the field (so the K-substep increment is a linear map of
the first substep's increment), and the prediction's increment scales with
the input increment."""
import numpy as np
import pytest

import nn_oracle as nno
from helpers import params as make_params

CFG = dict(rate_scale=0.1, out_scale=8.0)


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
    p1 = p - 0.05 * rng.random((N, N, N))
    rate = -25.0 * (rng.random((N, N, N)) < 0.1)
    # a K-substep increment of a few first increments: O(1) in units of s out_scale
    target = p + 8.0 * (p1 - p) * (1.0 + rng.random((N, N, N)))
    params = nno.init_params(5)
    w_off, b_off, _ = nno.param_layout()
    for l in range(len(nno.LAYERS)):  # non-zero biases so every bias gradient path is exercised
        params[b_off[l]:b_off[l] + nno.LAYERS[l][1]] = 0.05 * rng.normal(size=nno.LAYERS[l][1])
    return p, p1, rate, target, params


def test_network_gradient_matches_finite_differences(small_problem):
    p, p1, rate, target, params = small_problem
    loss, g = nno.loss_and_grad(params, p, p1, rate, target, CFG)
    assert loss > 0 and abs(loss - nno.loss(params, p, p1, rate, target, CFG)) <= 1e-12 * loss
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
        lp = nno.loss(pp, p, p1, rate, target, CFG)
        lm = nno.loss(pm, p, p1, rate, target, CFG)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1e-3, abs(fd)) + 1e-8, (i, fd, g[i])


def test_lod_substep_is_affine_in_the_field(oracle):
    """S(a p + (1 - a) q) = a S(p) + (1 - a) S(q) for a fixed uptake map
    (float64 oracle), so the K-substep increment from p is linear in the
    first substep's increment: two fields whose first increments differ by a
    factor have K-substep increments that differ by the same factor."""
    N = 32
    P = make_params(N)
    rng = np.random.default_rng(8)
    rate = np.where(rng.random(N ** 3) < 0.2, -20.0 * rng.random(N ** 3), 0.0)
    p = 38.0 - 10.0 * rng.random(N ** 3)
    q = 38.0 - 10.0 * rng.random(N ** 3)

    def step(f, k=1):
        f = f.copy()
        for _ in range(k):
            oracle.lod_substep(P, 0, f, rate)
        return f

    for a in (0.3, 2.5):
        lhs = step(a * p + (1 - a) * q)
        rhs = a * step(p) + (1 - a) * step(q)
        assert np.abs(lhs - rhs).max() <= 1e-12 * np.abs(rhs).max()
    # a fixed point of the substep under this map, then two fields on one
    # line through it: increments r and 3 r give K-substep increments D and 3 D
    fix = step(np.full(N ** 3, 38.0), 4000)
    assert np.abs(step(fix) - fix).max() <= 1e-10
    d = rng.normal(size=N ** 3)
    f1, f3 = fix + d, fix + 3.0 * d
    r1, r3 = step(f1) - f1, step(f3) - f3
    D1, D3 = step(f1, 50) - f1, step(f3, 50) - f3
    assert np.allclose(r3, 3.0 * r1, atol=1e-9) and np.allclose(D3, 3.0 * D1, atol=1e-9)
    assert np.abs(D1 - r1).max() > 1e-3  # 50 substeps are not one


def test_prediction_increment_scales_with_the_input_increment(small_problem):
    """Features are r / s and the output p0 + s out_scale head: scaling the
    first increment by a scales the predicted increment by a (interior),
    the outer layers are p1's, and r = 0 predicts p0."""
    p, p1, rate, _, params = small_problem
    m = nno.interior(p.shape[0])
    base = nno.forward(params, p, p1, rate, CFG) - p
    for a in (0.25, 4.0):
        q1 = p + a * (p1 - p)
        got = nno.forward(params, p, q1, rate, CFG)
        assert np.allclose((got - p)[m], a * base[m], rtol=1e-10, atol=1e-12)
        assert np.array_equal(got[~m], q1[~m])
    still = nno.forward(params, p, p.copy(), rate, CFG)
    assert np.array_equal(still, p)
