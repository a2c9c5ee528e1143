"""CPU restatement of the LOD surrogate network (gell_amd/csrc/surrogate.hip,
nn.hip) in numpy float64 -- test infrastructure only (imported by tests/ as
the checker; the product never imports it).

The network has no counterpart in the reference: the reference's loop runs
every LOD substep (Main.cu:121-123). It predicts the field K substeps after
p0 from the first substep's increment r = p1 - p0, given in units of its
interior RMS s, and the uptake map; the prediction is p0 + s out_scale head
in the interior and p1 on the two outer layers of every face. What pins this
restatement:
  * its analytic backward against central finite differences of its own
    forward (float64, no rounding) -- tests/test_nn_oracle.py;
  * each conv layer against a direct per-voxel triple loop on tiny grids;
  * the device kernels against it with BF16 rounding emulated at the points
    where the device rounds (``rnd=bf16``).

Layout: channels-last arrays a[z, y, x, c]; conv weights W[cout, cin, 27]
with tap = (dz 3 + dy) 3 + dx reading input voxel (x + dx - 1, y + dy - 1,
z + dz - 1), zero outside the grid.
"""
from __future__ import annotations

import numpy as np

LAYERS = [  # (cin, cout, level, relu)
    (16, 16, 1, True), (16, 32, 2, True), (32, 32, 2, True), (32, 64, 3, True), (64, 64, 3, True),
    (64, 128, 4, True), (128, 128, 4, True), (128, 64, 3, True), (64, 64, 3, True), (64, 32, 2, True),
    (32, 32, 2, True), (32, 16, 1, True), (16, 8, 1, False),
]


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def ident(a):
    return np.asarray(a, dtype=np.float64)


def param_layout():
    """(w_off, b_off, total) of the flat parameter vector."""
    w_off, b_off, off = [], [], 0
    for cin, cout, _, _ in LAYERS:
        w_off.append(off)
        off += cout * cin * 27
        b_off.append(off)
        off += cout
    return w_off, b_off, off


def unpack(params):
    w_off, b_off, _ = param_layout()
    out = []
    for l, (cin, cout, _, _) in enumerate(LAYERS):
        W = np.asarray(params[w_off[l]:w_off[l] + cout * cin * 27], np.float64).reshape(cout, cin, 27)
        b = np.asarray(params[b_off[l]:b_off[l] + cout], np.float64)
        out.append((W, b))
    return out


def _shifts(S):
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                yield (dz * 3 + dy) * 3 + dx, dz, dy, dx


def conv3(x, W, b, relu, rnd=ident):
    """y = act(conv(x) + b); pre-activation accumulated in float64."""
    S = x.shape[0]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    acc = np.zeros(x.shape[:3] + (W.shape[0],))
    for tap, dz, dy, dx in _shifts(S):
        acc += xp[dz:dz + S, dy:dy + S, dx:dx + S, :] @ W[:, :, tap].T
    acc += b
    if relu:
        acc = np.maximum(acc, 0.0)
    return rnd(acc)


def conv3_back(x, W, dA, rnd=ident, need_dx=True):
    """(dx, dW, db) for dA = dL/d(pre-activation)."""
    S = x.shape[0]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dW = np.zeros_like(W)
    dxp = np.zeros(xp.shape)
    d2 = dA.reshape(-1, dA.shape[-1])
    for tap, dz, dy, dx in _shifts(S):
        sl = (slice(dz, dz + S), slice(dy, dy + S), slice(dx, dx + S))
        dW[:, :, tap] = (xp[sl].reshape(-1, x.shape[-1]).T @ d2).T
        if need_dx:
            dxp[sl] += dA @ W[:, :, tap]
    db = d2.sum(0)
    dx_ = rnd(dxp[1:-1, 1:-1, 1:-1]) if need_dx else None
    return dx_, dW, db


def pool(x, rnd=ident):
    S = x.shape[0] // 2
    return rnd(x.reshape(S, 2, S, 2, S, 2, -1).mean(axis=(1, 3, 5)))


def pool_back(dy, rnd=ident):
    return rnd(np.repeat(np.repeat(np.repeat(dy, 2, 0), 2, 1), 2, 2) / 8.0)


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_back(dy, rnd=ident):
    S = dy.shape[0] // 2
    return rnd(dy.reshape(S, 2, S, 2, S, 2, -1).sum(axis=(1, 3, 5)))


def increment_scale(p0, p1):
    """s = RMS of p1 - p0 over the interior (voxels two or more from every
    face), and 1 / s (0 when s = 0)."""
    m = interior(p0.shape[0])
    r = (np.asarray(p1, np.float64) - np.asarray(p0, np.float64))[m]
    s = float(np.sqrt((r * r).mean()))
    return s, (1.0 / s if s > 0 else 0.0)


def features(p0, p1, rate, rate_scale, rnd=ident):
    """16-channel input on the half grid (fine arrays indexed [z, y, x]):
    q = (p1 - p0) / s at the 8 fine voxels of each coarse voxel, then rate
    times rate_scale at the same 8 (fine voxel (z & 1) 4 + (y & 1) 2 +
    (x & 1)), each rounded. With rounding (the device's emulation) q and the
    scaled rate are formed in float32, as the device forms them."""
    _, inv = increment_scale(p0, p1)
    if rnd is ident:
        q = fine_to_head((np.asarray(p1, np.float64) - np.asarray(p0, np.float64)) * inv)
        rr = fine_to_head(np.asarray(rate, np.float64) * rate_scale)
    else:
        d = (np.float32(p1) - np.float32(p0)) * np.float32(inv)
        q = fine_to_head(d.astype(np.float64))
        rr = fine_to_head((np.float32(rate) * np.float32(rate_scale)).astype(np.float64))
    return np.concatenate([rnd(q), rnd(rr)], axis=-1)


def head_to_fine(h):
    """(X, X, X, 8) head -> (2X)^3 field, sub = (z & 1) 4 + (y & 1) 2 + (x & 1)."""
    X = h.shape[0]
    return h.reshape(X, X, X, 2, 2, 2).transpose(0, 3, 1, 4, 2, 5).reshape(2 * X, 2 * X, 2 * X)


def fine_to_head(d):
    X = d.shape[0] // 2
    return d.reshape(X, 2, X, 2, X, 2).transpose(0, 2, 4, 1, 3, 5).reshape(X, X, X, 8)


def interior(N):
    m = np.zeros((N, N, N), bool)
    m[2:N - 2, 2:N - 2, 2:N - 2] = True
    return m


def forward(params, p, p1, rate, cfg, rnd=ident, keep=False):
    """Predicted field (N^3) K substeps after p, given p1 = p after one
    substep; with keep, also the activations for backward."""
    Wb = unpack(params)
    wr = [(rnd(W), b) for W, b in Wb]
    c = lambda l, x: conv3(x, wr[l][0], wr[l][1], LAYERS[l][3], rnd)  # noqa: E731
    t = {}
    t["f0"] = features(p, p1, rate, cfg["rate_scale"], rnd)
    t["e1"] = c(0, t["f0"])
    t["p1"] = pool(t["e1"], rnd)
    t["a2"] = c(1, t["p1"])
    t["e2"] = c(2, t["a2"])
    t["p2"] = pool(t["e2"], rnd)
    t["a4"] = c(3, t["p2"])
    t["e3"] = c(4, t["a4"])
    t["p3"] = pool(t["e3"], rnd)
    t["a6"] = c(5, t["p3"])
    t["e4"] = c(6, t["a6"])
    t["u3"] = up(t["e4"])
    t["c3"] = c(7, t["u3"])
    t["x3"] = rnd(t["c3"] + t["e3"])
    t["d3"] = c(8, t["x3"])
    t["u2"] = up(t["d3"])
    t["c2"] = c(9, t["u2"])
    t["x2"] = rnd(t["c2"] + t["e2"])
    t["d2"] = c(10, t["x2"])
    t["u1"] = up(t["d2"])
    t["c1"] = c(11, t["u1"])
    t["x1"] = rnd(t["c1"] + t["e1"])
    t["h"] = c(12, t["x1"])
    s, _ = increment_scale(p, p1)
    out = p + s * cfg["out_scale"] * head_to_fine(t["h"])
    out = np.where(interior(p.shape[0]), out, p1)
    return (out, t, wr) if keep else out


def loss(params, p, p1, rate, target, cfg, rnd=ident):
    """Mean squared interior error of the increment in units of s out_scale
    (forward only)."""
    out = forward(params, p, p1, rate, cfg, rnd)
    N = p.shape[0]
    _, inv = increment_scale(p, p1)
    e = np.where(interior(N), (out - target) * inv / cfg["out_scale"], 0.0)
    return float((e * e).sum() / float((N - 4) ** 3))


def loss_and_grad(params, p, p1, rate, target, cfg, rnd=ident):
    """Mean squared interior error of the increment in units of s out_scale,
    and d loss / d params."""
    out, t, wr = forward(params, p, p1, rate, cfg, rnd, keep=True)
    N = p.shape[0]
    m = interior(N)
    cnt = float((N - 4) ** 3)
    _, inv = increment_scale(p, p1)
    e = np.where(m, head_to_fine(t["h"]) - (target - p) * inv / cfg["out_scale"], 0.0)
    loss = float((e * e).sum() / cnt)
    dh = rnd(fine_to_head(2.0 * e / cnt))
    grads = [None] * len(LAYERS)
    relu_mask = lambda g, y: np.where(y > 0, g, 0.0)  # noqa: E731

    def back(l, x, dA, need_dx=True):
        dx, dW, db = conv3_back(x, wr[l][0], dA, rnd, need_dx)
        grads[l] = (dW, db)
        return dx

    dx1 = back(12, t["x1"], dh)
    de1 = dx1
    g = relu_mask(dx1, t["c1"])
    du1 = back(11, t["u1"], g)
    g = relu_mask(up_back(du1, rnd), t["d2"])
    dx2 = back(10, t["x2"], g)
    de2 = dx2
    g = relu_mask(dx2, t["c2"])
    du2 = back(9, t["u2"], g)
    g = relu_mask(up_back(du2, rnd), t["d3"])
    dx3 = back(8, t["x3"], g)
    de3 = dx3
    g = relu_mask(dx3, t["c3"])
    du3 = back(7, t["u3"], g)
    g = relu_mask(up_back(du3, rnd), t["e4"])
    g = relu_mask(back(6, t["a6"], g), t["a6"])
    dp3 = back(5, t["p3"], g)
    de3 = rnd(pool_back(dp3, rnd) + de3)
    g = relu_mask(de3, t["e3"])
    g = relu_mask(back(4, t["a4"], g), t["a4"])
    dp2 = back(3, t["p2"], g)
    de2 = rnd(pool_back(dp2, rnd) + de2)
    g = relu_mask(de2, t["e2"])
    g = relu_mask(back(2, t["a2"], g), t["a2"])
    dp1 = back(1, t["p1"], g)
    de1 = rnd(pool_back(dp1, rnd) + de1)
    g = relu_mask(de1, t["e1"])
    back(0, t["f0"], g, need_dx=False)
    w_off, b_off, total = param_layout()
    flat = np.zeros(total)
    for l, (dW, db) in enumerate(grads):
        flat[w_off[l]:w_off[l] + dW.size] = dW.ravel()
        flat[b_off[l]:b_off[l] + db.size] = db
    return loss, flat


def init_params(seed=7):
    """He-normal weights (head scaled by 0.01), zero biases."""
    rng = np.random.default_rng(seed)
    w_off, b_off, total = param_layout()
    out = np.zeros(total)
    for l, (cin, cout, _, _) in enumerate(LAYERS):
        sd = np.sqrt(2.0 / (27 * cin)) * (0.01 if l == len(LAYERS) - 1 else 1.0)
        out[w_off[l]:w_off[l] + cout * cin * 27] = rng.normal(0.0, sd, cout * cin * 27)
    return out
