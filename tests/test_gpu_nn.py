"""The surrogate network's device kernels against the CPU restatement
(oracle/nn_oracle.py) with BF16 rounding emulated where the device rounds:
each 3x3x3 layer shape (forward, input gradient, weight gradient), then the
whole network's prediction, loss and parameter gradients, then a few Adam
steps. Tolerances: one BF16 rounding step (2^-8 relative) on BF16 outputs,
1e-5 on FP32-accumulated weight gradients of exact BF16 products, and
5e-2 relative L2 on whole-network gradients (13 layers of BF16 activations
and gradients)."""
import numpy as np
import pytest

from gell_amd import runtime as rt
import nn_oracle as nno
from helpers import log, rel_l2

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 64), (64, 32),
          (32, 16), (16, 8)]


def to_bf16_bits(a):
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bits(h):
    return (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def dev(h):
    b = rt.DeviceBuffer(h.nbytes)
    b.upload(h)
    return b


def close_bf16(got, want):
    scale = max(np.abs(want).max(), 1e-6)
    err = np.abs(got - want)
    return np.all(err <= 2.0 ** -7 * np.abs(want) + 2.0 ** -9 * 1e-2 * scale), float(err.max() / scale)


@pytest.mark.parametrize("cin,cout,S", [(128, 64, 8), (64, 32, 8), (32, 16, 16), (64, 32, 32)])
def test_upconv_layer_equals_the_layer_on_the_upsampled_tensor(cin, cout, S):
    """The layers after an upsampling run on the coarse grid as 8 per-parity
    2x2x2 convolutions with summed FP32 taps (rounded to BF16 once); against
    the oracle's 3x3x3 convolution of the formed (2S)^3 upsampling with the
    unrounded weights, ReLU, BF16, + skip, BF16: within the rounding of the
    combined weights (2^-9 relative each) and of the outputs. (32, 16, 16)
    and (64, 32, 32) take the form whose blocks hold every output channel and
    stage their output in LDS; the others one channel tile per block."""
    rng = np.random.default_rng(cin * 7 + cout)
    xb = to_bf16_bits(rng.normal(size=(S, S, S, cin)))
    x = from_bits(xb)
    W = rng.normal(0, np.sqrt(2 / (27 * cin)), size=(cout, cin, 27)).astype(np.float32).astype(np.float64)
    b = rng.normal(0, 0.1, size=cout).astype(np.float32).astype(np.float64)
    F = 2 * S
    skb = to_bf16_bits(rng.normal(size=(F, F, F, cout)))
    xd, skd = dev(xb), dev(skb)
    yd = rt.DeviceBuffer(F ** 3 * cout * 2)
    conv = nno.conv3(nno.up(x), W, b, True)
    for with_skip in (False, True):
        rt.nn_upconv3(xd, W, b, yd, cin, cout, S, relu=True, skip=skd if with_skip else None)
        got = from_bits(yd.download(np.uint16, F ** 3 * cout)).reshape(F, F, F, cout)
        want = nno.bf16(conv)
        if with_skip:
            want = nno.bf16(want + from_bits(skb).reshape(F, F, F, cout))
        e = rel_l2(got, want)
        log(f"[nn] upconv {cin}->{cout} at {S}^3 -> {F}^3, skip {with_skip}: rel L2 {e:.2e}")
        assert e < 4e-3
    for buf in (xd, skd, yd):
        buf.free()


@pytest.mark.parametrize("cin,cout,S", [(16, 16, 32), (32, 32, 64), (16, 32, 16), (64, 64, 32)])
def test_conv_pooled_in_epilogue_equals_pooling_kernel(cin, cout, S):
    """The pooling formed in a convolution's epilogue (lanes' shuffles and
    pairs of z tiles) equals the pooling kernel on the stored output bit for
    bit: the mean of the eight BF16 values summed in float32 in x, y, z order
    and rounded to BF16. (64 -> 64 at 32^3 runs the K-split kernel, which does
    not pool: the call pools with the kernel.)"""
    rng = np.random.default_rng(cin + cout + S)
    xb = to_bf16_bits(rng.normal(size=(S, S, S, cin)))
    W = nno.bf16(rng.normal(0, np.sqrt(2 / (27 * cin)), size=(cout, cin, 27)))
    b = rng.normal(0, 0.1, size=cout).astype(np.float32).astype(np.float64)
    xd = dev(xb)
    yd = rt.DeviceBuffer(S ** 3 * cout * 2)
    pd = rt.DeviceBuffer((S // 2) ** 3 * cout * 2)
    fused = rt.nn_conv3_pooled(xd, W, b, yd, pd, cin, cout, S, relu=True)
    log(f"[nn] conv {cin}->{cout} at {S}^3: pooled in the epilogue {fused}")
    assert fused == (cin != 64)
    y = from_bits(yd.download(np.uint16, S ** 3 * cout)).reshape(S, S, S, cout).astype(np.float32)
    h = S // 2
    acc = np.zeros((h, h, h, cout), np.float32)
    for d in range(8):  # x fastest, then y, then z; arrays are [z, y, x, c]
        acc += y[(d >> 2)::2, ((d >> 1) & 1)::2, (d & 1)::2]
    want = to_bf16_bits(acc * np.float32(0.125))
    got = pd.download(np.uint16, h ** 3 * cout).reshape(h, h, h, cout)
    assert np.array_equal(got, want)
    for buf in (xd, yd, pd):
        buf.free()


@pytest.mark.parametrize("cin,cout", SHAPES)
def test_conv_layer_forward_and_gradients(cin, cout):
    S = 16 if cin <= 64 else 8
    rng = np.random.default_rng(cin * 1000 + cout)
    xb = to_bf16_bits(rng.normal(size=(S, S, S, cin)))
    x = from_bits(xb)
    W = nno.bf16(rng.normal(0, np.sqrt(2 / (27 * cin)), size=(cout, cin, 27)))
    b = rng.normal(0, 0.1, size=cout).astype(np.float32).astype(np.float64)
    xd = dev(xb)
    yd = rt.DeviceBuffer(S ** 3 * cout * 2)
    rt.nn_conv3(xd, W, b, yd, cin, cout, S, relu=True)
    got = from_bits(yd.download(np.uint16, S ** 3 * cout)).reshape(S, S, S, cout)
    want = nno.bf16(nno.conv3(x, W, b, True))
    ok, e = close_bf16(got, want)
    assert ok, f"forward: max err {e:.3e}"
    # input gradient: dx = conv of dy with the flipped weights
    dyb = to_bf16_bits(rng.normal(size=(S, S, S, cout)))
    dy = from_bits(dyb)
    dyd = dev(dyb)
    dxd = rt.DeviceBuffer(S ** 3 * cin * 2)
    rt.nn_conv3(dyd, W, np.zeros(cout), dxd, cin, cout, S, input_grad=True)
    gdx = from_bits(dxd.download(np.uint16, S ** 3 * cin)).reshape(S, S, S, cin)
    wdx, wdW, _ = nno.conv3_back(x, W, dy)
    ok, e = close_bf16(gdx, nno.bf16(wdx))
    assert ok, f"input gradient: max err {e:.3e}"
    # weight gradient (exact BF16 products, FP32 sums)
    gdW = rt.nn_conv3_wgrad(xd, dyd, cin, cout, S)
    assert rel_l2(gdW.astype(np.float64), wdW) < 1e-5
    for buf in (xd, yd, dyd, dxd):
        buf.free()


CFG = dict(rate_scale=0.1, out_scale=8.0)


@pytest.fixture(scope="module")
def net_problem():
    N = 128
    rng = np.random.default_rng(11)
    p = (38.0 - 6.0 * rng.random((N, N, N))).astype(np.float32)
    p1 = (p - 0.05 * rng.random((N, N, N))).astype(np.float32)
    rate = (-25.0 * (rng.random((N, N, N)) < 0.1)).astype(np.float32)
    target = (p + 8.0 * (p1 - p) * (1.0 + rng.random((N, N, N)))).astype(np.float32)
    sur = rt.Surrogate(N)
    params = sur.params().astype(np.float64)
    # non-zero biases so the bias paths carry signal
    w_off, b_off, _ = nno.param_layout()
    for l, (cin, cout, _, _) in enumerate(nno.LAYERS):
        params[b_off[l]:b_off[l] + cout] = 0.05 * rng.normal(size=cout)
    params = params.astype(np.float32)
    sur.set_params(params)
    bufs = [dev(a) for a in (p, p1, rate, target)]
    yield N, p, p1, rate, target, params.astype(np.float64), sur, bufs
    for b in bufs:
        b.free()
    sur.close()


def test_network_prediction_matches_oracle(net_problem):
    N, p, p1, rate, target, params, sur, (pd, p1d, rd, td) = net_problem
    out = rt.DeviceBuffer(N ** 3 * 4)
    sur.predict(pd, p1d, rd, out)
    got = out.download(np.float32, N ** 3).reshape(N, N, N).astype(np.float64)
    want = nno.forward(params, p.astype(np.float64), p1.astype(np.float64), rate.astype(np.float64), CFG,
                       rnd=nno.bf16)
    e = rel_l2(got - p, want - p)
    log(f"[nn] prediction rel L2 of the update vs oracle: {e:.3e}")
    assert e < 1e-2
    m = nno.interior(N)
    assert np.array_equal(got[~m], p1.astype(np.float64)[~m])  # outer layers: the substep's
    # out may alias field1: the prediction written over it is the same
    p1copy = dev(p1)
    sur.predict(pd, p1copy, rd, p1copy)
    assert np.array_equal(p1copy.download(np.float32, N ** 3), out.download(np.float32, N ** 3))
    p1copy.free()
    out.free()


def test_network_gradients_match_oracle(net_problem):
    N, p, p1, rate, target, params, sur, (pd, p1d, rd, td) = net_problem
    loss, g = sur.gradients(pd, p1d, rd, td, want_grad=True)
    wl, wg = nno.loss_and_grad(params, p.astype(np.float64), p1.astype(np.float64), rate.astype(np.float64),
                               target.astype(np.float64), CFG, rnd=nno.bf16)
    log(f"[nn] loss device {loss:.6f} oracle {wl:.6f}")
    assert abs(loss - wl) <= 1e-2 * wl
    w_off, b_off, _ = nno.param_layout()
    worst = 0.0
    for l, (cin, cout, _, _) in enumerate(nno.LAYERS):
        sw = slice(w_off[l], w_off[l] + cout * cin * 27)
        sb = slice(b_off[l], b_off[l] + cout)
        ew, eb = rel_l2(g[sw], wg[sw]), rel_l2(g[sb], wg[sb])
        worst = max(worst, ew, eb)
        log(f"[nn] layer {l} ({cin}->{cout}): weight grad rel L2 {ew:.2e}, bias grad {eb:.2e}")
    assert worst < 5e-2


def test_adam_step_matches_its_formula(net_problem):
    """One device Adam step from fresh moments equals the host formula on the
    returned gradients: w - lr (g (1 - b1) / c1) / (sqrt(g^2 (1 - b2) / c2) + eps)."""
    N, p, p1, rate, target, params, sur, (pd, p1d, rd, td) = net_problem
    sur.set_params(params.astype(np.float32))  # also clears the moments
    w0 = sur.params().astype(np.float64)
    _, g = sur.gradients(pd, p1d, rd, td, want_grad=True)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    sur.adam(lr)
    w1 = sur.params().astype(np.float64)
    g = g.astype(np.float64)
    m, v = (1 - b1) * g, (1 - b2) * g * g
    want = w0 - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
    assert np.max(np.abs(w1 - want)) <= 1e-6 * (1 + np.abs(want).max())


def test_training_reduces_the_loss(net_problem):
    """A uniform multiple of the first increment: 40 small Adam steps lower
    the loss."""
    N, p, p1, rate, target, params, sur, (pd, p1d, rd, td) = net_problem
    sur.set_params(params.astype(np.float32))
    shifted = dev((p + 20.0 * (p1 - p)).astype(np.float32))
    losses = []
    for _ in range(40):
        losses.append(sur.gradients(pd, p1d, rd, shifted))
        sur.adam(3e-4)
    log(f"[nn] losses {losses[0]:.5f} -> {min(losses[-5:]):.5f}")
    assert min(losses[-5:]) < 0.7 * losses[0]
    shifted.free()
