"""The two facts the force kernels' queued walk (fm.hip, cell_force_queued)
rests on, checked on the CPU against their definitions.

1. forcekernel's distance test (FMkernel.cuh:89-99): distance = sqrtf(x),
   x = (float)(ss + 0.0001), and a candidate is rejected when
   distance >= thres. The walk decides it without the square root:
   distance < t  <=>  (double)x < m^2, m = the midpoint of t and its float
   predecessor, m^2 formed in double. numpy's float32 sqrt is correctly
   rounded (IEEE 754), as the reference's sqrtf and the kernel's are.
2. A float sum that starts at +0 is never -0, so adding the +-0 force of a
   zero-displacement pair leaves it unchanged (round to nearest).
"""
from fractions import Fraction

import numpy as np

F32 = np.float32


def m2_of(t):
    """m^2 for thres t (float32 > 0), as the kernel forms it."""
    tp = np.frombuffer(np.array([t], F32).tobytes(), np.uint32)[0] - np.uint32(1)
    tp = np.frombuffer(np.array([tp], np.uint32).tobytes(), F32)[0]
    m = (np.float64(tp) + np.float64(t)) * 0.5
    return m * m


def x_of(ss):
    return F32(np.float64(ss) + 0.0001)


def below_sqrt(x, t):
    return bool(np.sqrt(F32(x)) < t)


def below_m2(x, t):
    return not (np.float64(x) >= m2_of(t))


def test_m2_is_exact_and_not_a_float():
    rng = np.random.default_rng(1)
    ts = np.concatenate([rng.uniform(0.0125, 30.0, 2000).astype(F32),
                         F32([0.0125, 0.5, 1.0, 2.0, 4.0, 8.0, 16.0, 25.0, 12.5, 13.75])])
    for t in ts:
        tp = F32(np.nextafter(t, F32(0)))
        m = (Fraction(float(tp)) + Fraction(float(t))) / 2
        assert Fraction(float(m2_of(t))) == m * m, t  # exact in double
        assert Fraction(float(F32(float(m * m)))) != m * m, t  # no float equals it: no tie


def test_sqrt_free_test_equals_sqrt_test_random():
    rng = np.random.default_rng(2)
    ss = rng.uniform(0.0, 900.0, 200000).astype(F32)
    ts = rng.uniform(0.0125, 30.0, 200000).astype(F32)
    x = (ss.astype(np.float64) + 0.0001).astype(F32)
    by_sqrt = np.sqrt(x) < ts
    tp = (ts.view(np.uint32) - 1).view(F32)
    m = (tp.astype(np.float64) + ts.astype(np.float64)) * 0.5
    by_m2 = ~(x.astype(np.float64) >= m * m)
    assert np.array_equal(by_sqrt, by_m2)


def test_sqrt_free_test_at_the_boundary():
    """x within a few ulps of t^2, where a dropped half-ulp or a wrong
    predecessor would show, for ordinary thresholds and powers of two."""
    rng = np.random.default_rng(3)
    ts = list(rng.uniform(0.0125, 30.0, 3000).astype(F32)) + [F32(v) for v in (0.5, 1, 2, 4, 8, 16, 0.25, 25)]
    n_diff = 0
    for t in ts:
        c = F32(np.float64(t) * np.float64(t))
        for k in range(-4, 5):
            x = c
            for _ in range(abs(k)):
                x = np.nextafter(x, F32(np.inf) if k > 0 else F32(0))
            a, b = below_sqrt(x, t), below_m2(x, t)
            assert a == b, (t, x)
            n_diff += a != below_sqrt(c, t)
    assert n_diff > 0  # the sweep crosses the boundary


def test_sqrt_free_test_nan_and_inf():
    t = F32(12.5)
    nan = F32(np.nan)
    # the reference: distance = NaN, `distance >= thres` false -> not rejected
    assert not (np.sqrt(nan) >= t) and below_m2(nan, t)
    assert not below_m2(F32(np.inf), t) and not below_sqrt(F32(np.inf), t)
    assert below_m2(x_of(F32(0.0)), F32(0.0125))  # the cell itself passes every cutoff


def test_plus_zero_sum_ignores_signed_zeros():
    rng = np.random.default_rng(4)
    for _ in range(2000):
        k = int(rng.integers(1, 12))
        terms = rng.normal(0, 1, k).astype(F32)
        terms[rng.random(k) < 0.3] = F32(0.0)
        terms[rng.random(k) < 0.2] *= F32(-1.0)
        if rng.random() < 0.3 and k > 1:
            terms[1] = -terms[0]  # an exact cancellation to +0 on the way
        for z_at in range(k + 1):
            # the same sum with a +-0 term inserted at position z_at
            for z in (F32(0.0), F32(-0.0)):
                a = F32(0.0)
                b = F32(0.0)
                for i, v in enumerate(terms):
                    if i == z_at:
                        b = F32(b + z)
                    a = F32(a + v)
                    b = F32(b + v)
                if z_at == k:
                    b = F32(b + z)
                assert a.tobytes() == b.tobytes()
                assert not (a == 0 and np.signbit(a))
