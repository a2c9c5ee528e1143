"""The partitioned line solve (GELL_KNOB_LOD_PARTS), emulated on the CPU with
the library's own host-built coefficient table, against a dense solve of the
reference's line system -- run before the kernels that use it, and without a
device."""
import numpy as np
import pytest

from gell_amd import runtime as rt
from helpers import params, rel_l2
from test_oracle_lod import coeffs64, dense_line_solve


def emulate(tab, d, N, P, side, bnd):
    """float32 replay of part_solve for a batch of lines d[lines, N]."""
    f = np.float32
    m = N // P
    El, iFl, W, V = (tab[q * N:(q + 1) * N] for q in range(4))
    Rinv = tab[4 * N:].reshape(2 * P, 2 * P)
    d = d.astype(np.float32).copy()
    d[:, 0] = bnd
    d[:, N - 1] = bnd
    y = np.zeros_like(d)
    for k in range(P):
        a = k * m
        v = d[:, a:a + m].copy()
        for i in range(1, m):
            v[:, i] = v[:, i] - El[a + i] * v[:, i - 1]
        v[:, m - 1] = v[:, m - 1] * iFl[a + m - 1]
        for i in range(m - 2, -1, -1):
            v[:, i] = (v[:, i] - f(side) * v[:, i + 1]) * iFl[a + i]
        y[:, a:a + m] = v
    g = np.zeros((d.shape[0], 2 * P), np.float32)
    for k in range(P):
        g[:, 2 * k] = y[:, k * m]
        g[:, 2 * k + 1] = y[:, k * m + m - 1]
    x = y.copy()
    for k in range(P):
        a = k * m
        um1 = np.zeros(d.shape[0], np.float32)
        tp1 = np.zeros(d.shape[0], np.float32)
        if k > 0:
            for q in range(2 * P):
                um1 = um1 + Rinv[2 * (k - 1) + 1, q] * g[:, q]
        if k < P - 1:
            for q in range(2 * P):
                tp1 = tp1 + Rinv[2 * (k + 1), q] * g[:, q]
        x[:, a:a + m] = y[:, a:a + m] - um1[:, None] * W[a:a + m] - tp1[:, None] * V[a:a + m]
    x[:, 0] = bnd
    return x


@pytest.mark.parametrize("N,P,dx", [(128, 2, 10.0), (128, 4, 10.0), (192, 2, 10.0), (256, 2, 10.0),
                                    (256, 4, 10.0), (256, 4, 25.0)])
def test_partitioned_solve_matches_dense(N, P, dx):
    Pm = params(N, dx=dx)
    tab = rt.lod_partition_table(Pm, 0, P)
    side, corner, diag = coeffs64(Pm)
    rng = np.random.default_rng(N + P)
    d = (5 + 33 * rng.random((6, N))).astype(np.float32)
    got = emulate(tab, d, N, P, np.float32(side), 38.0)
    for r in range(d.shape[0]):
        want = dense_line_solve(d[r].astype(np.float64), side, corner, diag, 38.0)
        assert rel_l2(got[r], want) < 1e-6, (r, rel_l2(got[r], want))
        assert got[r][0] == 38.0 and got[r][-1] == 38.0
