"""Pins of the oracle's diffusion (LODkernel.cuh:24-50,73-136,239-300).

The Thomas sweep is checked against a dense float64 solve of the system it
factorises; the substep against a numpy restatement built on those dense
solves; special cases against closed forms."""
import numpy as np
import pytest

from helpers import params, rel_l2


def dense_line_solve(p, side, corner, diag, bnd):
    """Interior values the reference's line sweep produces, from a dense
    solve of: row 0 [corner, side] = bnd (the injected Y[0]); rows 1..N-2
    [side, diag, side] = p[i]; x[N-1] = bnd."""
    N = len(p)
    A = np.zeros((N, N))
    r = np.zeros(N)
    A[0, 0], A[0, 1], r[0] = corner, side, bnd
    for i in range(1, N - 1):
        A[i, i - 1], A[i, i], A[i, i + 1] = side, diag, side
        r[i] = p[i]
    A[N - 1, N - 1], r[N - 1] = 1.0, bnd
    x = np.linalg.solve(A, r)
    x[0] = bnd
    return x


def coeffs64(P, s=0):
    dt, D, dx, lam = float(np.float32(P.dt_diff)), P.sub[s].D, P.dx, float(np.float32(P.sub[s].decay))
    side = -dt * D / (dx * dx)
    corner = 1.0 + dt * lam / 3.0 - side
    return side, corner, corner - side


@pytest.mark.parametrize("N,dx", [(8, 10.0), (33, 10.0), (64, 25.0), (100, 25.0)])
def test_thomas_line_matches_dense_solve(N, dx, oracle):
    P = params(N, dx=dx)
    E, F = oracle.lod_coeffs(P)
    side, corner, diag = coeffs64(P)
    # E/F are the LU factors of the same tridiagonal matrix
    assert abs(E[1] * F[0] - side) <= 1e-6 * abs(side)
    Fd = np.zeros(N)
    Fd[0] = corner
    for i in range(1, N):
        Fd[i] = (corner if i == N - 1 else diag) - side * side / Fd[i - 1]
    assert rel_l2(F, Fd) < 1e-6
    rng = np.random.default_rng(N)
    for trial in range(3):
        p = (30 + 10 * rng.random(N)).astype(np.float32)
        got = p.copy()
        oracle.thomas_line(got, E, F, 38.0)
        want = dense_line_solve(p.astype(np.float64), side, corner, diag, 38.0)
        assert rel_l2(got, want) < 2e-6
        assert got[0] == 38.0 and got[-1] == 38.0


def numpy_substep(p, rate, P, s=0):
    """LODsolver restated: boundary layers, implicit uptake, then x, y, z line
    solves by dense float64 solves."""
    N = P.N
    side, corner, diag = coeffs64(P, s)
    b = P.sub[s].boundary
    f = p.reshape(N, N, N).astype(np.float64).copy()  # [k, j, i]
    for l in (0, 1, N - 2, N - 1):
        f[:, :, l] = b
        f[:, l, :] = b
        f[l, :, :] = b
    f = f / (1.0 - rate.reshape(N, N, N).astype(np.float64) * float(np.float32(P.dt_diff)))
    for axis in (2, 1, 0):  # x, y, z
        g = np.moveaxis(f, axis, -1)
        flat = g.reshape(-1, N)
        for q in range(flat.shape[0]):
            flat[q] = dense_line_solve(flat[q], side, corner, diag, b)
        f = np.moveaxis(flat.reshape(g.shape), -1, axis)
    return f.ravel()


@pytest.mark.parametrize("N", [6, 12])
def test_substep_matches_dense_restatement(N, oracle):
    P = params(N)
    rng = np.random.default_rng(N)
    p = (20 + 18 * rng.random(N ** 3)).astype(np.float32)
    rate = (-3 * rng.random(N ** 3) * (rng.random(N ** 3) < 0.4)).astype(np.float32)
    got = p.copy()
    oracle.lod_substep(P, 0, got, rate)
    want = numpy_substep(p, rate, P)
    assert rel_l2(got, want) < 5e-6


def test_substep_without_diffusion_is_decay_and_uptake(oracle):
    # D = 0: E = 0, F = 1 + dt*lambda/3, so every interior voxel is
    # p / (1 - rate dt) / F^3 (x, y and z sweeps each divide once).
    N = 10
    P = params(N)
    P.sub[0].D = 0.0
    rng = np.random.default_rng(1)
    p = (10 + 20 * rng.random(N ** 3)).astype(np.float32)
    rate = (-rng.random(N ** 3)).astype(np.float32)
    got = p.copy()
    oracle.lod_substep(P, 0, got, rate)
    F = np.float32(1.0) + np.float32(P.dt_diff) * np.float32(P.sub[0].decay) / np.float32(3.0)
    dt = np.float32(P.dt_diff)
    g = got.reshape(N, N, N)
    inner = p.reshape(N, N, N)[2:-2, 2:-2, 2:-2] / (np.float32(1.0) - rate.reshape(N, N, N)[2:-2, 2:-2, 2:-2] * dt)
    want = ((inner / F) / F) / F
    assert np.array_equal(g[2:-2, 2:-2, 2:-2], want.astype(np.float32))
    # layer 1 of each face starts from the boundary value
    lay = np.float32(38.0) / (np.float32(1.0) - rate.reshape(N, N, N)[5, 5, 1] * dt)
    assert g[5, 5, 1] == ((lay / F) / F) / F
    # the z sweep runs last and pins its end planes (k = 0, N-1); the x and y
    # end planes are re-solved by the later sweeps
    assert np.all(g[0] == 38.0) and np.all(g[-1] == 38.0)
    assert not np.all(g[:, 2:-2, 0] == 38.0)


def test_uniform_boundary_field_is_a_fixed_point(oracle):
    N = 16
    P = params(N)
    P.sub[0].decay = 0.0
    p = np.full(N ** 3, 38.0, np.float32)
    oracle.lod_substep(P, 0, p, np.zeros(N ** 3, np.float32))
    assert np.max(np.abs(p - 38.0)) < 38.0 * 2e-6


def test_uptake_lowers_and_decay_relaxes(oracle):
    # a sink in the middle depletes it; the field stays within [min, boundary]
    N = 20
    P = params(N)
    p = np.full(N ** 3, 38.0, np.float32)
    rate = np.zeros(N ** 3, np.float32)
    mid = (N // 2) + N * ((N // 2) + N * (N // 2))
    rate[mid] = -50.0
    for _ in range(20):
        oracle.lod_substep(P, 0, p, rate)
    assert p[mid] < 38.0 - 0.1 and p[mid] == p.min()
    assert p.max() <= 38.0 + 1e-4 and p.min() > 0


def test_float64_oracle_agrees_with_float32(oracle):
    N = 24
    P = params(N)
    rng = np.random.default_rng(3)
    p32 = (20 + 18 * rng.random(N ** 3)).astype(np.float32)
    rate = (-2 * rng.random(N ** 3)).astype(np.float32)
    p64 = p32.astype(np.float64)
    for _ in range(5):
        oracle.lod_substep(P, 0, p32, rate)
        oracle.lod_substep(P, 0, p64, rate.astype(np.float64))
    assert rel_l2(p32, p64) < 1e-5  # five substeps of float32 rounding
