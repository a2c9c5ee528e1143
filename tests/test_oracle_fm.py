"""Pins of the oracle's mechanics (FMkernel.cuh:22-202) by independent
computations: numpy keys, numpy stable argsort, brute-force ranges, an
O(n^2) float64 force sum, and the AB2 closed form."""
import numpy as np
import pytest

from gell_amd import runtime as rt
from helpers import mixed_cells, params, rel_l2
from test_oracle_basic import morton_numpy


def keys_numpy(cells, n, N, dx):
    q = cells["pos"][:n] / np.float32(dx)
    ijk = np.trunc(q).astype(np.int64)
    inside = np.all((ijk >= 0) & (ijk < N), axis=1)
    key = np.where(inside, morton_numpy(ijk[:, 0], ijk[:, 1], ijk[:, 2]), -1)
    lin = ijk[:, 0] + N * (ijk[:, 1] + N * ijk[:, 2])
    return key, inside, lin


def test_keys_match_numpy(oracle):
    N, dx, n = 32, 10.0, 3000
    P = params(N, dx=dx, capacity=n)
    cells = mixed_cells(N, dx, n, n, seed=1, outside=50)
    cells["pos"][:10, 1] = np.float32(-5.0)   # -dx < y < 0 truncates to voxel 0: inside
    cells["pos"][10:20, 2] = np.float32(-15.0)  # voxel -1: outside
    want, inside, lin = keys_numpy(cells, n, N, dx)
    got_cells = cells.copy()
    keys = oracle.fm_keys(P, got_cells, n)
    assert np.array_equal(keys, want)
    assert np.all(got_cells["sign"][:n][~inside] == 0)
    assert np.array_equal(got_cells["sign"][:n][inside], cells["sign"][:n][inside])
    assert np.array_equal(got_cells["mesh_idx"][:n][inside], lin[inside])
    assert inside[:10].all() and not inside[10:20].any()


def test_sort_is_stable_and_ranges_brute_force(oracle):
    N, dx, n = 24, 10.0, 4000
    P = params(N, dx=dx, capacity=n)
    cells = mixed_cells(N, dx, n, n, seed=2, dense=True, dead_frac=0.2, outside=40)
    key, inside, lin = keys_numpy(cells, n, N, dx)
    want_order = np.argsort(key, kind="stable")
    work = cells.copy()
    keys, order, s, e = oracle.fm_update(P, work, n)
    assert np.array_equal(order, want_order)
    assert np.array_equal(keys, key[want_order])
    assert np.all(keys[:-1] <= keys[1:])
    # ranges: min/max sorted slot of live cells per voxel (after the key pass)
    live = (cells["sign"][:n] != 0) & inside
    ws = np.full(N ** 3, P.capacity + 1, np.int64)  # Max_Cell_num + 1 sentinel
    we = np.full(N ** 3, -1, np.int64)
    for pos, src in enumerate(want_order):
        if live[src]:
            v = lin[src]
            ws[v] = min(ws[v], pos)
            we[v] = max(we[v], pos)
    assert np.array_equal(s, ws) and np.array_equal(e, we)
    # Check_se (FMkernel.cuh:121-128,204-211): sum of range lengths counts every
    # sorted slot that lies inside some occupied range, dead ones included
    occ = we >= 0
    total = int((we[occ] - ws[occ] + 1).sum())
    covered = np.zeros(n, bool)
    for v in np.nonzero(occ)[0]:
        covered[ws[v]:we[v] + 1] = True
    assert total == covered.sum()
    assert covered[np.nonzero(live[want_order])[0]].all()


def force_brute(pos, r, live, N, dx, stencil=True):
    """O(n^2) float64 adhesion/repulsion with equal radii (so the order
    dependence of the carried threshold is inert), optionally restricted to
    cells in the 27 neighbouring voxels."""
    n = len(pos)
    p = pos.astype(np.float64)
    vox = np.trunc(pos / np.float32(dx)).astype(np.int64)
    f = np.zeros((n, 3))
    R = 2.0 * float(r)
    thres = R * 1.25
    for i in range(n):
        if not live[i]:
            continue
        d3 = p[i] - p
        dist = np.sqrt((d3 * d3).sum(1) + 1e-4)
        ok = live & (dist < thres)
        if stencil:
            ok &= np.all(np.abs(vox - vox[i]) <= 1, axis=1)
        dd = dist[ok]
        cr = np.where(dd <= R, 10.0 * (1 - dd / R) ** 2 / dd, 0.0)
        ca = 0.4 * (1 - dd / thres) ** 2 / dd
        f[i] = (d3[ok] * (cr - ca)[:, None]).sum(0)
    return f


@pytest.mark.parametrize("dx", [10.0, 25.0])
def test_force_equal_radii_matches_brute_force(dx, oracle):
    N, n = 16, 1500
    P = params(N, dx=dx, capacity=n)
    rng = np.random.default_rng(3)
    cells = mixed_cells(N, dx, n, n, seed=3, outside=0, dead_frac=0.1)
    c = np.float32(N * dx / 2)
    cells["pos"][:n] = (c + (rng.random((n, 3)) - 0.5) * np.float32(0.5 * N * dx)).astype(np.float32)
    cells["r"][:n] = 5.0
    work = cells.copy()
    keys, order, s, e = oracle.fm_update(P, work, n)
    # forces of the sorted state before movement: recompute on the sorted copy
    sorted_cells = cells[order].copy()
    oracle.fm_keys(P, sorted_cells, n)
    f = oracle.forces(P, sorted_cells, n, s, e)
    live = sorted_cells["sign"][:n] != 0
    want = force_brute(sorted_cells["pos"][:n], 5.0, live, N, dx, stencil=True)
    assert rel_l2(f[live], want[live]) < 1e-5
    # after movement a cell still inside carries that force in oldforce
    kept = work["sign"][:n] != 0
    assert np.array_equal(work["oldforce"][:n][kept], f[kept])
    if dx >= 12.5:  # every pair within the cutoff is inside the stencil
        want_all = force_brute(sorted_cells["pos"][:n], 5.0, live, N, dx, stencil=False)
        assert rel_l2(f[live], want_all[live]) < 1e-5
        # Newton's third law: the net force vanishes
        assert np.abs(f[live].astype(np.float64).sum(0)).max() < 1e-4 * np.abs(f[live]).sum()


def test_self_pair_contributes_nothing(oracle):
    N, dx = 8, 10.0
    P = params(N, dx=dx, capacity=4)
    cells = np.zeros(4, rt.CELL_DTYPE)
    cells["sign"][0] = 1
    cells["pos"][0] = (41.0, 41.0, 41.0)
    cells["r"][0] = 5.0
    cells["V"][0] = 523.0
    work = cells.copy()
    keys, order, s, e = oracle.fm_update(P, work, 1)
    assert np.all(work["oldforce"][0] == 0.0) and np.all(work["pos"][0] == cells["pos"][0])


def test_move_closed_form_and_domain_exit(oracle):
    N, dx = 8, 10.0
    P = params(N, dx=dx, capacity=8)
    cells = np.zeros(8, rt.CELL_DTYPE)
    cells["sign"][:5] = 1
    cells["pos"][:5] = [[40.0, 40.0, 40.0], [12.0, 33.0, 71.0], [0.2, 40.0, 40.0], [79.5, 40.0, 40.0],
                        [0.2, 40.0, 40.0]]
    cells["force"][:5] = [[1.0, -2.0, 0.5], [0.0, 3.0, -1.0], [-10.0, 0.0, 0.0], [20.0, 0.0, 0.0],
                          [-300.0, 0.0, 0.0]]
    cells["oldforce"][:5] = [[0.5, 0.5, 0.5], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]
    got = cells.copy()
    oracle.move(P, got, 5)
    h = np.float32(0.1) / np.float32(2.0)
    # cell 2 lands at x = -1.3: (int)(-0.13) truncates to voxel 0, so it stays
    # inside (BasicFunctions.cuh:60-66)
    for i in range(3):
        want = cells["pos"][i] + h * (np.float32(3.0) * cells["force"][i] - cells["oldforce"][i])
        assert np.array_equal(got["pos"][i], want.astype(np.float32))
        assert np.array_equal(got["oldforce"][i], cells["force"][i])
        assert np.all(got["force"][i] == 0.0) and got["sign"][i] == 1
        ijk = np.trunc(want / np.float32(dx)).astype(int)
        assert got["mesh_idx"][i] == ijk[0] + N * (ijk[1] + N * ijk[2])
    assert got["pos"][2][0] < 0
    for i in (3, 4):  # pushed out of the domain: flagged dead, state kept
        assert got["sign"][i] == 0
        assert np.array_equal(got["pos"][i], cells["pos"][i])
        assert np.array_equal(got["force"][i], cells["force"][i])
        assert np.array_equal(got["oldforce"][i], cells["oldforce"][i])
