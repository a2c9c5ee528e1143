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


# ---------------------------------------------------------------------------
# Carried threshold with unequal radii (FMkernel.cuh:72, 91, 98-99)
# ---------------------------------------------------------------------------
# forcekernel keeps `thres` from the last pair that passed the first distance
# test: it starts at Tumor_CC_Max_Dist (25), each such pair sets it to
# 1.25 (r_i + r_j), and the first test of the next candidate (:91) compares
# against that carried value before the candidate's own radii are looked at.
# The stencil is walked x outer, z inner, slots ascending, and the self pair
# (distance 0.01) is visited like any other. The two clusters below are laid
# out so that the carried value decides one pair in each; the accepted sets
# are worked out by hand in the comments, and the forces are then summed in
# float64 from the pair formula (FMkernel.cuh:101-109).
#
# Cluster 1 (around voxel (3, 3, 3)):
#   A r=5.0 (35, 35, 35)     voxel (3,3,3)
#   B r=5.0 (35, 35, 46.5)   voxel (3,3,4): 11.5 from A
#   C r=5.5 (47.8, 35, 35)   voxel (4,3,3): 12.8 from A, 17.2 from B
#   A's walk: self (stencil index 13) -> thres 12.5; B (index 14) at 11.5 is
#     accepted -> thres 12.5; C (index 22) at 12.8 >= 12.5 is rejected at :91.
#     With its own radii C would have thres 13.125 > 12.8: the carried value
#     decides.  A <- {B}.
#   B's walk: A (index 12) at 11.5 < 25 -> thres 12.5, accepted; self; C far.
#     B <- {A}.
#   C's walk: A (index 4) at 12.8 < 25 -> thres 13.125 > 12.8, accepted; self.
#     C <- {A}: the pair is accepted one way and not the other.
# Cluster 2 (around voxel (11, 11, 11)):
#   A2 r=5.5 (115, 115, 115) voxel (11,11,11)
#   F  r=4.5 (115, 127, 115) voxel (11,12,11): 12.0 from A2
#   G  r=5.5 (128, 115, 115) voxel (12,11,11): 13.0 from A2, 17.7 from F
#   A2's walk: self -> thres 13.75; F (index 16) at 12.0 < 13.75 passes :91,
#     its own thres is 12.5 > 12.0: accepted, thres 12.5; G (index 22) at
#     13.0 >= 12.5 is rejected at :91 (its own thres would be 13.75).
#     A2 <- {F}.
#   F's walk: A2 (index 10) at 12.0 -> thres 12.5, accepted; self; G far.
#   G's walk: A2 (index 4) at 13.0 < 25 -> thres 13.75, accepted; F (index 7)
#     at 17.7 rejected; self.  G <- {A2}.
CARRY_CELLS = [  # name, r, position
    ("A", 5.0, (35.0, 35.0, 35.0)),
    ("B", 5.0, (35.0, 35.0, 46.5)),
    ("C", 5.5, (47.8, 35.0, 35.0)),
    ("A2", 5.5, (115.0, 115.0, 115.0)),
    ("F", 4.5, (115.0, 127.0, 115.0)),
    ("G", 5.5, (128.0, 115.0, 115.0)),
]
CARRY_ACCEPTED = {"A": ["B"], "B": ["A"], "C": ["A"], "A2": ["F"], "F": ["A2"], "G": ["A2"]}
# what an implementation that tests every candidate against its own radii
# only would accept instead
NO_CARRY_ACCEPTED = {"A": ["B", "C"], "B": ["A"], "C": ["A"], "A2": ["F", "G"], "F": ["A2"], "G": ["A2"]}


def carry_case():
    """Cells of the two clusters (N = 16, dx = 10), unsorted on purpose."""
    N, dx = 16, 10.0
    order = [4, 2, 0, 5, 3, 1]
    cells = np.zeros(len(CARRY_CELLS), rt.CELL_DTYPE)
    for slot, q in enumerate(order):
        _, r, p = CARRY_CELLS[q]
        cells["sign"][slot] = 1
        cells["pos"][slot] = p
        cells["r"][slot] = r
        cells["V"][slot] = 2494.0
    return N, dx, cells


def pair_force64(p1, r1, p2, r2):
    d3 = np.asarray(p1, np.float64) - np.asarray(p2, np.float64)
    d = np.sqrt((d3 * d3).sum() + 1e-4)
    R = max(r1 + r2, 0.01)
    thres = 1.25 * R
    assert d < thres
    cr = 10.0 * (1 - d / R) ** 2 / d if d <= R else 0.0
    ca = 0.4 * (1 - d / thres) ** 2 / d
    return d3 * (cr - ca)


def carry_expected(cells, accepted):
    byname = {}
    for slot in range(len(cells)):
        p = tuple(float(x) for x in cells["pos"][slot])
        name = next(nm for nm, _, q in CARRY_CELLS if np.allclose(q, p))
        byname[name] = slot
    f = np.zeros((len(cells), 3))
    for name, others in accepted.items():
        i = byname[name]
        for o in others:
            j = byname[o]
            f[i] += pair_force64(cells["pos"][i], float(cells["r"][i]), cells["pos"][j], float(cells["r"][j]))
    return f


def test_force_carried_threshold_unequal_radii(oracle):
    N, dx, cells = carry_case()
    n = len(cells)
    P = params(N, dx=dx, capacity=n)
    work = cells.copy()
    keys, order, s, e = oracle.fm_update(P, work, n)
    sorted_cells = cells[order].copy()
    oracle.fm_keys(P, sorted_cells, n)
    f = oracle.forces(P, sorted_cells, n, s, e).astype(np.float64)
    want = carry_expected(sorted_cells, CARRY_ACCEPTED)
    alt = carry_expected(sorted_cells, NO_CARRY_ACCEPTED)
    # near the cutoff (1 - d / thres) cancels: float32 rounding leaves a few
    # 1e-6 relative; leaving out or adding one pair moves a force by > 9 %
    for i in range(n):
        assert np.linalg.norm(f[i] - want[i]) <= 2e-5 * np.linalg.norm(want[i]), (i, f[i], want[i])
    moved = [i for i in range(n) if np.linalg.norm(alt[i] - want[i]) > 0.05 * np.linalg.norm(want[i])]
    assert len(moved) == 2
    # the step stores the force in oldforce (every cell stays inside)
    assert np.all(work["sign"][:n] == 1)
    assert np.array_equal(work["oldforce"][:n], f.astype(np.float32))
