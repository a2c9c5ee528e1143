"""Mechanics step on the GPU against the oracle (FM_update, FMkernel.cuh:181-202).

From a shared input state: Morton keys, sort order and per-voxel ranges must
be bit-exact; records after force + movement must match (integer fields
exactly, float fields within 1e-5 relative L2 and expected bit-exact)."""
import numpy as np
import pytest

from gell_amd import runtime as rt
from helpers import GpuState, compare_cells, log, mixed_cells, params

pytestmark = pytest.mark.gpu


def _check_step(st: GpuState, cells: np.ndarray, n: int, oracle, label: str):
    P = st.P
    st.upload(cells)
    st.ctx.fm_update(st.buf, n)
    got = st.cells()
    keys = st.ctx.sorted_keys(n)
    order = st.ctx.sorted_order(n)
    s, e = st.ctx.voxel_ranges()
    want = cells.copy()
    wk, wo, ws, we = oracle.fm_update(P, want, n)
    assert np.array_equal(keys, wk), f"{label}: keys"
    assert np.array_equal(order, wo), f"{label}: order"
    assert np.array_equal(s, ws), f"{label}: s"
    assert np.array_equal(e, we), f"{label}: e"
    rep = compare_cells(got, want, n)
    log(f"[fm] {label}: n={n} " + " ".join(f"{k}={v[0]:.1e}/{v[1]}" for k, v in rep.items()))
    return want


@pytest.mark.parametrize("N,n,dense", [(32, 2000, True), (64, 10000, False), (64, 6000, True), (64, 16384, False),
                                       (64, 16385, False), (96, 20000, True)])
def test_fm_steps_match_oracle(N, n, dense, oracle):
    log(f"[fm] N={N} n={n} dense={dense}")
    P = params(N, capacity=n + 64)
    cells = mixed_cells(N, 10.0, n, P.capacity, seed=N + n, dense=dense)
    st = GpuState(P, cells)
    state = cells
    for step in range(4):
        # each step starts from the oracle's state of the previous one
        state = _check_step(st, state, n, oracle, f"N={N} step {step}")
    st.close()


@pytest.mark.parametrize("N", [13, 30])
def test_fm_faces_of_odd_grids_match_oracle(N, oracle):
    """Grid sides that are not multiples of 4 (the occupancy words cover
    aligned 4^3 blocks, which then stick out of the grid), with a third of the
    cells in the first or last voxel layer of some axis, where the stencil
    is clipped."""
    n = 3000
    log(f"[fm] faces N={N}")
    P = params(N, capacity=n)
    cells = mixed_cells(N, 10.0, n, P.capacity, seed=N, dense=True)
    rng = np.random.default_rng(N)
    k = n // 3
    axis, high = rng.integers(0, 3, k), rng.integers(0, 2, k) == 1
    layer = np.where(high, rng.uniform((N - 1) * 10.0, N * 10.0 - 0.01, k), rng.uniform(0.0, 10.0, k))
    cells["pos"][np.arange(k), axis] = layer.astype(np.float32)
    st = GpuState(P, cells)
    state = cells
    for step in range(3):
        state = _check_step(st, state, n, oracle, f"N={N} faces step {step}")
    st.close()


def test_fm_chained_steps_stay_identical(oracle):
    # GPU state is never re-synchronised: ten steps back to back.
    N, n = 48, 5000
    P = params(N, capacity=n)
    cells = mixed_cells(N, 10.0, n, P.capacity, seed=5, dense=True)
    st = GpuState(P, cells)
    want = cells.copy()
    for _ in range(10):
        st.ctx.fm_update(st.buf, n)
        oracle.fm_update(P, want, n)
    compare_cells(st.cells(), want, n)
    st.close()


def test_fm_large_sort_256(oracle):
    # 256^3 grid, 25-bit biased keys: three radix passes of 9/8/8 bits.
    N, n = 256, 200_000
    log(f"[fm] N={N} n={n}")
    P = params(N, capacity=n)
    cells = mixed_cells(N, 10.0, n, P.capacity, seed=11, radius=1200.0)
    st = GpuState(P, cells)
    _check_step(st, cells, n, oracle, "N=256")
    st.close()


def test_fm_edge_cases(oracle):
    N = 16
    P = params(N, capacity=64)
    # n = 0
    st = GpuState(P, np.zeros(0, rt.CELL_DTYPE))
    st.ctx.fm_update(st.buf, 0)
    st.ctx.sync()
    # one cell; all dead; everything outside
    for kind in ("one", "dead", "outside"):
        cells = mixed_cells(N, 10.0, 40, 64, seed=3, outside=0)
        n = 1 if kind == "one" else 40
        if kind == "dead":
            cells["sign"][:n] = 0
        if kind == "outside":
            cells["pos"][:n, 2] = -50.0
        _check_step(st, cells, n, oracle, kind)
    st.close()


def test_fm_bad_count_is_an_error():
    P = params(16, capacity=8)
    st = GpuState(P, np.zeros(0, rt.CELL_DTYPE))
    with pytest.raises(rt.GellError):
        st.ctx.fm_update(st.buf, 9)
    st.close()


def test_fm_carried_threshold_unequal_radii(oracle):
    """The hand-built clusters whose accepted pairs the carried threshold
    decides (FMkernel.cuh:72, 91, 98-99): every force kernel variant against
    the oracle bit for bit and against the hand-derived float64 forces."""
    from test_oracle_fm import CARRY_ACCEPTED, carry_case, carry_expected

    N, dx, cells = carry_case()
    n = len(cells)
    P = params(N, dx=dx, capacity=n)
    want = cells.copy()
    _, order, _, _ = oracle.fm_update(P, want, n)
    sorted_in = cells[order].copy()
    oracle.fm_keys(P, sorted_in, n)
    hand = carry_expected(sorted_in, CARRY_ACCEPTED)
    st = GpuState(P, cells)
    for v in range(10):
        log(f"[fm] carried thres, force variant {v}")
        st.ctx.tune("fm_prefetch", v)
        st.upload(cells)
        st.ctx.fm_update(st.buf, n)
        got = st.cells()
        compare_cells(got, want, n)
        f = got["oldforce"][:n].astype(np.float64)
        for i in range(n):
            assert np.linalg.norm(f[i] - hand[i]) <= 2e-5 * np.linalg.norm(hand[i]), (v, i)
    st.close()


@pytest.mark.parametrize("N,n,dense", [(30, 140000, True), (64, 140000, False), (96, 150000, True),
                                       (128, 131072, False)])
def test_fm_rank_list_matches_oracle(N, n, dense, oracle):
    """Force variant 9 keeps the rank-compacted cell list instead of the
    Morton-indexed ranges (from 131072 cells on); gell_voxel_ranges rebuilds
    the reference's s/e from it. Keys, order, ranges and records against the
    oracle, with the variant switched to 6 and back between steps (each
    clears what the other left). N = 30: occupancy words clipped by the
    grid's faces, ~5 cells per voxel."""
    P = params(N, capacity=n + 64)
    cells = mixed_cells(N, 10.0, n, P.capacity, seed=N + n + 1, dense=dense)
    st = GpuState(P, cells)
    state = cells
    for step, v in enumerate((9, 9, 6, 9)):
        st.ctx.tune("fm_prefetch", v)
        state = _check_step(st, state, n, oracle, f"N={N} variant {v} step {step}")
    st.close()
