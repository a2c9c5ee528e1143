"""Uptake map, birth and death on the GPU against the oracle
(PreLOD LODkernel.cuh:270-278; CellBirth_kernel / CellDeath_kernel
PDkernel.cuh:27-117).

Uptake maps are bit-exact (integer fixed-point sums). Birth events (which
mothers divide, which slot each daughter gets, the new count) are bit-exact;
records match exactly except the division geometry, whose sinf/cosf/acosf
may differ by ulps between the device and host maths libraries."""
import numpy as np
import pytest

from gell_amd import runtime as rt
from helpers import GpuState, compare_cells, log, mixed_cells, params, rel_l2, ulp_diff

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_sub", [1, 2])
def test_prelod_matches_oracle(n_sub, oracle):
    N, n = 32, 8000
    P = params(N, capacity=n, n_sub=n_sub)
    cells = mixed_cells(N, 10.0, n, n, seed=21, dense=True)
    st = GpuState(P, cells)
    for step in range(3):
        st.ctx.fm_update(st.buf, n)
        for s in range(n_sub):
            st.ctx.prelod(s, st.buf, n)
        got_cells = st.cells()
        for s in range(n_sub):
            want = oracle.prelod(P, s, got_cells, n)
            got = st.ctx.uptake_field(s)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"sub {s} step {step}"
    log(f"[prelod] n_sub={n_sub} ok")
    st.close()


def test_prelod_span_clear_matches_oracle(oracle):
    """From the second mechanics step on, PreLOD clears the previous map
    along its x-line spans (rebuilt by the substeps in between) instead of
    voxel by voxel. After four steps of a moving spheroid the maps equal the
    oracle's bit for bit (a voxel left uncleared would show), and cells and
    fields equal those of a run that clears the whole map (handing out the
    map pointer before each PreLOD makes the next PreLOD clear all of it)."""
    from gell_amd import spheroid

    N, n = 128, 30000
    P = params(N, capacity=n, n_sub=2)
    cells = spheroid.make_cells(n, N, 10.0, 400.0, 3, capacity=n)
    outs = []
    for listed in (False, True):
        st = GpuState(P, cells)
        for _ in range(4):
            st.ctx.fm_update(st.buf, n)
            for s in range(2):
                if listed:
                    st.ctx.lib.gell_uptake_field(st.ctx.h, s)
                st.ctx.prelod(s, st.buf, n)
            for s in range(2):
                for _ in range(3):
                    st.ctx.lod_substep(s)
        got_cells = st.cells()[:n].copy()
        maps = [st.ctx.uptake_field(s) for s in range(2)]
        fields = [st.ctx.field(s) for s in range(2)]
        st.close()
        for s in range(2):
            want = oracle.prelod(P, s, got_cells, n)
            assert np.array_equal(maps[s].view(np.uint32), want.view(np.uint32)), f"listed={listed} sub {s}"
        outs.append([got_cells.tobytes()] + [f.tobytes() for f in fields])
    assert outs[0] == outs[1]
    log("[prelod] span clear ok")


def _birth_state(N, n, capacity, seed):
    cells = mixed_cells(N, 10.0, n, capacity, seed=seed, outside=0)
    rng = np.random.default_rng(seed)
    live = cells[:n]
    # phase 1 near the end of its 780-min window divides this step
    ph = rng.choice([0, 1, 1, 2, 3, 4, 5], n).astype(np.int32)
    live["phase"] = ph
    clk = np.where(ph == 1, 780.0 - 6.0 * rng.integers(0, 2, n), rng.random(n) * 150.0)
    live["cell_clock"] = clk.astype(np.float32)
    live["Vcs"] = np.where(rng.random(n) < 0.1, 20.0, live["Vcs"]).astype(np.float32)
    p0 = (rng.random(N ** 3) * 40.0).astype(np.float32)
    return cells, p0


@pytest.mark.parametrize("edit", ["zero_region", "write_outside"])
def test_prelod_after_caller_wrote_the_map(edit, oracle):
    """The caller may write the uptake map through the pointer
    gell_uptake_field hands out (the reference's public p_rate). Either it
    zeroes a slab that holds tissue between a PreLOD and the substeps (which
    then rebuild the x-line spans from the edited map; the spans no longer
    cover every nonzero accumulator), or it writes rates into voxels that
    hold no cell after the substeps (outside the previous voxel list). The
    next PreLOD must still give the oracle's map bit for bit: it clears the
    whole map, as the reference's oxy_consumption_reset (LODkernel.cuh:272)
    always does."""
    from gell_amd import spheroid

    N, n = 128, 30000
    P = params(N, capacity=n, n_sub=1)
    cells = spheroid.make_cells(n, N, 10.0, 400.0, 5, capacity=n)
    st = GpuState(P, cells)

    def write_map(fn):
        m = st.ctx.uptake_field(0).reshape(N, N, N)
        fn(m)
        rt.upload_ptr(st.ctx.lib.gell_uptake_field(st.ctx.h, 0), m.ravel())

    for step in range(2):
        st.ctx.fm_update(st.buf, n)
        st.ctx.prelod(0, st.buf, n)
        if step == 1 and edit == "zero_region":
            write_map(lambda m: m.__setitem__(slice(N // 2 - 8, N // 2 + 8), 0.0))
        for _ in range(3):
            st.ctx.lod_substep(0)
        if step == 1 and edit == "write_outside":
            write_map(lambda m: m.__setitem__(slice(0, 4), -3.0))  # z < 4: far outside the ball
    st.ctx.fm_update(st.buf, n)
    st.ctx.prelod(0, st.buf, n)
    got_cells = st.cells()
    want = oracle.prelod(P, 0, got_cells, n)
    assert np.array_equal(st.ctx.uptake_field(0).view(np.uint32), want.view(np.uint32)), edit
    st.close()


@pytest.mark.parametrize("capacity_room", [None, 37])
def test_birth_matches_oracle(capacity_room, oracle):
    N, n = 32, 5000
    cap = 2 * n if capacity_room is None else n + capacity_room
    P = params(N, capacity=cap, seed=987654321)
    cells, p0 = _birth_state(N, n, cap, seed=8)
    st = GpuState(P, cells)
    st.ctx.set_field(0, p0)
    n_gpu = st.ctx.birth(st.buf, n, 17)
    got = st.cells()
    slots = st.ctx.birth_slots(n)
    want = cells.copy()
    n_cpu, wslots = oracle.birth(P, p0, want, n, 17)
    log(f"[birth] cap={cap} n {n} -> gpu {n_gpu} cpu {n_cpu}")
    assert n_gpu == n_cpu
    assert np.array_equal(slots, wslots)
    assert n_cpu > n
    compare_cells(got, want, n_cpu, exact=False, skip=("pos",))
    # all non-geometry fields bit-exact
    for f in ("r", "V", "Vf", "Vcs", "Vns", "cell_clock", "force", "oldforce"):
        assert ulp_diff(got[f][:n_cpu], want[f][:n_cpu]) == 0, f
    div = np.concatenate([np.nonzero(wslots >= 0)[0], wslots[wslots >= 0]])
    same = np.setdiff1d(np.arange(n_cpu), div)
    assert ulp_diff(got["pos"][same], want["pos"][same]) == 0
    assert rel_l2(got["pos"][:n_cpu], want["pos"][:n_cpu]) <= 1e-6
    st.close()


def test_birth_full_capacity_does_nothing(oracle):
    N, n = 16, 100
    P = params(N, capacity=n)
    cells, p0 = _birth_state(N, n, n, seed=9)
    st = GpuState(P, cells)
    st.ctx.set_field(0, p0)
    assert st.ctx.birth(st.buf, n, 3) == n
    got = st.cells()
    assert np.array_equal(got.view(np.uint8), cells.view(np.uint8))
    assert np.all(st.ctx.birth_slots(n) == -1)
    st.close()


@pytest.mark.parametrize("n", [0, 1, 777, 100_000])
def test_death_matches_oracle(n, oracle):
    N = 32
    P = params(N, capacity=max(n, 1))
    cells = mixed_cells(N, 10.0, n, max(n, 1), seed=n + 1, dead_frac=0.4, outside=0)
    st = GpuState(P, cells)
    got_n = st.ctx.death(st.buf, n)
    got = st.cells()
    want = cells.copy()
    want_n = oracle.death(want, n)
    assert got_n == want_n == int(cells["sign"][:n].sum())
    assert np.array_equal(got[:got_n].view(np.uint8), want[:want_n].view(np.uint8))
    st.close()
