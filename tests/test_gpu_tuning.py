"""Every kernel variant the autotuner may pick gives bit-identical results."""
import numpy as np
import pytest

from gell_amd import runtime as rt
from helpers import GpuState, compare_cells, mixed_cells, params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [64, 128, 192])
def test_lod_variants_identical(N):
    P = params(N)
    rng = np.random.default_rng(N)
    p = (20 + 18 * rng.random(N ** 3)).astype(np.float32)
    rate = (-2 * rng.random(N ** 3)).astype(np.float32)
    outs = []
    for v in (0, 1):
        st = GpuState(P, np.zeros(0, rt.CELL_DTYPE), exact=True)
        st.ctx.tune("lod_yz_lds", v)
        assert st.ctx.tune_get("lod_yz_lds") == v
        st.ctx.set_field(0, p)
        rt.upload_ptr(st.ctx.lib.gell_uptake_field(st.ctx.h, 0), rate)
        for _ in range(2):
            st.ctx.lod_substep(0)
        outs.append(st.ctx.field(0))
        st.close()
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("N,n", [(64, 20000), (64, 20001), (100, 30000), (13, 3000), (64, 140001), (192, 530001)])
def test_fm_variants_identical(N, n):
    """All force-kernel variants give the same bits; an odd count exercises
    the staged variant's 8-byte record tail, N = 100 Morton blocks that
    stick out of the grid (block variant), N = 13 occupancy words (aligned
    4^3 blocks) clipped by the grid's faces, and the mixed cells include dead
    and outside ones. Variant 10 (auto) runs 7 in all but the last case,
    where (>= 524288 cells, <= 1/8 cell per grid voxel) it runs 9."""
    P = params(N, capacity=n)
    cells = mixed_cells(N, 10.0, n, n, seed=4, dense=True)
    _all_fm_variants_equal(P, cells, n)


def _all_fm_variants_equal(P, cells, n, steps=3):
    outs = []
    for v in range(11):
        st = GpuState(P, cells)
        st.ctx.tune("fm_prefetch", v)
        for _ in range(steps):
            st.ctx.fm_update(st.buf, n)
        outs.append(st.cells())
        st.close()
    for v in range(1, len(outs)):
        assert np.array_equal(outs[0].view(np.uint8), outs[v].view(np.uint8)), f"variant {v}"


@pytest.mark.parametrize("n", [12000, 40000])
def test_fm_variants_identical_dense_hds(n, oracle):
    """The reference's HDS tissue (radius 8.4 um cells, ~4 per 25 um voxel):
    every variant gives the same bits, equal to the oracle's."""
    N = 60
    P = params(N, dx=25.0, capacity=n)
    cells = rt.sphere_init(P, n, n, 8.4, seed=n)
    _all_fm_variants_equal(P, cells, n)
    st = GpuState(P, cells)
    st.ctx.fm_update(st.buf, n)
    got = st.cells()
    st.close()
    want = cells.copy()
    oracle.fm_update(P, want, n)
    compare_cells(got, want, n)


def test_bad_knob_is_an_error():
    st = GpuState(params(16), np.zeros(0, rt.CELL_DTYPE))
    with pytest.raises(rt.GellError):
        st.ctx.tune("lod_yz_lds", 7)
    with pytest.raises(rt.GellError):
        st.ctx.tune("fm_prefetch", 11)
    with pytest.raises(KeyError):
        st.ctx.tune("nope", 1)
    st.close()


@pytest.mark.parametrize("N,xp,yzp", [(64, 2, 2), (64, 4, 4), (96, 4, 2), (100, 2, 2), (100, 4, 4), (100, 4, 1),
                                      (100, 1, 4), (100, 2, 4),
                                      (128, 2, 2), (128, 4, 4), (192, 2, 2), (256, 2, 1), (256, 1, 2),
                                      (256, 4, 4), (256, 2, 4), (256, 8, 16), (512, 4, 4), (512, 8, 8),
                                      (512, 8, 16), (512, 1, 1)])
def test_lod_partitioned_within_tolerance(N, xp, yzp, oracle):
    """Partitioned line solves reproduce the serial sweep within 1e-6
    relative L2 per substep (the stated bar for fields is 1e-5) and give the
    same bits when repeated."""
    from helpers import rel_l2

    P = params(N)
    rng = np.random.default_rng(N + xp + 10 * yzp)
    p = (20 + 18 * rng.random(N ** 3)).astype(np.float32)
    rate = (-2 * rng.random(N ** 3) * (rng.random(N ** 3) < 0.3)).astype(np.float32)
    st = GpuState(P, np.zeros(0, rt.CELL_DTYPE))
    st.ctx.tune("lod_fused", 0)
    st.ctx.tune("lod_x_parts", xp)
    st.ctx.tune("lod_parts", yzp)
    outs = []
    for _ in range(2):
        st.ctx.set_field(0, p)
        rt.upload_ptr(st.ctx.lib.gell_uptake_field(st.ctx.h, 0), rate)
        st.ctx.lod_substep(0)
        outs.append(st.ctx.field(0))
    st.close()
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    want = p.copy()
    oracle.lod_substep(P, 0, want, rate)
    e = rel_l2(outs[0], want)
    assert e < 1e-6, e
    assert np.max(np.abs(outs[0] - want)) < 1e-4


@pytest.mark.parametrize("N", [96, 100, 256])
def test_autotune_keeps_state_and_picks_valid_variant(N, oracle):
    from helpers import rel_l2

    P = params(N)
    rng = np.random.default_rng(5)
    p = (20 + 18 * rng.random(N ** 3)).astype(np.float32)
    rate = (-2 * rng.random(N ** 3)).astype(np.float32)
    st = GpuState(P, np.zeros(0, rt.CELL_DTYPE))
    st.ctx.set_field(0, p)
    rt.upload_ptr(st.ctx.lib.gell_uptake_field(st.ctx.h, 0), rate)
    exact = st.ctx.autotune(bit_exact=True)
    assert exact["lod_parts"] == 1 and exact["lod_x_parts"] == 1 and exact["lod_us"] > 0
    assert np.array_equal(st.ctx.field(0).view(np.uint32), p.view(np.uint32))
    best = st.ctx.autotune()
    assert np.array_equal(st.ctx.field(0).view(np.uint32), p.view(np.uint32))
    assert best["lod_us"] <= exact["lod_us"] * 1.05
    st.ctx.lod_substep(0)
    got = st.ctx.field(0)
    st.close()
    want = p.copy()
    oracle.lod_substep(P, 0, want, rate)
    assert rel_l2(got, want) < 1e-6


@pytest.mark.parametrize("N,zp", [(128, 2), (128, 4), (256, 4), (256, 2), (256, 8), (512, 8), (512, 16), (512, 4)])
def test_lod_fused(N, zp, oracle):
    """The fused substep (x + y blocks in one pass, y coupling applied in the
    z sweep) gives the bits of the unfused partitioned kernels with x and y in
    N / 64 blocks and z in zp blocks when zp == N / 64, and stays within the
    partitioned tolerance of the serial sweep for every zp."""
    from helpers import rel_l2

    P = params(N)
    rng = np.random.default_rng(N + zp)
    p = (20 + 18 * rng.random(N ** 3)).astype(np.float32)
    rate = (-2 * rng.random(N ** 3) * (rng.random(N ** 3) < 0.3)).astype(np.float32)
    st = GpuState(P, np.zeros(0, rt.CELL_DTYPE))
    outs = {}
    for fused in (1, 0):
        st.ctx.tune("lod_fused", fused)
        st.ctx.tune("lod_x_parts", N // 64)
        st.ctx.tune("lod_parts", zp if fused else N // 64)
        for rep in range(2 if fused else 1):
            st.ctx.set_field(0, p)
            rt.upload_ptr(st.ctx.lib.gell_uptake_field(st.ctx.h, 0), rate)
            st.ctx.lod_substep(0)
            outs[(fused, rep)] = st.ctx.field(0)
    st.close()
    got = outs[(1, 0)]
    assert np.array_equal(got.view(np.uint32), outs[(1, 1)].view(np.uint32))
    if zp == N // 64:
        assert np.array_equal(got.view(np.uint32), outs[(0, 0)].view(np.uint32))
    want = p.copy()
    oracle.lod_substep(P, 0, want, rate)
    e = rel_l2(got, want)
    assert e < 1e-6, e
    assert np.max(np.abs(got - want)) < 1e-4


def test_substrate_streams_give_identical_results():
    """Substrate 1's substeps on their own stream (overlapping substrate 0's)
    leave both fields bit-identical to running everything on one stream,
    over mechanics steps with a moving spheroid (uptake maps rebuilt each step)."""
    from helpers import mixed_cells

    N, n = 128, 20_000
    P = params(N, capacity=n, n_sub=2)
    cells = mixed_cells(N, 10.0, n, n, seed=9, outside=0)
    outs = []
    for streams in (1, 0):
        st = GpuState(P, cells)
        st.ctx.tune("substrate_streams", streams)
        for _ in range(3):
            st.ctx.fm_update(st.buf, n)
            for s in range(2):
                st.ctx.prelod(s, st.buf, n)
            for s in range(2):
                for _ in range(11):
                    st.ctx.lod_substep(s)
        outs.append([st.ctx.field(s) for s in range(2)] + [st.cells()])
        st.close()
    for a, b in zip(outs[0][:2], outs[1][:2]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(outs[0][2].view(np.uint8), outs[1][2].view(np.uint8))


def test_rate_spans_rebuilt_after_prelod_equal_spans_rebuilt_every_substep(oracle):
    """The partitioned x sweep reads uptake rates only inside each x line's
    span of voxels that may hold a nonzero rate, rebuilt from the map by the
    first substep after each PreLOD. Over several mechanics steps (cells move,
    the previous map is cleared) this must give the same bits as spans
    rebuilt before every substep (handing out the uptake-map pointer forces
    the rebuild), and the field must stay within the 1e-5 relative-L2 bar of
    the oracle (a skipped nonzero rate would break it; the partitioned
    solve's rounding, under 1e-6 per substep, accumulates over 66 substeps)."""
    from gell_amd import spheroid
    from helpers import rel_l2

    N, n = 128, 40000
    P = params(N, capacity=n)
    cells = spheroid.make_cells(n, N, 10.0, 420.0, 11, capacity=n)
    outs = []
    for rebuild in (False, True):
        st = GpuState(P, cells)
        assert st.ctx.tune_get("lod_fused") == 1 and st.ctx.tune_get("lod_parts") > 1
        for _ in range(6):
            st.ctx.fm_update(st.buf, n)
            st.ctx.prelod(0, st.buf, n)
            for _ in range(11):
                if rebuild:
                    st.ctx.lib.gell_uptake_field(st.ctx.h, 0)
                st.ctx.lod_substep(0)
        outs.append((st.cells()[:n].copy(), st.ctx.field(0)))
        st.close()
    assert outs[0][0].tobytes() == outs[1][0].tobytes()
    assert outs[0][1].tobytes() == outs[1][1].tobytes()
    want = cells.copy()
    f = np.full(N ** 3, 38.0, np.float32)
    oracle.run_mech(P, want, n, f, 6)
    assert rel_l2(outs[0][1], f) < 1e-5
