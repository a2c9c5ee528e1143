"""gell_mech_step (the mechanics step with every substrate's uptake map
accumulated inside its force/movement kernel) against the two calls it
stands for, gell_fm_update followed by gell_prelod per substrate
(FMkernel.cuh:181-202, LODkernel.cuh:270-278): cells, uptake maps and fields
after the substeps bit for bit, over several steps, for every force variant
(7-9 fuse, the others fall back to the separate calls), with the two paths
mixed and with a map the caller was handed in between; and the maps against
the oracle."""
import numpy as np
import pytest

from gell_amd import runtime as rt
from gell_amd import spheroid
from helpers import GpuState, log, mixed_cells, params

pytestmark = pytest.mark.gpu


def _run(P, cells, n, variant, plan, substeps=2):
    """plan: per step 'f' (fused), 'u' (unfused) or 'h' (fused after handing
    out the maps and writing one of them)."""
    st = GpuState(P, cells)
    st.ctx.tune("fm_prefetch", variant)
    for k, how in enumerate(plan):
        if how == "u":
            st.ctx.fm_update(st.buf, n)
            for s in range(P.n_substrates):
                st.ctx.prelod(s, st.buf, n)
        else:
            if how == "h":
                m = st.ctx.uptake_field(0)
                m[: len(m) // 8] = -1.0
                rt.upload_ptr(st.ctx.lib.gell_uptake_field(st.ctx.h, 0), m)
            st.ctx.mech_step(st.buf, n)
        for s in range(P.n_substrates):
            for _ in range(substeps):
                st.ctx.lod_substep(s)
    cells_out = st.cells()[:n].copy()
    maps = [st.ctx.uptake_field(s) for s in range(P.n_substrates)]
    fields = [st.ctx.field(s) for s in range(P.n_substrates)]
    st.close()
    return cells_out, maps, fields


@pytest.mark.parametrize("N,n,dense", [(64, 20000, False), (48, 15000, True), (128, 140001, False)])
@pytest.mark.parametrize("variant", [6, 7, 8, 9])
def test_mech_step_equals_fm_update_and_prelod(N, n, dense, variant, oracle):
    log(f"[mech] N={N} n={n} dense={dense} variant={variant}")
    P = params(N, capacity=n, n_sub=2)
    cells = mixed_cells(N, 10.0, n, n, seed=N + n, dense=dense)
    want = _run(P, cells, n, variant, "uuuu")
    for plan in ("ffff", "ufuf", "fhff"):
        got = _run(P, cells, n, variant, plan)
        assert np.array_equal(got[0].view(np.uint8), want[0].view(np.uint8)), plan
        # ('h' edits a map just before a step rewrites it: nothing downstream sees the edit)
        for s in range(2):
            assert np.array_equal(got[1][s].view(np.uint32), want[1][s].view(np.uint32)), (plan, s)
            assert np.array_equal(got[2][s].view(np.uint32), want[2][s].view(np.uint32)), (plan, s)
            # the last map against the oracle on the final cells, whatever the path
            ref = oracle.prelod(P, s, got[0], n)
            assert np.array_equal(got[1][s].view(np.uint32), ref.view(np.uint32)), (plan, s)


def test_mech_step_after_caller_wrote_the_map(oracle):
    """A handed-out map the caller wrote, then fused steps: each new map is
    the oracle's (the whole map is cleared, accumulators untouched)."""
    N, n = 128, 30000
    P = params(N, capacity=n, n_sub=2)
    cells = spheroid.make_cells(n, N, 10.0, 400.0, 9, capacity=n)
    st = GpuState(P, cells)
    st.ctx.tune("fm_prefetch", 9)
    for step in range(4):
        if step == 2:
            m = st.ctx.uptake_field(1).reshape(N, N, N)
            m[N // 2 - 8:N // 2 + 8] = 0.0
            m[:3] = 2.0
            rt.upload_ptr(st.ctx.lib.gell_uptake_field(st.ctx.h, 1), m.ravel())
        st.ctx.mech_step(st.buf, n)
        got = st.cells()
        for s in range(2):
            ref = oracle.prelod(P, s, got, n)
            assert np.array_equal(st.ctx.uptake_field(s).view(np.uint32), ref.view(np.uint32)), (step, s)
        for s in range(2):
            st.ctx.lod_substep(s)
    st.close()
