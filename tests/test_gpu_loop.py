"""The whole time loop (Main.cu:108-127) on the GPU against the oracle's:
birth, 61 mechanics steps each with its uptake map and 11 diffusion substeps
per substrate, death -- for two biology steps at 32^3 with two substrates."""
import numpy as np
import pytest

from gell_amd import runtime as rt
from helpers import GpuState, log, mixed_cells, params, rel_l2

pytestmark = pytest.mark.gpu


def run_gpu(P, cells, n, bio_steps):
    st = GpuState(P, cells)
    mech = rt.float_loop_count(P.dt_bio, P.dt_mech)
    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
    for b in range(bio_steps):
        n = st.ctx.birth(st.buf, n, b)
        for _ in range(mech):
            st.ctx.fm_update(st.buf, n)
            for s in range(P.n_substrates):
                st.ctx.prelod(s, st.buf, n)
            for _ in range(diff):
                for s in range(P.n_substrates):
                    st.ctx.lod_substep(s)
        n = st.ctx.death(st.buf, n)
    out = st.cells(), [st.ctx.field(s) for s in range(P.n_substrates)], n
    st.close()
    return out


def test_two_bio_steps_match_oracle(oracle):
    N, n = 32, 3000
    P = params(N, capacity=2 * n, n_sub=2, seed=2024)
    cells = mixed_cells(N, 10.0, n, P.capacity, seed=77, dead_frac=0.0, outside=0)
    cells["phase"][:n] = np.where(np.arange(n) % 5 == 0, 1, 0)
    cells["cell_clock"][:n] = np.where(np.arange(n) % 5 == 0, 774.0, 0.0).astype(np.float32)
    got, fields, n_gpu = run_gpu(P, cells, n, 2)
    want = cells.copy()
    f = np.concatenate([np.full(N ** 3, P.sub[s].initial, np.float32) for s in range(P.n_substrates)])
    n_cpu = oracle.run(P, want, n, f, 2)
    log(f"[loop] n {n} -> gpu {n_gpu} cpu {n_cpu}")
    assert n_gpu == n_cpu
    for s in range(P.n_substrates):
        e = rel_l2(fields[s], f[s * N ** 3:(s + 1) * N ** 3])
        log(f"[loop] field {s} relL2 {e:.2e}")
        assert e <= 1e-5
    assert np.array_equal(got["sign"][:n_cpu], want["sign"][:n_cpu])
    assert np.array_equal(got["phase"][:n_cpu], want["phase"][:n_cpu])
    e = rel_l2(got["pos"][:n_cpu], want["pos"][:n_cpu])
    log(f"[loop] pos relL2 {e:.2e}")
    assert e <= 1e-5
