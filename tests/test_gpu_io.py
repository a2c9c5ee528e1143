"""Restart and the reference's example run, on the GPU.

* A checkpoint (gell_checkpoint_save / _load) holds the whole state a step
  depends on: a run restarted from it continues bit for bit as the
  uninterrupted run, and a checkpoint of another grid is refused.
* The HDS example (Definitions.h as shipped: 100^3 voxels of 25 um, 2347
  cells of radius 8.4 um, Main.cu's loop) run by the C++ driver for one
  biology step with the bit-exact kernel variants writes the reference's
  files, and its cells and field equal the oracle's after the same step --
  as files, byte for byte."""
import subprocess

import numpy as np
import pytest

from gell_amd import build as hip_build
from gell_amd import runtime as rt
from helpers import GpuState, log, mixed_cells, params

pytestmark = pytest.mark.gpu


def bio_steps(st, n, first, count):
    P = st.P
    mech = rt.float_loop_count(P.dt_bio, P.dt_mech)
    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
    for b in range(first, first + count):
        n = st.ctx.birth(st.buf, n, b)
        for _ in range(mech):
            st.ctx.fm_update(st.buf, n)
            for s in range(P.n_substrates):
                st.ctx.prelod(s, st.buf, n)
            for _ in range(diff):
                for s in range(P.n_substrates):
                    st.ctx.lod_substep(s)
        n = st.ctx.death(st.buf, n)
    return n


def snapshot(st, n):
    cells = st.cells()[:n].copy()
    fields = [st.ctx.field(s) for s in range(st.P.n_substrates)]
    return cells, fields


def test_checkpoint_restart_continues_bit_for_bit(tmp_path):
    N, n = 32, 3000
    P = params(N, capacity=2 * n, n_sub=2, seed=99)
    cells = mixed_cells(N, 10.0, n, P.capacity, seed=5, dead_frac=0.0, outside=0)
    # premitotic cells whose clocks reach duration_pre2post (780 min) in each
    # of the four steps, so that divisions happen before and after the restart
    i = np.arange(n)
    cells["phase"][:n] = np.where(i % 4 == 0, 1, 0)
    cells["cell_clock"][:n] = np.where(i % 4 == 0, 774.0 - 6.0 * ((i // 4) % 4), 0.0).astype(np.float32)
    path = tmp_path / "state.bin"

    a = GpuState(P, cells)
    n2 = bio_steps(a, n, 0, 2)
    a.ctx.checkpoint_save(a.buf, n2, 2, path)
    saved = snapshot(a, n2)
    n4 = bio_steps(a, n2, 2, 2)
    want = snapshot(a, n4)
    a.close()
    log(f"[ckpt] n {n} -> {n2} -> {n4}, file {path.stat().st_size} bytes")
    assert n4 > n2 > n

    b = GpuState(P, np.zeros(P.capacity, rt.CELL_DTYPE))
    m, step = b.ctx.checkpoint_load(b.buf, path)
    assert (m, step) == (n2, 2)
    got0 = snapshot(b, m)
    assert got0[0].tobytes() == saved[0].tobytes()
    for s in range(P.n_substrates):
        assert got0[1][s].tobytes() == saved[1][s].tobytes()
    m4 = bio_steps(b, m, step, 2)
    got = snapshot(b, m4)
    b.close()
    assert m4 == n4
    assert got[0].tobytes() == want[0].tobytes()
    for s in range(P.n_substrates):
        assert got[1][s].tobytes() == want[1][s].tobytes()


def test_checkpoint_refuses_other_grid_and_other_files(tmp_path):
    P = params(32, capacity=1000)
    cells = mixed_cells(32, 10.0, 500, P.capacity, seed=1, outside=0)
    a = GpuState(P, cells)
    a.ctx.checkpoint_save(a.buf, 500, 0, tmp_path / "a.bin")
    small = GpuState(params(32, capacity=400), cells[:400])
    with pytest.raises(rt.GellError, match="capacity"):
        small.ctx.checkpoint_load(small.buf, tmp_path / "a.bin")
    other = GpuState(params(40, capacity=1000), cells)
    with pytest.raises(rt.GellError, match="grid"):
        other.ctx.checkpoint_load(other.buf, tmp_path / "a.bin")
    other_seed = GpuState(params(32, capacity=1000, seed=124), cells)
    with pytest.raises(rt.GellError, match="seed"):
        other_seed.ctx.checkpoint_load(other_seed.buf, tmp_path / "a.bin")
    other_dt = GpuState(params(32, capacity=1000, dt_mech=0.05), cells)
    with pytest.raises(rt.GellError, match="time steps"):
        other_dt.ctx.checkpoint_load(other_dt.buf, tmp_path / "a.bin")
    p_sub = params(32, capacity=1000)
    p_sub.sub[0].uptake = -5.0
    other_sub = GpuState(p_sub, cells)
    with pytest.raises(rt.GellError, match="substrate constants"):
        other_sub.ctx.checkpoint_load(other_sub.buf, tmp_path / "a.bin")
    (tmp_path / "junk.bin").write_bytes(b"not a checkpoint" * 100)
    with pytest.raises(rt.GellError, match="format"):
        a.ctx.checkpoint_load(a.buf, tmp_path / "junk.bin")
    raw = (tmp_path / "a.bin").read_bytes()
    (tmp_path / "cut.bin").write_bytes(raw[: len(raw) // 2])
    with pytest.raises(rt.GellError, match="truncated"):
        a.ctx.checkpoint_load(a.buf, tmp_path / "cut.bin")
    for st in (a, small, other, other_seed, other_dt, other_sub):
        st.close()


def test_hds_driver_first_bio_step_matches_oracle(tmp_path, oracle_omp):
    exe = hip_build.OUT / "gell_sim"
    assert exe.exists(), f"{exe} not built"
    out = tmp_path / "hds"
    out.mkdir()
    cmd = [str(exe), "hds", "--steps", "1", "--exact", "--save-cells", "--checkpoint-every", "1", "--out", str(out)]
    log("[hds] " + " ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    log(r.stdout[-2000:] + r.stderr[-2000:])
    assert r.returncode == 0

    P = rt.default_params()  # the reference's Definitions.h
    assert (P.N, P.dx, P.capacity) == (100, 25.0, 2_000_000)
    n0 = 2347
    cells = rt.sphere_init(P, n0, seed=12345)
    ref = tmp_path / "ref"
    ref.mkdir()
    rt.savecsv_cell(cells, n0, ref / "Gell_0.csv")
    field = np.full(P.N ** 3, 38.0, np.float32)
    rt.savecsv_meshslice(field, P.N, ref / "Mesh_slice_0.csv")
    n1 = oracle_omp.run(P, cells, n0, field, 1)
    log(f"[hds] oracle: {n0} -> {n1} cells")
    rt.savecsv_cell(cells, n1, ref / "Gell_Simuend.csv")
    rt.savecsv_meshslice(field, P.N, ref / "Mesh_slice_Simuend.csv")
    for name in ("Gell_0.csv", "Mesh_slice_0.csv", "Gell_Simuend.csv", "Mesh_slice_Simuend.csv"):
        assert (out / name).read_bytes() == (ref / name).read_bytes(), name

    rows = (out / "Gell_Time_record.csv").read_text().splitlines()
    assert rows[0] == "Index,Cell Num,Time" and len(rows) == 3
    assert [int(r.split(",")[1]) for r in rows[1:]] == [n0, n1]
    gap = (out / "Gell_GapTime_record.csv").read_text().splitlines()
    assert len(gap) == 1 + 4560 // 360

    st = GpuState(P, np.zeros(16, rt.CELL_DTYPE))
    m, step = st.ctx.checkpoint_load(st.buf, out / "checkpoint.bin")
    assert (m, step) == (n1, 1)
    assert st.cells()[:m].tobytes() == cells[:n1].tobytes()
    assert st.ctx.field(0).tobytes() == field.tobytes()
    st.close()
