"""Host-side code either side of the per-step path, on the CPU: the seeded
spheroid initializer (Initialization.cuh:15-43) and the CSV writers
(FileSave.cuh:24-86). The library is loaded but no device call is made.

The CSV files are checked against the reference's format strings applied
with C printf semantics (Python's %-formatting of %d and %f is C's), on
values chosen so that a wrong index, plane, column order or row order shows:
the field of voxel (i, j, k) is i + 10 j + 100 k, counts are their own
indices. The cell columns include those the reference's ParaView state reads
(visualization/HDS_Slice.pvsm: x, y, z, r, Phase)."""
import math

import numpy as np
import pytest

from gell_amd import runtime as rt

CELL_ROW = "%d,%d,%d,%f,%f,%f,%f,%f,%f,%f\n"  # FileSave.cuh:29


def read(path):
    with open(path, newline="") as f:
        return f.read()


def some_cells(n, cap, seed=3):
    rng = np.random.default_rng(seed)
    c = np.zeros(cap, rt.CELL_DTYPE)
    c["sign"][:n] = rng.integers(0, 2, n)
    c["phase"][:n] = rng.integers(0, 6, n)
    c["pos"][:n] = rng.uniform(0, 2500, (n, 3)).astype(np.float32)
    c["r"][:n] = rng.uniform(4, 9, n).astype(np.float32)
    c["V"][:n] = rng.uniform(1000, 5000, n).astype(np.float32)
    c["Vns"][:n] = rng.uniform(100, 200, n).astype(np.float32)
    c["Vcs"][:n] = rng.uniform(400, 600, n).astype(np.float32)
    c["Vf"][:n] = -7.0  # not a column of the file
    return c


def test_cells_csv(tmp_path):
    n = 57
    c = some_cells(n, 64)
    path = tmp_path / "Gell_1.csv"
    rt.savecsv_cell(c, n, path)
    text = read(path)
    lines = text.split("\n")
    assert lines[0] == "Index,Sign,Phase,x,y,z,r,V,Vns,Vcs"
    assert {"x", "y", "z", "r", "Phase"} <= set(lines[0].split(","))
    want = lines[0] + "\n" + "".join(
        CELL_ROW % (i, c["sign"][i], c["phase"][i], *c["pos"][i], c["r"][i], c["V"][i], c["Vns"][i], c["Vcs"][i])
        for i in range(n))
    assert text == want
    back = np.loadtxt(path, delimiter=",", skiprows=1)
    assert back.shape == (n, 10)
    assert np.allclose(back[:, 3:6], c["pos"][:n], atol=1e-6 * 2500)
    assert np.array_equal(back[:, 0], np.arange(n))


def test_cells_csv_empty_and_errors(tmp_path):
    path = tmp_path / "empty.csv"
    rt.savecsv_cell(np.zeros(4, rt.CELL_DTYPE), 0, path)
    assert read(path) == "Index,Sign,Phase,x,y,z,r,V,Vns,Vcs\n"
    with pytest.raises(rt.GellError, match="cannot open"):
        rt.savecsv_cell(some_cells(3, 3), 3, tmp_path / "no_such_dir" / "x.csv")


def coded_field(N):
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return (i + 10 * j + 100 * k).astype(np.float32).ravel()  # x fastest


def test_mesh_csv(tmp_path):
    N = 5
    path = tmp_path / "Mesh.csv"
    rt.savecsv_mesh(coded_field(N), N, path)
    rows = read(path).splitlines()
    assert rows[0] == "i,j,k,Concentration"
    assert len(rows) == N ** 3 + 1
    for r in rows[1:]:
        i, j, k, v = r.split(",")
        assert float(v) == int(i) + 10 * int(j) + 100 * int(k)
        assert v == "%f" % float(v)
    assert rows[1] == "0,0,0,0.000000" and rows[2] == "1,0,0,1.000000" and rows[N + 1] == "0,1,0,10.000000"


@pytest.mark.parametrize("N", [4, 7])
def test_meshslice_csv_is_the_middle_plane(tmp_path, N):
    path = tmp_path / "Mesh_slice.csv"
    rt.savecsv_meshslice(coded_field(N), N, path)
    rows = read(path).splitlines()
    assert rows[0] == "x,y,Concentration"
    assert len(rows) == N * N + 1
    for idx, r in enumerate(rows[1:]):
        x, y, v = r.split(",")
        assert (int(x), int(y)) == (idx % N, idx // N)  # x fastest
        assert float(v) == int(x) + 10 * int(y) + 100 * (N // 2)


def test_counts_csv(tmp_path):
    counts = np.array([2347, 2350, 2361, 2400], np.int32)
    times = np.array([0.0, 0.012, 0.0135, 1.5])
    path = tmp_path / "Gell_Time_record.csv"
    rt.savecsv_array(counts, times, path)
    assert read(path) == ("Index,Cell Num,Time\n0,2347,0.000000\n1,2350,0.012000\n2,2361,0.013500\n"
                          "3,2400,1.500000\n")


def test_gap_counts_csv_uses_the_reference_indexing(tmp_path):
    # rows i < n_iter / gap hold counts[(i + 1) gap - 1] and gap_times[i]
    n_iter, gap = 4560, 360
    counts = np.arange(n_iter + 1, dtype=np.int32)
    gap_times = np.arange(77, dtype=np.float64) * 0.5
    path = tmp_path / "Gell_GapTime_record.csv"
    rt.savecsv_gaparray(counts, gap_times, n_iter, gap, path)
    rows = read(path).splitlines()
    assert rows[0] == "Index,Cell Num,Time"
    assert len(rows) == 1 + 12
    assert rows[1] == "0,359,0.000000"
    assert rows[12] == "11,4319,5.500000"


HDS_BALL = 8.4 * 2347 ** (1 / 3) / 1.612 * 1.87  # Initialization.cuh:18, about 129.5 um


def test_sphere_init_hds():
    P = rt.make_params()  # Voxel_num 100, Voxel_length 25, Max_Cell_num 2M
    n, cap = 2347, 4096
    c = rt.sphere_init(P, n, cap, 8.4, seed=12345)
    assert abs(HDS_BALL - 129.497) < 1e-3
    live = c[:n]
    assert (live["sign"] == 1).all() and (live["phase"] == 0).all()
    assert (live["r"] == np.float32(8.4)).all()
    for f, v in (("V", 2494), ("Vf", 1871), ("Vns", 135), ("Vcs", 489), ("cell_clock", 0)):
        assert (live[f] == v).all(), f
    assert not live["force"].any() and not live["oldforce"].any()
    d = np.linalg.norm(live["pos"].astype(np.float64) - 1250.0, axis=1)
    assert d.max() < HDS_BALL * (1 + 1e-6)
    assert d.max() > 0.97 * HDS_BALL
    # uniform in a ball: E[d^2] = 3/5 R^2, E[d] = 3/4 R; sd of the sample means ~0.6% and ~0.4%
    assert abs((d ** 2).mean() / HDS_BALL ** 2 - 0.6) < 0.02
    assert abs(d.mean() / HDS_BALL - 0.75) < 0.015
    # octants balanced
    octant = ((live["pos"] > 1250).astype(int) * [1, 2, 4]).sum(1)
    assert np.bincount(octant, minlength=8).min() > n / 8 * 0.8
    ijk = np.floor(live["pos"] / np.float32(25)).astype(np.int64)
    assert np.array_equal(live["mesh_idx"], ijk[:, 0] + 100 * (ijk[:, 1] + 100 * ijk[:, 2]))
    rest = c[n:]
    assert (rest["sign"] == 0).all() and (rest["phase"] == -1).all() and (rest["mesh_idx"] == -1).all()
    assert (rest["r"] == -1).all() and (rest["pos"] == -1).all() and (rest["cell_clock"] == 0).all()


def test_sphere_init_seeded():
    P = rt.make_params(N=40, dx=10.0)
    a = rt.sphere_init(P, 500, 512, 5.0, seed=7)
    b = rt.sphere_init(P, 500, 512, 5.0, seed=7)
    c = rt.sphere_init(P, 500, 512, 5.0, seed=8)
    assert a.tobytes() == b.tobytes()
    assert not np.array_equal(a["pos"][:500], c["pos"][:500])
    R = 5.0 * 500 ** (1 / 3) / 1.612 * 1.87
    d = np.linalg.norm(a["pos"][:500].astype(np.float64) - 200.0, axis=1)
    assert d.max() < R * (1 + 1e-6) and math.isclose(R, 46.036, rel_tol=1e-4)


def test_sphere_init_rejects_bad_sizes():
    P = rt.make_params()
    with pytest.raises(rt.GellError):
        rt.sphere_init(P, 10, 5)
