"""Parity at the benchmark workloads' full size, against the OpenMP oracle
(bit-identical to the serial one).

L: 256^3 voxels at 10 um, 1M cells in a 1.2 mm ball, oxygen + glucose.
H: 512^3 voxels, 10M cells at about one cell per voxel with the radial
   phase profile (proliferating rim, quiescent shell, necrotic core).

Both start from the workload bench.py times, with the radial phase profile
(so that births happen), 2 % of the cells dead and a few outside the domain
(so that death compacts). Each stage is compared from the same input state,
with the autotuned (default) kernel variants:

  mechanics step (FM_update, FMkernel.cuh:181-202): keys, sort order,
      per-voxel ranges and every record bit-exact;
  uptake maps (PreLOD, LODkernel.cuh:270-278): bit-exact, every substrate;
  11 LOD substeps per substrate (LODsolver, LODkernel.cuh:280-300): within
      1e-5 relative L2 (the partitioned line solves round differently);
  birth (CellBirth_kernel, PDkernel.cuh:88-98) on the GPU's field: count,
      daughter slots and records (division geometry within 1e-6);
  death (CellDeath_kernel, PDkernel.cuh:100-117): count and records byte for
      byte.
"""
import numpy as np
import pytest

import bench
from gell_amd import runtime as rt
from gell_amd import spheroid
from helpers import compare_cells, log, rel_l2, ulp_diff

pytestmark = pytest.mark.gpu

SEED = 20240611


def _workload(name):
    cfg = bench.CONFIGS[name]
    P, cells = bench.make_workload(cfg, SEED)
    n = cfg["cells"]
    c = cfg["N"] * cfg["dx"] / 2.0
    if not cfg.get("profile"):
        spheroid.radial_profile(cells, n, (c, c, c), cfg["radius"], SEED + 1)
    rng = np.random.default_rng(SEED + 2)
    live = cells[:n]
    live["sign"][rng.random(n) < 0.02] = 0
    out = rng.choice(n, 16, replace=False)
    live["pos"][out, 0] = np.float32(cfg["N"] * cfg["dx"] + 3.0)
    # phase-1 cells at the end of their window divide in this biology step
    ph1 = np.nonzero(live["phase"] == 1)[0]
    live["cell_clock"][ph1[: len(ph1) // 4]] = np.float32(774.0)
    return cfg, P, cells, n


def _scale_case(name, oracle):
    cfg, P, cells, n = _workload(name)
    N = cfg["N"]
    log(f"[scale {name}] N={N} n={n} substrates={P.n_substrates}")
    ctx = rt.Context(P)
    tuned = ctx.autotune()
    log(f"[scale {name}] tuned {tuned}")
    buf = rt.DeviceBuffer(P.capacity * 72)
    buf.upload(cells)

    # -- mechanics step -------------------------------------------------------
    ctx.fm_update(buf, n)
    got = buf.download(rt.CELL_DTYPE, P.capacity)
    keys, order = ctx.sorted_keys(n), ctx.sorted_order(n)
    s, e = ctx.voxel_ranges()
    want = cells.copy()
    wk, wo, ws, we = oracle.fm_update(P, want, n)
    assert np.array_equal(keys, wk), "keys"
    assert np.array_equal(order, wo), "order"
    assert np.array_equal(s, ws) and np.array_equal(e, we), "ranges"
    del s, e, ws, we
    compare_cells(got, want, n)
    log(f"[scale {name}] mechanics step bit-exact ({int(got['sign'][:n].sum())} live)")

    # -- uptake maps ------------------------------------------------------------
    for sub in range(P.n_substrates):
        ctx.prelod(sub, buf, n)
    rates = []
    for sub in range(P.n_substrates):
        r = oracle.prelod(P, sub, want, n)
        assert np.array_equal(ctx.uptake_field(sub).view(np.uint32), r.view(np.uint32)), f"map {sub}"
        rates.append(r)
    log(f"[scale {name}] uptake maps bit-exact")

    # -- LOD substeps -------------------------------------------------------------
    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
    for sub in range(P.n_substrates):
        f = np.full(N ** 3, P.sub[sub].initial, np.float32)
        for _ in range(diff):
            ctx.lod_substep(sub)
            oracle.lod_substep(P, sub, f, rates[sub])
        g = ctx.field(sub)
        err = rel_l2(g, f)
        log(f"[scale {name}] substrate {sub}: {diff} substeps rel L2 {err:.2e}, "
            f"max abs {float(np.abs(g - f).max()):.2e}")
        assert err <= 1e-5
    del rates

    # -- birth, on the GPU's oxygen field and cells ----------------------------------
    p0 = ctx.field(0)
    state = buf.download(rt.CELL_DTYPE, P.capacity)
    n_gpu = ctx.birth(buf, n, 0)
    got = buf.download(rt.CELL_DTYPE, P.capacity)
    slots = ctx.birth_slots(n)
    want = state
    n_cpu, wslots = oracle.birth(P, p0, want, n, 0)
    log(f"[scale {name}] birth n {n} -> gpu {n_gpu} cpu {n_cpu}")
    assert n_gpu == n_cpu and n_cpu > n
    assert np.array_equal(slots, wslots)
    compare_cells(got, want, n_cpu, exact=False, skip=("pos",), float_tol=1e-6)
    for f in ("r", "V", "Vf", "Vcs", "Vns", "cell_clock", "force", "oldforce"):
        assert ulp_diff(got[f][:n_cpu], want[f][:n_cpu]) == 0, f
    div = np.concatenate([np.nonzero(wslots >= 0)[0], wslots[wslots >= 0]])
    same = np.ones(n_cpu, bool)
    same[div] = False
    assert ulp_diff(got["pos"][:n_cpu][same], want["pos"][:n_cpu][same]) == 0
    assert rel_l2(got["pos"][:n_cpu], want["pos"][:n_cpu]) <= 1e-6
    del want, state

    # -- death, from the GPU's state ------------------------------------------------
    m_gpu = ctx.death(buf, n_gpu)
    after = buf.download(rt.CELL_DTYPE, P.capacity)
    m_cpu = oracle.death(got, n_gpu)
    log(f"[scale {name}] death n {n_gpu} -> gpu {m_gpu} cpu {m_cpu}")
    assert m_gpu == m_cpu < n_gpu
    assert np.array_equal(after[:m_gpu].view(np.uint8), got[:m_cpu].view(np.uint8))
    buf.free()
    ctx.close()


@pytest.mark.timeout(300)
def test_L_workload_stages_match_oracle(oracle_omp):
    _scale_case("L", oracle_omp)


@pytest.mark.timeout(600)
def test_H_workload_stages_match_oracle(oracle_omp):
    _scale_case("H", oracle_omp)
