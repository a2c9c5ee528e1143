"""Pins of the oracle's uptake map and proliferation/death
(Cells.cuh:74-224, LODkernel.cuh:12-22, PDkernel.cuh:27-117,
RandomFunctions.cuh:23-32) by closed forms, hand-set thresholds and
invariants computed independently here."""
import numpy as np
import pytest

from gell_amd import runtime as rt
from helpers import mixed_cells, params, rel_l2

f32 = np.float32


def one_cell(phase=0, V=2494.0, Vf=1871.0, Vns=135.0, Vcs=489.0, clock=0.0, pos=(40.0, 40.0, 40.0)):
    c = np.zeros(1, rt.CELL_DTYPE)
    c["sign"] = 1
    c["phase"] = phase
    c["V"], c["Vf"], c["Vns"], c["Vcs"] = V, Vf, Vns, Vcs
    c["cell_clock"] = clock
    c["pos"] = pos
    c["r"] = (0.23873241 * V) ** (1 / 3)
    return c


def test_uptake_map_closed_form(oracle):
    N, dx = 16, 10.0
    P = params(N, dx=dx, capacity=500)
    cells = mixed_cells(N, dx, 500, 500, seed=5, outside=0)
    cells["mesh_idx"][:500] = np.random.default_rng(1).integers(0, 40, 500)  # crowd a few voxels
    rate = oracle.prelod(P, 0, cells, 500)
    want = np.zeros(N ** 3)
    for c in cells[:500]:
        if not c["sign"]:
            continue
        u = -10.0 * float(c["V"]) / dx ** 3
        if c["phase"] >= 3:
            u *= 0.1
        elif c["phase"] < 0:
            u = 0.0
        want[c["mesh_idx"]] += u
    assert rel_l2(rate, want) < 1e-6
    assert np.count_nonzero(rate[40:]) == 0
    # the per-cell value is the reference's float expression, scaled by phase
    c = one_cell(phase=4, V=1000.0)
    assert oracle.cell_uptake(P, 0, c) == f32(f32(f32(-10.0) * f32(f32(f32(f32(1000.0) / f32(10)) / f32(10)) / f32(10))) * f32(1.0)) * f32(0.1)


@pytest.mark.parametrize("phase", [0, 1, 2, 3, 4, 5])
def test_volume_relaxation_closed_form(phase, oracle):
    # per phase: rates and nuclear target (Cells.cuh:91-133); targets that do
    # not depend on V relax geometrically: x_k = x0_t + (x - x0_t)(1 - dt r)^k
    rates = {0: (3 / 60, .33 / 60, .33 / 60, 135.0, .7502), 1: (3 / 60, .27 / 60, .33 / 60, 270.0, .7502),
             2: (3 / 60, .33 / 60, .33 / 60, 135.0, .7502), 3: (3 / 60, 1 / 60, .35 / 60, 0.0, 0.0),
             4: (.67 / 60, .0032 / 60, .013 / 60, 0.0, 1.0), 5: (.05 / 60, .0032 / 60, .013 / 60, 0.0, 0.0)}
    rf, rcs, rns, vns0, ff = rates[phase]
    P = params(8)
    c = one_cell(phase=phase, V=2600.0, Vf=1900.0, Vns=150.0, Vcs=550.0)
    k = 40
    for _ in range(k):
        oracle.volume_update(P, c)
    dt = 6.0
    Vns = vns0 + (150.0 - vns0) * (1 - dt * rns) ** k
    # Vcs relaxes toward 3.615 * Vns_0 (constant target)
    Vcs = 3.615 * vns0 + (550.0 - 3.615 * vns0) * (1 - dt * rcs) ** k
    assert abs(float(c["Vns"][0]) - Vns) <= 1e-4 * max(1.0, abs(Vns))
    assert abs(float(c["Vcs"][0]) - Vcs) <= 1e-4 * max(1.0, abs(Vcs))
    if ff == 0.0:  # fluid volume decays to 0
        Vf = 1900.0 * (1 - dt * rf) ** k
        assert abs(float(c["Vf"][0]) - Vf) <= 1e-4 * max(1.0, Vf)
    V = float(c["Vf"][0]) + float(c["Vns"][0]) + float(c["Vcs"][0])
    assert abs(float(c["V"][0]) - V) <= 1e-4 * V
    assert abs(float(c["r"][0]) - (0.23873241 * float(c["V"][0])) ** (1 / 3)) <= 1e-5 * float(c["r"][0])


def probs(O2):
    """The phase probabilities with the reference's promotions (Cells.cuh:150-165)."""
    pa = f32((1.0 / 7.0 / 24.0 / 60.0) * 6.0)
    if O2 > 10:
        pp = f32((1.0 / 8.5 / 60.0) * 6.0)
    elif O2 < 5:
        pp = f32(0.0)
    else:
        pp = f32(float(f32(f32(f32(1.0) * (f32(O2) - f32(5.0))) / f32(5.0))) * (1.0 / 8.5 / 60.0) * 6.0)
    if O2 > 5:
        pn = f32(0.0)
    elif O2 < 2.5:
        pn = f32((1.0 / 6.0 / 60.0) * 6.0)
    else:
        pn = f32(float(f32(f32(f32(1.0) * (f32(5.0) - f32(O2))) / f32(2.5))) * (1.0 / 6.0 / 60.0) * 6.0)
    return pp, pa, pn


@pytest.mark.parametrize("O2", [20.0, 7.5, 3.0, 2.0])
def test_phase_thresholds(O2, oracle):
    P = params(8)
    pp, pa, pn = probs(O2)
    edges = [(pp, 1, pp), (f32(pp + pa), 3, pa), (f32(f32(pp + pa) + pn), 4, pn)]
    for edge, phase, width in edges:
        if width == 0:  # an empty window: the draw falls to the branch before it
            continue
        c = one_cell(phase=0)
        assert oracle.phase_update(P, c, float(edge), O2) is False
        assert c["phase"][0] == phase and c["cell_clock"][0] == 0.0
        # just above the edge: the next branch (or nothing)
        c = one_cell(phase=0)
        oracle.phase_update(P, c, float(np.nextafter(edge, f32(2.0))), O2)
        assert c["phase"][0] != phase
    c = one_cell(phase=0)
    oracle.phase_update(P, c, 1.0, O2)
    assert c["phase"][0] == 0 and c["cell_clock"][0] == 6.0


def test_deterministic_transitions(oracle):
    P = params(8)
    c = one_cell(phase=1, clock=774.0)   # 780 min premitotic -> divides
    assert oracle.phase_update(P, c, 1.0, 20.0) is True
    assert c["phase"][0] == 2 and c["cell_clock"][0] == 0.0
    c = one_cell(phase=1, clock=768.0)
    assert oracle.phase_update(P, c, 1.0, 20.0) is False and c["phase"][0] == 1
    c = one_cell(phase=2, clock=144.0)   # 150 min postmitotic -> quiescent
    oracle.phase_update(P, c, 1.0, 20.0)
    assert c["phase"][0] == 0 and c["cell_clock"][0] == 0.0
    c = one_cell(phase=4, V=4988.0)      # early -> late necrotic at 2 x 2494
    oracle.phase_update(P, c, 0.5, 1.0)
    assert c["phase"][0] == 5
    c = one_cell(phase=3, Vcs=24.0)      # lysed below 0.05 * 488
    oracle.phase_update(P, c, 0.5, 20.0)
    assert c["sign"][0] == 0


def test_random_direction(oracle):
    for u, v in [(0.25, 0.5), (0.7, 0.1), (1.0, 1.0), (1e-6, 0.999)]:
        d = oracle.random_dir(u, v)
        assert abs(np.linalg.norm(d.astype(np.float64)) - 1.0) < 1e-6
    d = oracle.random_dir(0.25, 0.5)          # theta = pi/2, phi = pi/2 -> +x
    assert np.allclose(d, [1.0, 0.0, 0.0], atol=1e-6)
    d = oracle.random_dir(0.5, 1.0)           # phi = acos(1) = 0 -> +z
    assert np.allclose(d, [0.0, 0.0, 1.0], atol=1e-6)


@pytest.mark.parametrize("room", [None, 25])
def test_birth_invariants(room, oracle):
    N, n = 16, 400
    cap = 2 * n if room is None else n + room
    P = params(N, capacity=cap, seed=99)
    cells = mixed_cells(N, 10.0, n, cap, seed=6, outside=0, dead_frac=0.1)
    cells["phase"][:n] = np.where(np.arange(n) % 3 == 0, 1, 0)
    cells["cell_clock"][:n] = np.where(np.arange(n) % 3 == 0, 774.0, 0.0).astype(np.float32)
    p0 = np.full(N ** 3, 38.0, np.float32)
    before = cells.copy()
    n2, slots = oracle.birth(P, p0, cells, n, 4)
    div = np.nonzero(slots >= 0)[0]
    assert n2 == n + len(div)
    assert np.array_equal(slots[div], np.arange(n, n2))             # slots in mother order
    wanted = ((before["phase"][:n] == 1) & (before["sign"][:n] != 0)).sum()
    assert len(div) == (wanted if room is None else min(wanted, cap - 1 - n))
    for m in div:
        d = slots[m]
        mom, dau = cells[m], cells[d]
        assert mom["phase"] == 2 == dau["phase"] and dau["cell_clock"] == 0 and dau["mesh_idx"] == -1
        assert dau["sign"] == 1 and np.all(dau["force"] == 0) and np.all(dau["oldforce"] == 0)
        for f in ("V", "Vf", "Vcs", "Vns", "r"):
            assert mom[f] == dau[f]
        r_before = mom["r"] / np.float32(0.7937)
        sep = np.linalg.norm(dau["pos"].astype(np.float64) - mom["pos"])
        assert abs(sep - 2 * 0.2063 * r_before) < 1e-3
        mid = (dau["pos"].astype(np.float64) + mom["pos"]) / 2
        assert np.allclose(mid, before["pos"][m], atol=1e-3)
    # refused divisions (capacity) keep phase 2 / clock 0 without halving
    if room is not None:
        refused = np.nonzero((before["phase"][:n] == 1) & (before["sign"][:n] != 0) & (slots < 0))[0]
        assert len(refused) > 0
        assert np.all(cells["phase"][refused] == 2)


def test_death_is_a_stable_filter(oracle):
    n = 1000
    cells = mixed_cells(16, 10.0, n, n, seed=7, dead_frac=0.3, outside=0)
    keep = cells[:n][cells["sign"][:n] != 0].copy()
    work = cells.copy()
    m = oracle.death(work, n)
    assert m == len(keep)
    assert np.array_equal(work[:m].view(np.uint8), keep.view(np.uint8))
