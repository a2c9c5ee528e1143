"""Seeded synthetic spheroids in the 72-byte cell layout.

Follows the reference's initializer (Initialization.cuh:15-43): rejection
sampling of uniform points in a ball centred in the domain, phase 0, clock 0,
zero forces, volumes in the reference ratio Vf : Vns : Vcs = 1871 : 135 : 489
(Initialization.cuh:34-37) and mesh_idx of the containing voxel. The radius
is drawn per cell (r = r_lo + U(0, r_hi - r_lo)) and V = 4/3 pi r^3, so that
the reference's Radius_update (r = (0.23873241 V)^(1/3), Cells.cuh:69-71)
returns r.

``radial_profile`` gives the heterogeneous-spheroid layout: a proliferating
rim (phases 0-2), a quiescent shell (phase 0) and a necrotic core (phases
4/5), by distance from the centre.
"""
from __future__ import annotations

import numpy as np

from .runtime import CELL_DTYPE

VF, VNS, VCS = 1871.0, 135.0, 489.0
VTOT = VF + VNS + VCS


def ball_points(rng: np.random.Generator, n: int, radius: float) -> np.ndarray:
    """n points uniform in a ball, by rejection from the cube (float32)."""
    out = np.empty((n, 3), np.float32)
    got = 0
    while got < n:
        m = int((n - got) * 2.0) + 64
        p = rng.random((m, 3), dtype=np.float32) * np.float32(2.0) - np.float32(1.0)
        p = p[(p * p).sum(1) < 1.0]
        take = min(len(p), n - got)
        out[got:got + take] = p[:take]
        got += take
    return out * np.float32(radius)


def make_cells(n: int, N: int, dx: float, radius: float, seed: int, capacity: int | None = None,
               r_lo: float = 4.5, r_hi: float = 5.5, center=None) -> np.ndarray:
    """A spheroid of n cells; returns a record array of ``capacity`` slots
    (slots >= n hold the reference's default Cell values)."""
    capacity = n if capacity is None else capacity
    assert capacity >= n
    rng = np.random.default_rng(seed)
    c = np.float32(N * dx / 2.0) if center is None else np.asarray(center, np.float32)
    cells = np.zeros(capacity, CELL_DTYPE)
    # Cell's default member values for unused slots (Cells.cuh:10-33).
    cells["mesh_idx"] = -1
    cells["pos"] = -1.0
    cells["r"] = -1.0
    cells["V"] = cells["Vf"] = cells["Vcs"] = cells["Vns"] = -1.0
    cells["phase"] = -1
    if n == 0:
        return cells
    pos = ball_points(rng, n, radius) + c
    r = (np.float32(r_lo) + rng.random(n, dtype=np.float32) * np.float32(r_hi - r_lo)).astype(np.float32)
    V = (np.float32(4.0 / 3.0 * np.pi) * r * r * r).astype(np.float32)
    live = cells[:n]
    live["sign"] = 1
    live["pos"] = pos
    live["r"] = r
    live["V"] = V
    live["Vf"] = V * np.float32(VF / VTOT)
    live["Vns"] = V * np.float32(VNS / VTOT)
    live["Vcs"] = V * np.float32(VCS / VTOT)
    live["phase"] = 0
    live["cell_clock"] = 0.0
    live["force"] = 0.0
    live["oldforce"] = 0.0
    ijk = (pos / np.float32(dx)).astype(np.int64)  # truncation toward zero, as Find_Mesh
    inside = np.all((ijk >= 0) & (ijk < N), axis=1)
    live["mesh_idx"] = np.where(inside, ijk[:, 0] + N * (ijk[:, 1] + N * ijk[:, 2]), -1)
    return cells


def radial_profile(cells: np.ndarray, n: int, center, radius: float, seed: int,
                   rim: float = 0.15, core: float = 0.5) -> None:
    """Assign phases by radius in place: outer `rim` fraction of the radius
    proliferating (phases 0, 1, 2 mixed, clocks spread over their windows),
    the next shell quiescent, and the inner `core` fraction necrotic (4, 5)."""
    rng = np.random.default_rng(seed)
    live = cells[:n]
    d = np.sqrt(((live["pos"] - np.asarray(center, np.float32)) ** 2).sum(1)) / np.float32(radius)
    ph = np.zeros(n, np.int32)
    clock = np.zeros(n, np.float32)
    rim_m = d > 1.0 - rim
    core_m = d < core
    u = rng.random(n)
    ph[rim_m & (u < 0.3)] = 1
    ph[rim_m & (u >= 0.3) & (u < 0.6)] = 2
    clock[ph == 1] = (rng.random(int((ph == 1).sum())) * 780.0).astype(np.float32)
    clock[ph == 2] = (rng.random(int((ph == 2).sum())) * 150.0).astype(np.float32)
    nec = core_m & (u < 0.7)
    ph[nec] = 4
    ph[core_m & (u >= 0.7)] = 5
    live["phase"] = ph
    live["cell_clock"] = clock
