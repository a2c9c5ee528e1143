"""Shared helpers for the parity tests: synthetic states and comparisons."""
from __future__ import annotations

import sys

import numpy as np

from gell_amd import runtime as rt
from gell_amd import spheroid

FLOAT_FIELDS = ["pos", "force", "oldforce", "r", "V", "Vf", "Vcs", "Vns", "cell_clock"]
INT_FIELDS = ["sign", "mesh_idx", "phase"]


def log(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def params(N, dx=10.0, capacity=None, n_sub=1, seed=123, **kw):
    subs = [dict(boundary=38.0, D=1e5, decay=0.1, uptake=-10.0, dead_factor=0.1, initial=38.0)]
    if n_sub > 1:
        subs.append(dict(boundary=5.0, D=5e4, decay=0.0, uptake=-2.0, dead_factor=0.1, initial=5.0))
    p = rt.make_params(N=N, dx=dx, capacity=capacity or 4096, seed=seed, substrates=subs[:max(1, n_sub)], **kw)
    return p


def mixed_cells(N, dx, n, capacity, seed, radius=None, dead_frac=0.05, outside=3, dense=False):
    """A spheroid with some dead cells, some cells outside the domain and
    (when dense) a packed core so that cells have many neighbours."""
    radius = radius if radius is not None else 0.35 * N * dx
    cells = spheroid.make_cells(n, N, dx, radius, seed, capacity=capacity)
    rng = np.random.default_rng(seed + 1)
    live = cells[:n]
    if dense:
        m = n // 2
        c = np.float32(N * dx / 2)
        live["pos"][:m] = spheroid.ball_points(rng, m, 2.5 * dx) + c
    dead = rng.random(n) < dead_frac
    live["sign"][dead] = 0
    if outside and n > outside:
        idx = rng.choice(n, outside, replace=False)
        live["pos"][idx, 0] = np.float32(N * dx + 5.0)
    # varied phases so that uptake sees all branches
    live["phase"] = rng.choice([0, 0, 0, 1, 2, 3, 4, 5], n).astype(np.int32)
    live["oldforce"] = (rng.standard_normal((n, 3)) * 0.5).astype(np.float32)
    return cells


def as_bits(a):
    a = np.ascontiguousarray(a, np.float32)
    return a.view(np.uint32)


def ulp_diff(a, b) -> int:
    """Largest distance in units in the last place between two float32 arrays."""
    ia = as_bits(a).astype(np.int64)
    ib = as_bits(b).astype(np.int64)
    ia = np.where(ia >= 2 ** 31, 2 ** 31 - ia, ia)
    ib = np.where(ib >= 2 ** 31, 2 ** 31 - ib, ib)
    return int(np.abs(ia - ib).max()) if ia.size else 0


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def compare_cells(got, want, n, float_tol=1e-5, exact=True, skip=()):
    """Integer fields must match exactly; float fields within float_tol
    relative L2 (and bit-exact when `exact`). Returns a report dict."""
    rep = {}
    for f in INT_FIELDS:
        if f in skip:
            continue
        assert np.array_equal(got[f][:n], want[f][:n]), f"{f} differs at {np.nonzero(got[f][:n] != want[f][:n])[0][:10]}"
    for f in FLOAT_FIELDS:
        if f in skip:
            continue
        g, w = got[f][:n], want[f][:n]
        e = rel_l2(g, w)
        u = ulp_diff(g, w)
        rep[f] = (e, u)
        assert e <= float_tol, f"{f}: rel L2 {e:.3e} > {float_tol}"
        if exact:
            assert u == 0, f"{f}: max ulp diff {u} (rel L2 {e:.3e})"
    return rep


class GpuState:
    """Context + device cell array for one parameter set. With exact=True the
    LOD uses the serial line kernels, which are bit-identical to the oracle."""

    def __init__(self, P, cells: np.ndarray, exact: bool = False):
        self.P = P
        self.ctx = rt.Context(P)
        if exact:
            self.ctx.tune("lod_parts", 1)
            self.ctx.tune("lod_x_parts", 1)
        self.buf = rt.DeviceBuffer(P.capacity * 72)
        self.upload(cells)

    def upload(self, cells: np.ndarray) -> None:
        c = np.zeros(self.P.capacity, rt.CELL_DTYPE)
        c[: len(cells)] = cells[: self.P.capacity]
        self.buf.upload(c)

    def cells(self) -> np.ndarray:
        self.ctx.sync()
        return self.buf.download(rt.CELL_DTYPE, self.P.capacity)

    def close(self):
        self.buf.free()
        self.ctx.close()
