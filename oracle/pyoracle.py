"""ctypes binding of the CPU oracle (``oracle/_build/liboracle*.so``).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by bench.py's CPU-baseline leg. The product never
imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BUILD = HERE / "_build"


def build() -> None:
    """Compile the oracle libraries (serial and OpenMP)."""
    r = subprocess.run(["make", "-s", "-C", str(HERE), "-j2"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")


_libs: dict[str, C.CDLL] = {}


def load(omp: bool = False) -> C.CDLL:
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = BUILD / name
    if not path.exists():
        build()
    lib = C.CDLL(str(path))
    vp, i32, u32, u64, f32 = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64, C.c_float
    sig = {
        "orc_omp_threads": (C.c_int, []),
        "orc_philox4x32_10": (None, [vp, vp, vp]),
        "orc_birth_uniforms": (None, [u64, u32, u64, vp]),
        "orc_morton": (i32, [i32, i32, i32]),
        "orc_morton_decode": (None, [i32, vp]),
        "orc_float_loop_count": (i32, [f32, f32]),
        "orc_fm_keys": (None, [vp, vp, i32, vp]),
        "orc_fm_update": (None, [vp, vp, i32, vp, vp, vp, vp]),
        "orc_forces": (None, [vp, vp, i32, vp, vp, vp]),
        "orc_move": (None, [vp, vp, i32]),
        "orc_prelod": (None, [vp, i32, vp, i32, vp]),
        "orc_cell_uptake": (f32, [vp, i32, vp]),
        "orc_lod_coeffs": (None, [vp, i32, vp, vp]),
        "orc_lod_coeffs_f64": (None, [vp, i32, vp, vp]),
        "orc_thomas_line": (None, [vp, i32, vp, vp, f32]),
        "orc_lod_substep": (None, [vp, i32, vp, vp]),
        "orc_lod_substep_f64": (None, [vp, i32, vp, vp]),
        "orc_volume_update": (None, [vp, vp]),
        "orc_phase_update": (i32, [vp, vp, f32, f32]),
        "orc_random_dir": (None, [f32, f32, vp]),
        "orc_birth": (i32, [vp, vp, vp, i32, u64, vp]),
        "orc_death": (i32, [vp, i32]),
        "orc_run": (i32, [vp, vp, i32, vp, i32, i32, i32, u64, vp]),
        "orc_run_f64": (i32, [vp, vp, i32, vp, i32, i32, i32, u64]),
        "orc_run_mech": (i32, [vp, vp, i32, vp, i32]),
        "orc_run_mech_f64": (i32, [vp, vp, i32, vp, i32]),
    }
    for n, (res, args) in sig.items():
        f = getattr(lib, n)
        f.restype = res
        f.argtypes = args
    _libs[name] = lib
    return lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _pp(params) -> int:
    return C.addressof(params)


class Oracle:
    """Thin numpy-facing wrapper; arrays are modified in place where the
    reference modifies its state in place."""

    def __init__(self, omp: bool = False):
        self.lib = load(omp)

    def threads(self) -> int:
        return self.lib.orc_omp_threads()

    def philox(self, ctr, key) -> np.ndarray:
        c = np.asarray(ctr, np.uint32)
        k = np.asarray(key, np.uint32)
        out = np.zeros(4, np.uint32)
        self.lib.orc_philox4x32_10(_p(c), _p(k), _p(out))
        return out

    def birth_uniforms(self, seed: int, slot: int, step: int) -> np.ndarray:
        out = np.zeros(3, np.float32)
        self.lib.orc_birth_uniforms(seed, slot, step, _p(out))
        return out

    def morton(self, i, j, k) -> int:
        return self.lib.orc_morton(i, j, k)

    def morton_decode(self, key) -> tuple[int, int, int]:
        out = np.zeros(3, np.int32)
        self.lib.orc_morton_decode(key, _p(out))
        return tuple(int(x) for x in out)

    def float_loop_count(self, end, step) -> int:
        return self.lib.orc_float_loop_count(end, step)

    def fm_keys(self, P, cells: np.ndarray, n: int) -> np.ndarray:
        keys = np.zeros(n, np.int32)
        self.lib.orc_fm_keys(_pp(P), _p(cells), n, _p(keys))
        return keys

    def fm_update(self, P, cells: np.ndarray, n: int):
        V = P.N ** 3
        keys = np.zeros(n, np.int32)
        order = np.zeros(n, np.int32)
        s = np.zeros(V, np.int32)
        e = np.zeros(V, np.int32)
        self.lib.orc_fm_update(_pp(P), _p(cells), n, _p(keys), _p(order), _p(s), _p(e))
        return keys, order, s, e

    def forces(self, P, cells: np.ndarray, n: int, s: np.ndarray, e: np.ndarray) -> np.ndarray:
        f = np.zeros((n, 3), np.float32)
        self.lib.orc_forces(_pp(P), _p(cells), n, _p(s), _p(e), _p(f))
        return f

    def move(self, P, cells: np.ndarray, n: int) -> None:
        self.lib.orc_move(_pp(P), _p(cells), n)

    def prelod(self, P, sub: int, cells: np.ndarray, n: int) -> np.ndarray:
        out = np.zeros(P.N ** 3, np.float32)
        self.lib.orc_prelod(_pp(P), sub, _p(cells), n, _p(out))
        return out

    def cell_uptake(self, P, sub: int, cell: np.ndarray) -> float:
        c = np.ascontiguousarray(cell)
        return self.lib.orc_cell_uptake(_pp(P), sub, _p(c))

    def lod_coeffs(self, P, sub: int = 0, f64: bool = False):
        dt = np.float64 if f64 else np.float32
        E = np.zeros(P.N, dt)
        F = np.zeros(P.N, dt)
        (self.lib.orc_lod_coeffs_f64 if f64 else self.lib.orc_lod_coeffs)(_pp(P), sub, _p(E), _p(F))
        return E, F

    def thomas_line(self, p: np.ndarray, E: np.ndarray, F: np.ndarray, bnd: float) -> None:
        self.lib.orc_thomas_line(_p(p), len(p), _p(E), _p(F), bnd)

    def lod_substep(self, P, sub: int, p: np.ndarray, p_rate: np.ndarray) -> None:
        if p.dtype == np.float64:
            self.lib.orc_lod_substep_f64(_pp(P), sub, _p(p), _p(np.ascontiguousarray(p_rate, np.float64)))
        else:
            self.lib.orc_lod_substep(_pp(P), sub, _p(p), _p(np.ascontiguousarray(p_rate, np.float32)))

    def volume_update(self, P, cell: np.ndarray) -> None:
        self.lib.orc_volume_update(_pp(P), _p(cell))

    def phase_update(self, P, cell: np.ndarray, rnd: float, O2: float) -> bool:
        return bool(self.lib.orc_phase_update(_pp(P), _p(cell), rnd, O2))

    def random_dir(self, u: float, v: float) -> np.ndarray:
        out = np.zeros(3, np.float32)
        self.lib.orc_random_dir(u, v, _p(out))
        return out

    def birth(self, P, p0: np.ndarray, cells: np.ndarray, n: int, step: int):
        """cells must hold P.capacity records; returns (new n, slots[n])."""
        assert len(cells) >= P.capacity
        slots = np.zeros(max(n, 1), np.int32)
        p0 = np.ascontiguousarray(p0, np.float32)
        nn = self.lib.orc_birth(_pp(P), _p(p0), _p(cells), n, step, _p(slots))
        return nn, slots[:n]

    def death(self, cells: np.ndarray, n: int) -> int:
        return self.lib.orc_death(_p(cells), n)

    def run(self, P, cells: np.ndarray, n: int, fields: np.ndarray, bio_steps: int, mech_limit: int = -1,
            with_pd: bool = True, first_step: int = 0, times: np.ndarray | None = None) -> int:
        assert len(cells) >= P.capacity
        if fields.dtype == np.float64:
            return self.lib.orc_run_f64(_pp(P), _p(cells), n, _p(fields), bio_steps, mech_limit, int(with_pd),
                                        first_step)
        t = times if times is not None else np.zeros(3, np.float64)
        return self.lib.orc_run(_pp(P), _p(cells), n, _p(fields), bio_steps, mech_limit, int(with_pd), first_step,
                                _p(t))

    def run_mech(self, P, cells: np.ndarray, n: int, fields: np.ndarray, steps: int) -> int:
        """Exactly `steps` mechanics steps without PD; float or double fields."""
        fn = self.lib.orc_run_mech_f64 if fields.dtype == np.float64 else self.lib.orc_run_mech
        return fn(_pp(P), _p(cells), n, _p(fields), steps)
