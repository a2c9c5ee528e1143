"""ctypes binding of the C ABI in ``include/gell_abi.h``.

Python-side mirror of the reference's host interface for the per-step path:
``Context`` owns what ``MechanicsMesh_struct`` and ``Mesh_struct`` own in the
reference (FMkernel.cuh:160-212, LODkernel.cuh:192-319) and exposes the five
operations Main.cu's time loop calls (Main.cu:108-127). Device memory is
handled through the library's own allocation helpers, so no other GPU
binding is needed.

The HIP library is required: if ``gell_amd/_build/libgell_hip.so`` is missing
every entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libgell_hip.so"
MAX_SUBSTRATES = 4
N_PHASES = 8
PHASE_NAMES = ["fm_keys", "fm_sort", "fm_gather", "fm_force_move", "prelod", "lod", "birth", "death"]

# The 72-byte record of include/gell_abi.h (reference Cells.cuh:9-47).
CELL_DTYPE = np.dtype(
    [
        ("sign", "u1"),
        ("pad_", "u1", (3,)),
        ("mesh_idx", "<i4"),
        ("pos", "<f4", (3,)),
        ("force", "<f4", (3,)),
        ("oldforce", "<f4", (3,)),
        ("r", "<f4"),
        ("V", "<f4"),
        ("Vf", "<f4"),
        ("Vcs", "<f4"),
        ("Vns", "<f4"),
        ("phase", "<i4"),
        ("cell_clock", "<f4"),
    ]
)
assert CELL_DTYPE.itemsize == 72


class Substrate(C.Structure):
    _fields_ = [
        ("boundary", C.c_float),
        ("D", C.c_float),
        ("decay", C.c_float),
        ("uptake", C.c_float),
        ("dead_factor", C.c_float),
        ("initial", C.c_float),
    ]


class Params(C.Structure):
    _fields_ = [
        ("N", C.c_int32),
        ("dx", C.c_float),
        ("capacity", C.c_int32),
        ("n_substrates", C.c_int32),
        ("dt_bio", C.c_float),
        ("dt_mech", C.c_float),
        ("dt_diff", C.c_float),
        ("reserved_", C.c_int32),
        ("seed", C.c_uint64),
        ("sub", Substrate * MAX_SUBSTRATES),
    ]


class SurrogateConfig(C.Structure):
    _fields_ = [
        ("N", C.c_int32),
        ("K", C.c_int32),
        ("rate_scale", C.c_float),
        ("out_scale", C.c_float),
        ("check_every", C.c_int32),
        ("seed", C.c_uint64),
    ]


class GellError(RuntimeError):
    pass


_lib = None


def library_path() -> Path:
    return LIB_PATH


def load_library() -> C.CDLL:
    """Load the gfx950 library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise GellError(f"HIP library not built: {LIB_PATH} (run `python -m gell_amd.build`)")
    lib = C.CDLL(str(LIB_PATH))
    vp, i32, i64, u64, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_size_t
    P = C.POINTER
    sig = {
        "gell_default_params": (None, [P(Params)]),
        "gell_create": (C.c_int, [P(Params), vp, P(vp)]),
        "gell_destroy": (None, [vp]),
        "gell_last_error": (C.c_char_p, [vp]),
        "gell_status_string": (C.c_char_p, [C.c_int]),
        "gell_fm_update": (C.c_int, [vp, vp, i32]),
        "gell_mech_step": (C.c_int, [vp, vp, i32]),
        "gell_prelod": (C.c_int, [vp, i32, vp, i32]),
        "gell_lod_substep": (C.c_int, [vp, i32]),
        "gell_birth": (C.c_int, [vp, vp, i32, u64, P(i32)]),
        "gell_death": (C.c_int, [vp, vp, i32, P(i32)]),
        "gell_death_noalloc": (C.c_int, [vp, i32, vp, P(i32)]),
        "gell_field": (vp, [vp, i32]),
        "gell_uptake_field": (vp, [vp, i32]),
        "gell_sorted_keys": (vp, [vp]),
        "gell_sorted_order": (vp, [vp]),
        "gell_voxel_ranges": (vp, [vp]),
        "gell_birth_slots": (vp, [vp]),
        "gell_sync": (C.c_int, [vp]),
        "gell_profile_enable": (C.c_int, [vp, C.c_int]),
        "gell_tune": (C.c_int, [vp, C.c_int, C.c_int]),
        "gell_tune_get": (C.c_int, [vp, C.c_int]),
        "gell_autotune": (C.c_int, [vp, C.c_int, P(C.c_double)]),
        "gell_lod_partition_table": (C.c_int, [P(Params), i32, i32, P(C.c_float), i32]),
        "gell_profile_read": (C.c_int, [vp, P(C.c_double), P(i64), C.c_int]),
        "gell_dev_alloc": (vp, [sz]),
        "gell_dev_free": (None, [vp]),
        "gell_copy_h2d": (C.c_int, [vp, vp, sz]),
        "gell_copy_d2h": (C.c_int, [vp, vp, sz]),
        "gell_copy_d2d": (C.c_int, [vp, vp, sz]),
        "gell_device_sync": (C.c_int, []),
        "gell_stream_copy": (C.c_int, [vp, vp, sz, vp]),
        "gell_device_name": (C.c_int, [C.c_char_p, C.c_int]),
        "gell_float_loop_count": (i32, [C.c_float, C.c_float]),
        "gell_ensemble_create": (C.c_int, [P(Params), i32, vp, P(vp)]),
        "gell_ensemble_destroy": (None, [vp]),
        "gell_ensemble_size": (i32, [vp]),
        "gell_ensemble_member": (vp, [vp, i32]),
        "gell_ensemble_last_error": (C.c_char_p, [vp]),
        "gell_ensemble_set_mode": (C.c_int, [vp, i32]),
        "gell_ensemble_autotune": (C.c_int, [vp, C.c_int]),
        "gell_ensemble_mech_steps": (C.c_int, [vp, P(vp), P(i32), i32]),
        "gell_ensemble_bio_step": (C.c_int, [vp, P(vp), P(i32), u64]),
        "gell_ensemble_sync": (C.c_int, [vp]),
        "gell_ensemble_stats": (C.c_int, [vp, P(i64), P(i64)]),
        "gell_surrogate_default_config": (None, [P(SurrogateConfig)]),
        "gell_surrogate_create": (C.c_int, [P(SurrogateConfig), vp, P(vp)]),
        "gell_surrogate_destroy": (None, [vp]),
        "gell_surrogate_last_error": (C.c_char_p, [vp]),
        "gell_surrogate_param_count": (i64, [vp]),
        "gell_surrogate_get_params": (C.c_int, [vp, vp, i64]),
        "gell_surrogate_set_params": (C.c_int, [vp, vp, i64]),
        "gell_surrogate_predict": (C.c_int, [vp, vp, vp, vp, vp]),
        "gell_surrogate_gradients": (C.c_int, [vp, vp, vp, vp, vp, i32, P(C.c_float), vp]),
        "gell_surrogate_adam": (C.c_int, [vp, C.c_float]),
        "gell_surrogate_advance": (C.c_int, [vp, vp, i32, C.c_float, P(i32), P(C.c_float)]),
        "gell_surrogate_reset_budget": (C.c_int, [vp]),
        "gell_surrogate_gate_stats": (C.c_int, [vp, P(i64), P(i64), P(i64), P(C.c_double), P(C.c_double)]),
        "gell_nn_conv3": (C.c_int, [vp, vp, vp, vp, i32, i32, i32, i32, i32]),
        "gell_nn_conv3_pooled": (C.c_int, [vp, vp, vp, vp, vp, i32, i32, i32, i32, vp]),
        "gell_nn_conv3_wgrad": (C.c_int, [vp, vp, vp, i32, i32, i32]),
        "gell_nn_upconv3": (C.c_int, [vp, vp, vp, vp, vp, i32, i32, i32, i32]),
        "gell_io_last_error": (C.c_char_p, []),
        "gell_sphere_init": (C.c_int, [vp, i32, i32, P(Params), C.c_float, u64]),
        "gell_write_cells_csv": (C.c_int, [vp, i32, C.c_char_p]),
        "gell_write_field_csv": (C.c_int, [vp, i32, C.c_char_p]),
        "gell_write_slice_csv": (C.c_int, [vp, i32, C.c_char_p]),
        "gell_write_counts_csv": (C.c_int, [vp, vp, i32, C.c_char_p]),
        "gell_write_gap_counts_csv": (C.c_int, [vp, vp, i32, i32, C.c_char_p]),
        "gell_checkpoint_save": (C.c_int, [vp, vp, i32, u64, C.c_char_p]),
        "gell_checkpoint_load": (C.c_int, [vp, vp, P(i32), P(u64), C.c_char_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    """Names every build of the library must export (those of gell_abi.h)."""
    hdr = Path(__file__).resolve().parent.parent / "include" / "gell_abi.h"
    import re

    names = []
    for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(gell_\w+)\s*\(", hdr.read_text(), re.M):
        names.append(m.group(1))
    return sorted(set(names))


def default_params(**overrides) -> Params:
    """Params filled with the reference's Definitions.h values, then overrides.

    ``substrates`` may be a list of dicts with Substrate field names.
    """
    p = Params()
    load_library().gell_default_params(C.byref(p))
    subs = overrides.pop("substrates", None)
    for k, v in overrides.items():
        setattr(p, k, v)
    if subs is not None:
        p.n_substrates = len(subs)
        for s, d in enumerate(subs):
            for k, v in d.items():
                setattr(p.sub[s], k, v)
    return p


def make_params(**kw) -> Params:
    """Params without touching the library (for CPU-only code paths)."""
    p = Params()
    p.N, p.dx, p.capacity, p.n_substrates = 100, 25.0, 2_000_000, 1
    p.dt_bio, p.dt_mech, p.dt_diff, p.seed = 6.0, 0.1, 0.01, 123
    p.sub[0].boundary, p.sub[0].D, p.sub[0].decay = 38.0, 1e5, 0.1
    p.sub[0].uptake, p.sub[0].dead_factor, p.sub[0].initial = -10.0, 0.1, 38.0
    subs = kw.pop("substrates", None)
    for k, v in kw.items():
        setattr(p, k, v)
    if subs is not None:
        p.n_substrates = len(subs)
        for s, d in enumerate(subs):
            for k, v in d.items():
                setattr(p.sub[s], k, v)
    return p


class DeviceBuffer:
    """A device allocation made through the library."""

    def __init__(self, nbytes: int):
        self.lib = load_library()
        self.nbytes = int(nbytes)
        self.ptr = self.lib.gell_dev_alloc(max(1, self.nbytes))
        if not self.ptr:
            raise GellError(f"device allocation of {nbytes} bytes failed")

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        if a.nbytes and self.lib.gell_copy_h2d(self.ptr + offset, a.ctypes.data, a.nbytes) != 0:
            raise GellError("host-to-device copy failed")

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        if out.nbytes and self.lib.gell_copy_d2h(out.ctypes.data, self.ptr + offset, out.nbytes) != 0:
            raise GellError("device-to-host copy failed")
        return out

    def free(self) -> None:
        if self.ptr:
            self.lib.gell_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download_ptr(ptr: int, dtype, count: int) -> np.ndarray:
    out = np.empty(count, dtype=dtype)
    if out.nbytes and load_library().gell_copy_d2h(out.ctypes.data, ptr, out.nbytes) != 0:
        raise GellError("device-to-host copy failed")
    return out


def upload_ptr(ptr: int, a: np.ndarray) -> None:
    a = np.ascontiguousarray(a)
    if a.nbytes and load_library().gell_copy_h2d(ptr, a.ctypes.data, a.nbytes) != 0:
        raise GellError("host-to-device copy failed")


class Context:
    """One simulation context: cell-list, fields, uptake maps, scratch."""

    def __init__(self, params: Params, stream: int | None = None):
        self.lib = load_library()
        self.params = params
        h = C.c_void_p()
        st = self.lib.gell_create(C.byref(params), stream, C.byref(h))
        if st != 0:
            msg = self.lib.gell_last_error(None).decode()
            raise GellError(f"gell_create: {msg} ({st})")
        self.h = h.value
        self.N = params.N
        self.V = params.N ** 3

    def _check(self, st: int, what: str) -> None:
        if st != 0:
            raise GellError(f"{what}: {self.lib.gell_last_error(self.h).decode()} ({st})")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.gell_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- the five operations -------------------------------------------------
    def fm_update(self, cells: DeviceBuffer | int, n: int) -> None:
        self._check(self.lib.gell_fm_update(self.h, _ptr(cells), n), "gell_fm_update")

    def mech_step(self, cells: DeviceBuffer | int, n: int) -> None:
        """fm_update followed by prelod of every substrate, the uptake rates
        accumulated by the mechanics step's own force kernel."""
        self._check(self.lib.gell_mech_step(self.h, _ptr(cells), n), "gell_mech_step")

    def prelod(self, s: int, cells: DeviceBuffer | int, n: int) -> None:
        self._check(self.lib.gell_prelod(self.h, s, _ptr(cells), n), "gell_prelod")

    def lod_substep(self, s: int) -> None:
        self._check(self.lib.gell_lod_substep(self.h, s), "gell_lod_substep")

    def birth(self, cells: DeviceBuffer | int, n: int, step: int) -> int:
        out = C.c_int32()
        self._check(self.lib.gell_birth(self.h, _ptr(cells), n, step, C.byref(out)), "gell_birth")
        return out.value

    def death(self, cells: DeviceBuffer | int, n: int) -> int:
        out = C.c_int32()
        self._check(self.lib.gell_death(self.h, _ptr(cells), n, C.byref(out)), "gell_death")
        return out.value

    def sync(self) -> None:
        self._check(self.lib.gell_sync(self.h), "gell_sync")

    # -- state access ----------------------------------------------------------
    def field_ptr(self, s: int) -> int:
        return self.lib.gell_field(self.h, s)

    def field(self, s: int) -> np.ndarray:
        self.sync()
        return download_ptr(self.field_ptr(s), np.float32, self.V)

    def set_field(self, s: int, a: np.ndarray) -> None:
        self.sync()
        upload_ptr(self.field_ptr(s), np.asarray(a, np.float32).ravel())

    def uptake_field(self, s: int) -> np.ndarray:
        self.sync()
        return download_ptr(self.lib.gell_uptake_field(self.h, s), np.float32, self.V)

    def sorted_keys(self, n: int) -> np.ndarray:
        self.sync()
        return download_ptr(self.lib.gell_sorted_keys(self.h), np.int32, n)

    def sorted_order(self, n: int) -> np.ndarray:
        self.sync()
        return download_ptr(self.lib.gell_sorted_order(self.h), np.int32, n)

    def voxel_ranges(self) -> tuple[np.ndarray, np.ndarray]:
        self.sync()
        se = download_ptr(self.lib.gell_voxel_ranges(self.h), np.int32, 2 * self.V).reshape(-1, 2)
        return se[:, 0].copy(), se[:, 1].copy()

    def birth_slots(self, n: int) -> np.ndarray:
        self.sync()
        return download_ptr(self.lib.gell_birth_slots(self.h), np.int32, n)

    # -- restart ---------------------------------------------------------------
    def checkpoint_save(self, cells: DeviceBuffer | int, n: int, bio_step: int, path: str) -> None:
        """Write the state a step depends on: parameters, step counter,
        cells[0, n), every substrate's field."""
        st = self.lib.gell_checkpoint_save(self.h, _ptr(cells), n, bio_step, str(path).encode())
        if st != 0:
            raise GellError(f"gell_checkpoint_save: {self.lib.gell_io_last_error().decode()} ({st})")

    def checkpoint_load(self, cells: DeviceBuffer | int, path: str) -> tuple[int, int]:
        """Restore a checkpoint into this context and `cells`; returns (n, bio_step)."""
        n, step = C.c_int32(), C.c_uint64()
        st = self.lib.gell_checkpoint_load(self.h, _ptr(cells), C.byref(n), C.byref(step), str(path).encode())
        if st != 0:
            raise GellError(f"gell_checkpoint_load: {self.lib.gell_io_last_error().decode()} ({st})")
        return n.value, step.value

    # -- kernel variants -------------------------------------------------------
    KNOBS = {"lod_yz_lds": 0, "fm_prefetch": 1, "lod_parts": 2, "lod_x_parts": 3, "lod_fused": 4,
             "substrate_streams": 5}

    def tune(self, knob: str, value: int) -> None:
        self._check(self.lib.gell_tune(self.h, self.KNOBS[knob], int(value)), "gell_tune")

    def tune_get(self, knob: str) -> int:
        return self.lib.gell_tune_get(self.h, self.KNOBS[knob])

    def autotune(self, bit_exact: bool = False) -> dict:
        """Pick the fastest kernel variants for this context's grid; returns
        the chosen knob values and the best LOD substep time."""
        us = C.c_double()
        self._check(self.lib.gell_autotune(self.h, 1 if bit_exact else 0, C.byref(us)), "gell_autotune")
        out = {k: self.tune_get(k) for k in self.KNOBS}
        out["lod_us"] = us.value
        return out

    # -- profiling -------------------------------------------------------------
    def profile(self, enable: bool = True) -> None:
        self._check(self.lib.gell_profile_enable(self.h, int(enable)), "gell_profile_enable")

    def profile_read(self, reset: bool = True) -> dict:
        ms = (C.c_double * N_PHASES)()
        calls = (C.c_int64 * N_PHASES)()
        self._check(self.lib.gell_profile_read(self.h, ms, calls, int(reset)), "gell_profile_read")
        return {PHASE_NAMES[q]: (ms[q], calls[q]) for q in range(N_PHASES)}


class _Member(Context):
    """A context owned by an Ensemble (not destroyed on its own)."""

    def __init__(self, handle: int, params: Params):
        self.lib = load_library()
        self.params = params
        self.h = handle
        self.N = params.N
        self.V = params.N ** 3

    def close(self) -> None:
        self.h = None


class Ensemble:
    """Independent replicates advanced together (gell_ensemble_* of the C
    ABI): one context per replicate on its own stream, mechanics steps
    replayed as one hipGraph per set of cell counts."""

    STREAMS, GRAPH = 0, 1

    def __init__(self, params: list[Params], stream: int | None = None):
        self.lib = load_library()
        self.params = list(params)
        arr = (Params * len(params))(*params)
        h = C.c_void_p()
        st = self.lib.gell_ensemble_create(arr, len(params), stream, C.byref(h))
        if st != 0:
            raise GellError(f"gell_ensemble_create: {self.lib.gell_ensemble_last_error(None).decode()} ({st})")
        self.h = h.value
        self.members = [_Member(self.lib.gell_ensemble_member(self.h, r), p) for r, p in enumerate(params)]

    def __len__(self) -> int:
        return len(self.members)

    def _check(self, st: int, what: str) -> None:
        if st != 0:
            raise GellError(f"{what}: {self.lib.gell_ensemble_last_error(self.h).decode()} ({st})")

    def set_mode(self, mode: int) -> None:
        self._check(self.lib.gell_ensemble_set_mode(self.h, mode), "gell_ensemble_set_mode")

    def autotune(self, bit_exact: bool = False) -> None:
        self._check(self.lib.gell_ensemble_autotune(self.h, 1 if bit_exact else 0), "gell_ensemble_autotune")

    @staticmethod
    def _arrays(cells, n):
        ptrs = (C.c_void_p * len(cells))(*[_ptr(c) for c in cells])
        counts = (C.c_int32 * len(n))(*[int(x) for x in n])
        return ptrs, counts

    def mech_steps(self, cells, n, steps: int) -> None:
        ptrs, counts = self._arrays(cells, n)
        self._check(self.lib.gell_ensemble_mech_steps(self.h, ptrs, counts, steps), "gell_ensemble_mech_steps")

    def bio_step(self, cells, n, step: int) -> list[int]:
        ptrs, counts = self._arrays(cells, n)
        self._check(self.lib.gell_ensemble_bio_step(self.h, ptrs, counts, step), "gell_ensemble_bio_step")
        return list(counts)

    def sync(self) -> None:
        self._check(self.lib.gell_ensemble_sync(self.h), "gell_ensemble_sync")

    def stats(self) -> dict:
        cap, lau = C.c_int64(), C.c_int64()
        self._check(self.lib.gell_ensemble_stats(self.h, C.byref(cap), C.byref(lau)), "gell_ensemble_stats")
        return {"graph_captures": cap.value, "graph_launches": lau.value}

    def close(self) -> None:
        if getattr(self, "h", None):
            for m in self.members:
                m.close()
            self.lib.gell_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Surrogate:
    """The LOD surrogate network (gell_surrogate_* of the C ABI)."""

    def __init__(self, N: int, K: int = 50, stream: int | None = None, **cfg):
        self.lib = load_library()
        c = SurrogateConfig()
        self.lib.gell_surrogate_default_config(C.byref(c))
        c.N, c.K = N, K
        for k, v in cfg.items():
            setattr(c, k, v)
        self.cfg = c
        h = C.c_void_p()
        st = self.lib.gell_surrogate_create(C.byref(c), stream, C.byref(h))
        if st != 0:
            raise GellError(f"gell_surrogate_create: {self.lib.gell_surrogate_last_error(None).decode()} ({st})")
        self.h = h.value
        self.N = N
        self.nparam = self.lib.gell_surrogate_param_count(self.h)

    def _check(self, st: int, what: str) -> None:
        if st != 0:
            raise GellError(f"{what}: {self.lib.gell_surrogate_last_error(self.h).decode()} ({st})")

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparam, np.float32)
        self._check(self.lib.gell_surrogate_get_params(self.h, out.ctypes.data, self.nparam), "get_params")
        return out

    def set_params(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, np.float32)
        self._check(self.lib.gell_surrogate_set_params(self.h, a.ctypes.data, a.size), "set_params")

    def predict(self, field, field1, rate, out) -> None:
        """out = the field K substeps after `field` under `rate`; field1 is
        `field` after one real substep under it (device pointers)."""
        self._check(self.lib.gell_surrogate_predict(self.h, _ptr(field), _ptr(field1), _ptr(rate), _ptr(out)),
                    "predict")

    def gradients(self, field, field1, rate, target, want_grad: bool = False, accumulate: bool = False):
        loss = C.c_float()
        g = np.zeros(self.nparam, np.float32) if want_grad else None
        self._check(self.lib.gell_surrogate_gradients(self.h, _ptr(field), _ptr(field1), _ptr(rate), _ptr(target),
                                                      int(accumulate), C.byref(loss),
                                                      g.ctypes.data if want_grad else None),
                    "gradients")
        return (loss.value, g) if want_grad else loss.value

    def adam(self, lr: float) -> None:
        self._check(self.lib.gell_surrogate_adam(self.h, lr), "adam")

    def advance(self, ctx: "Context", s: int, tol: float = 1e-3) -> tuple[bool, float]:
        used, res = C.c_int32(), C.c_float()
        self._check(self.lib.gell_surrogate_advance(self.h, ctx.h, s, tol, C.byref(used), C.byref(res)), "advance")
        return bool(used.value), res.value

    def reset_budget(self) -> None:
        self._check(self.lib.gell_surrogate_reset_budget(self.h), "reset_budget")

    def gate_stats(self) -> dict:
        c, a, f = C.c_int64(), C.c_int64(), C.c_int64()
        r, sp = C.c_double(), C.c_double()
        self._check(self.lib.gell_surrogate_gate_stats(self.h, C.byref(c), C.byref(a), C.byref(f), C.byref(r),
                                                       C.byref(sp)), "gate_stats")
        return {"checked": c.value, "accepted": a.value, "fallback": f.value, "ratio": r.value, "spent": sp.value}

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.gell_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nn_conv3(x: DeviceBuffer, w: np.ndarray, b: np.ndarray, y: DeviceBuffer, cin: int, cout: int, S: int,
             relu: bool = False, input_grad: bool = False) -> None:
    """One network layer alone (x, y device BF16 tensors, w [cout][cin][27], b[cout] host FP32)."""
    w = np.ascontiguousarray(w, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    st = load_library().gell_nn_conv3(_ptr(x), w.ctypes.data, b.ctypes.data, _ptr(y), cin, cout, S, int(relu),
                                      int(input_grad))
    if st != 0:
        raise GellError(f"gell_nn_conv3: {load_library().gell_surrogate_last_error(None).decode()} ({st})")


def nn_conv3_pooled(x: DeviceBuffer, w: np.ndarray, b: np.ndarray, y: DeviceBuffer, pooled: DeviceBuffer, cin: int,
                    cout: int, S: int, relu: bool = True) -> bool:
    """nn_conv3 that also writes the 2x2x2 pooling of y; True when the
    convolution's own epilogue formed it."""
    w = np.ascontiguousarray(w, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    flag = C.c_int32(0)
    st = load_library().gell_nn_conv3_pooled(_ptr(x), w.ctypes.data, b.ctypes.data, _ptr(y), _ptr(pooled), cin, cout,
                                             S, int(relu), C.byref(flag))
    if st != 0:
        raise GellError(f"gell_nn_conv3_pooled: {load_library().gell_surrogate_last_error(None).decode()} ({st})")
    return bool(flag.value)


def nn_upconv3(x: DeviceBuffer, w: np.ndarray, b: np.ndarray, y: DeviceBuffer, cin: int, cout: int, S: int,
               relu: bool = True, skip: DeviceBuffer | None = None) -> None:
    """A layer after an upsampling, on x's S^3 grid (y is (2S)^3 x cout)."""
    w = np.ascontiguousarray(w, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    st = load_library().gell_nn_upconv3(_ptr(x), w.ctypes.data, b.ctypes.data, _ptr(skip) if skip else None, _ptr(y),
                                        cin, cout, S, int(relu))
    if st != 0:
        raise GellError(f"gell_nn_upconv3: {load_library().gell_surrogate_last_error(None).decode()} ({st})")


def nn_conv3_wgrad(x: DeviceBuffer, dy: DeviceBuffer, cin: int, cout: int, S: int) -> np.ndarray:
    out = np.zeros((cout, cin, 27), np.float32)
    st = load_library().gell_nn_conv3_wgrad(_ptr(x), _ptr(dy), out.ctypes.data, cin, cout, S)
    if st != 0:
        raise GellError(f"gell_nn_conv3_wgrad: {load_library().gell_surrogate_last_error(None).decode()} ({st})")
    return out


def _ptr(x) -> int:
    return x.ptr if isinstance(x, DeviceBuffer) else int(x)


def float_loop_count(end: float, step: float) -> int:
    """Trip count of ``for (float t = 0; t < end; t += step)`` (Main.cu:108,114,121)."""
    t = np.float32(0.0)
    end, step = np.float32(end), np.float32(step)
    c = 0
    while t < end:
        c += 1
        t = np.float32(t + step)
    return c


def lod_partition_table(params: Params, substrate: int, parts: int) -> np.ndarray:
    """Host-built coefficient table of the partitioned line solve."""
    n = 4 * params.N + 4 * parts * parts
    out = np.zeros(n, np.float32)
    st = load_library().gell_lod_partition_table(C.byref(params), substrate, parts,
                                                 out.ctypes.data_as(C.POINTER(C.c_float)), n)
    if st != 0:
        raise GellError(f"gell_lod_partition_table: {st}")
    return out


def device_name() -> str:
    buf = C.create_string_buffer(256)
    if load_library().gell_device_name(buf, 256) != 0:
        raise GellError("no device")
    return buf.value.decode()


# -- initial state and output files (host arrays) -------------------------------
def sphere_init(params: Params, n: int, capacity: int | None = None, cell_radius: float = 8.4,
                seed: int = 12345) -> np.ndarray:
    """Cell_sphere_Initialization (Initialization.cuh:15-43): `capacity`
    records, the first n quiescent cells uniformly in the ball of radius
    cell_radius n^(1/3) / 1.612 * 1.87 at the domain centre (SplitMix64 draws)."""
    cap = params.capacity if capacity is None else capacity
    cells = np.zeros(cap, CELL_DTYPE)
    st = load_library().gell_sphere_init(cells.ctypes.data, cap, n, C.byref(params), cell_radius, seed)
    _io_check(st, "gell_sphere_init")
    return cells



def _io_check(st: int, what: str) -> None:
    if st != 0:
        raise GellError(f"{what}: {load_library().gell_io_last_error().decode()} ({st})")


def savecsv_cell(cells: np.ndarray, n: int, path: str) -> None:
    """savecsv_cell (FileSave.cuh:24-33) of records cells[0, n) (CELL_DTYPE)."""
    c = np.ascontiguousarray(cells[:n], dtype=CELL_DTYPE)
    _io_check(load_library().gell_write_cells_csv(c.ctypes.data if n else None, n, str(path).encode()),
              "gell_write_cells_csv")


def savecsv_mesh(field: np.ndarray, N: int, path: str) -> None:
    """savecsv_mesh (FileSave.cuh:35-49) of an N^3 field (x fastest)."""
    f = np.ascontiguousarray(field, dtype=np.float32).ravel()
    assert f.size == N ** 3
    _io_check(load_library().gell_write_field_csv(f.ctypes.data, N, str(path).encode()), "gell_write_field_csv")


def savecsv_meshslice(field: np.ndarray, N: int, path: str) -> None:
    """savecsv_meshslice (FileSave.cuh:50-64) of the plane z = N / 2."""
    f = np.ascontiguousarray(field, dtype=np.float32).ravel()
    assert f.size == N ** 3
    _io_check(load_library().gell_write_slice_csv(f.ctypes.data, N, str(path).encode()), "gell_write_slice_csv")


def savecsv_array(counts: np.ndarray, times: np.ndarray, path: str) -> None:
    """savecsv_array (FileSave.cuh:66-75): one row per entry."""
    c = np.ascontiguousarray(counts, dtype=np.int32)
    t = np.ascontiguousarray(times, dtype=np.float64)
    assert c.size == t.size
    _io_check(load_library().gell_write_counts_csv(c.ctypes.data, t.ctypes.data, c.size, str(path).encode()),
              "gell_write_counts_csv")


def savecsv_gaparray(counts: np.ndarray, gap_times: np.ndarray, n_iter: int, gap: int, path: str) -> None:
    """savecsv_gaparray (FileSave.cuh:77-86): rows i < n_iter // gap of
    counts[(i + 1) gap - 1] and gap_times[i]."""
    c = np.ascontiguousarray(counts, dtype=np.int32)
    t = np.ascontiguousarray(gap_times, dtype=np.float64)
    rows = n_iter // gap
    assert c.size >= rows * gap and t.size >= rows
    _io_check(load_library().gell_write_gap_counts_csv(c.ctypes.data, t.ctypes.data, n_iter, gap,
                                                        str(path).encode()), "gell_write_gap_counts_csv")
