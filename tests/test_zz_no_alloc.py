"""gell_death_noalloc (the context-free death that allocates nothing; its
scratch is a static device array of the library) against gell_death, which
uses a context's buffers: the same records and count, byte for byte, on two
calls in a row (the second on fewer cells, so the array still holds the
first call's data), and a clear error above the stated maximum."""
import ctypes as C

import numpy as np
import pytest

from gell_amd import runtime as rt
from helpers import GpuState, log, mixed_cells, params

pytestmark = pytest.mark.gpu

MAX_CELLS = 131072  # GELL_DEATH_NOALLOC_MAX_CELLS


def _noalloc(lib, buf, n):
    out = C.c_int32(-1)
    st = lib.gell_death_noalloc(buf.ptr, n, None, C.byref(out))
    return st, out.value


@pytest.mark.parametrize("sizes", [(MAX_CELLS, 777), (50_000, 1)])
def test_death_noalloc_matches_death(sizes):
    lib = rt.load_library()
    for k, n in enumerate(sizes):
        log(f"[noalloc] call {k}: n={n}")
        P = params(32, capacity=n)
        cells = mixed_cells(32, 10.0, n, n, seed=100 + n, dead_frac=0.3, outside=0)
        ref = GpuState(P, cells)
        want_n = ref.ctx.death(ref.buf, n)
        want = ref.cells()[:want_n].copy()
        ref.close()
        buf = rt.DeviceBuffer(n * 72)
        buf.upload(cells[:n])
        st, got_n = _noalloc(lib, buf, n)
        assert st == 0, lib.gell_last_error(None).decode()
        got = buf.download(rt.CELL_DTYPE, n)
        buf.free()
        assert got_n == want_n == int(cells["sign"][:n].sum())
        assert np.array_equal(got[:got_n].view(np.uint8), want.view(np.uint8))


def test_death_noalloc_limits():
    lib = rt.load_library()
    buf = rt.DeviceBuffer(72)
    st, got_n = _noalloc(lib, buf, 0)
    assert st == 0 and got_n == 0
    st, _ = _noalloc(lib, buf, MAX_CELLS + 1)
    assert st == -1 and "GELL_DEATH_NOALLOC_MAX_CELLS" in lib.gell_last_error(None).decode()
    st, _ = _noalloc(lib, buf, -1)
    assert st == -1
    buf.free()
