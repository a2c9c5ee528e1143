"""LOD substep on the GPU against the oracle (LODsolver, LODkernel.cuh:280-300).

Same seeded field and uptake map on both sides; fields must agree within
1e-5 relative L2 per substep (the stated bar). The serial line kernels keep
the reference's operation order without contraction and are checked
bit-exact; the default partitioned kernels are checked to the bar."""
import numpy as np
import pytest

from gell_amd import runtime as rt
from helpers import GpuState, log, params, rel_l2, ulp_diff

pytestmark = pytest.mark.gpu


def _field_and_rate(N, seed):
    rng = np.random.default_rng(seed)
    V = N ** 3
    p = (20.0 + 18.0 * rng.random(V)).astype(np.float32)
    rate = (-2.5 * rng.random(V) * (rng.random(V) < 0.3)).astype(np.float32)
    return p, rate


def _gpu_substeps(P, p, rate, steps, s=0, exact=True):
    st = GpuState(P, np.zeros(0, rt.CELL_DTYPE), exact=exact)
    st.ctx.set_field(s, p)
    rt.upload_ptr(st.ctx.lib.gell_uptake_field(st.ctx.h, s), rate)
    for _ in range(steps):
        st.ctx.lod_substep(s)
    out = st.ctx.field(s)
    st.close()
    return out


@pytest.mark.parametrize("N", [16, 32, 50, 64, 96, 100, 128, 200])
def test_lod_substeps_match_oracle(N, oracle):
    log(f"[lod] N={N}")
    P = params(N)
    p, rate = _field_and_rate(N, N)
    got = _gpu_substeps(P, p, rate, 3)
    want = p.copy()
    for _ in range(3):
        oracle.lod_substep(P, 0, want, rate)
    e, u = rel_l2(got, want), ulp_diff(got, want)
    log(f"[lod] N={N} relL2={e:.3e} max_ulp={u}")
    assert e <= 1e-5
    assert u == 0


@pytest.mark.parametrize("N", [64, 96, 100])
@pytest.mark.parametrize("plane", [0, 1])
@pytest.mark.parametrize("spill", [0, 1])
def test_lod_serial_plane_and_line_variants(N, spill, plane, oracle):
    """Serial lines at N = 64, 96 and 100 (the reference's Voxel_num): the
    separate x, y, z kernels, or the x + y sweeps per z plane in one
    workgroup, with or without the LDS spill of the z sweep: all bit-exact."""
    P = params(N, dx=25.0)
    p, rate = _field_and_rate(N, N + spill + 2 * plane)
    st = GpuState(P, np.zeros(0, rt.CELL_DTYPE), exact=True)
    st.ctx.tune("lod_yz_lds", spill)
    st.ctx.tune("lod_fused", plane)
    st.ctx.set_field(0, p)
    rt.upload_ptr(st.ctx.lib.gell_uptake_field(st.ctx.h, 0), rate)
    for _ in range(4):
        st.ctx.lod_substep(0)
    got = st.ctx.field(0)
    st.close()
    want = p.copy()
    for _ in range(4):
        oracle.lod_substep(P, 0, want, rate)
    assert ulp_diff(got, want) == 0


def test_lod_256_one_substep(oracle):
    N = 256
    log("[lod] N=256")
    P = params(N)
    p, rate = _field_and_rate(N, 7)
    got = _gpu_substeps(P, p, rate, 1)
    want = p.copy()
    oracle.lod_substep(P, 0, want, rate)
    e, u = rel_l2(got, want), ulp_diff(got, want)
    log(f"[lod] N=256 relL2={e:.3e} max_ulp={u}")
    assert e <= 1e-5
    assert u == 0


@pytest.mark.parametrize("N", [128, 256])
def test_lod_default_variant_within_tolerance(N, oracle):
    # the default (partitioned) kernels: 1e-5 relative L2 per substep is the bar
    P = params(N)
    p, rate = _field_and_rate(N, N + 1)
    got = _gpu_substeps(P, p, rate, 2, exact=False)
    want = p.copy()
    for _ in range(2):
        oracle.lod_substep(P, 0, want, rate)
    e = rel_l2(got, want)
    log(f"[lod] default N={N} relL2={e:.3e}")
    assert e <= 1e-5


def test_lod_second_substrate_uses_its_own_constants(oracle):
    N = 32
    P = params(N, n_sub=2)
    p, rate = _field_and_rate(N, 3)
    got = _gpu_substeps(P, p, rate, 2, s=1)
    want = p.copy()
    for _ in range(2):
        oracle.lod_substep(P, 1, want, rate)
    assert ulp_diff(got, want) == 0
    # and it differs from the oxygen constants' result
    o2 = p.copy()
    oracle.lod_substep(P, 0, o2, rate)
    assert rel_l2(o2, want) > 1e-3


def test_lod_fixed_point_without_decay_or_uptake():
    # lambda = 0, no uptake, field == boundary everywhere: the field stays put.
    N = 64
    P = params(N)
    P.sub[0].decay = 0.0
    p = np.full(N ** 3, 38.0, np.float32)
    got = _gpu_substeps(P, p, np.zeros(N ** 3, np.float32), 5)
    assert np.max(np.abs(got - 38.0)) <= 38.0 * 1e-5
