"""Pins of the oracle's building blocks: Morton keys, Philox4x32-10, the
uniform map, the float loop counts of Main.cu, and the ABI library's exports."""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

from gell_amd import runtime as rt

GOLDEN = Path(__file__).resolve().parent / "golden"


def morton_numpy(i, j, k):
    """Independent vectorised bit interleave (x -> bit 0, y -> 1, z -> 2)."""
    i, j, k = (np.asarray(a, np.int64) for a in (i, j, k))
    key = np.zeros(np.broadcast(i, j, k).shape, np.int64)
    for b in range(10):
        key |= ((i >> b) & 1) << (3 * b)
        key |= ((j >> b) & 1) << (3 * b + 1)
        key |= ((k >> b) & 1) << (3 * b + 2)
    return key


def test_morton_known_values(oracle):
    assert oracle.morton(0, 0, 0) == 0
    assert oracle.morton(1, 0, 0) == 1
    assert oracle.morton(0, 1, 0) == 2
    assert oracle.morton(0, 0, 1) == 4
    assert oracle.morton(2, 0, 0) == 8
    assert oracle.morton(0, 0, 2) == 32
    # i=011, j=101, k=110: triples (k j i) from bit 2 down: 110 101 011
    assert oracle.morton(3, 5, 6) == 0b110_101_011
    assert oracle.morton(1023, 1023, 1023) == (1 << 30) - 1


def test_morton_bijection_and_octants(oracle):
    n = 32
    g = np.arange(n)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    keys = np.array([oracle.morton(int(a), int(b), int(c)) for a, b, c in
                     zip(I.ravel()[::97], J.ravel()[::97], K.ravel()[::97])])
    assert np.array_equal(keys, morton_numpy(I.ravel()[::97], J.ravel()[::97], K.ravel()[::97]))
    full = morton_numpy(I, J, K).ravel()
    assert len(np.unique(full)) == n ** 3 and full.max() == n ** 3 - 1
    # Z-order: the 8 cells of every aligned 2x2x2 block have 8 consecutive keys.
    blk = (I // 2) + n * ((J // 2) + n * (K // 2))
    order = np.argsort(full)
    b = blk.ravel()[order].reshape(-1, 8)
    assert np.all(b == b[:, :1])
    for key in (0, 1, 7, 12345, (1 << 30) - 1):
        assert oracle.morton(*oracle.morton_decode(key)) == key


def test_philox_random123_kat(oracle):
    # Published Random123 known-answer vectors for philox4x32-10.
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(oracle.philox(ctr, key)) == want


def test_philox_rocrand_engine_vectors(oracle):
    for v in json.loads((GOLDEN / "philox_rocrand_kat.json").read_text()):
        assert list(oracle.philox(v["ctr"], v["key"])) == v["out"], v


def test_uniforms_in_half_open_unit_interval(oracle):
    us = np.array([oracle.birth_uniforms(123, s, t) for s in range(200) for t in range(3)])
    assert us.min() > 0.0 and us.max() <= 1.0
    # u = ((w >> 8) + 1) / 2^24, so 2^24 * u is an integer in [1, 2^24].
    q = us.astype(np.float64) * 2 ** 24
    assert np.all(q == np.round(q))
    # the three words of one block are the three draws
    w = oracle.philox([5, 9, 0, 0], [123, 0])
    u = oracle.birth_uniforms(123, 5, 9)
    assert np.array_equal(u, ((w[:3] >> 8) + 1).astype(np.float32) / np.float32(2 ** 24))
    # mean/variance of U(0,1]
    big = np.array([oracle.birth_uniforms(7, s, 0)[0] for s in range(4000)])
    assert abs(big.mean() - 0.5) < 0.02 and abs(big.var() - 1 / 12) < 0.01


def test_reference_loop_counts(oracle):
    # Main.cu:108,114,121 accumulate float loop variables.
    assert oracle.float_loop_count(19.0 * 24 * 60, 6.0) == 4560
    assert oracle.float_loop_count(6.0, 0.1) == 61
    assert oracle.float_loop_count(0.1, 0.01) == 11
    assert rt.float_loop_count(6.0, 0.1) == 61 and rt.float_loop_count(0.1, 0.01) == 11


def test_abi_library_exports_header_symbols():
    lib = rt.load_library()
    names = rt.exported_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert ctypes.sizeof(rt.Params) == 136
    assert rt.CELL_DTYPE.itemsize == 72
    off = {k: rt.CELL_DTYPE.fields[k][1] for k in rt.CELL_DTYPE.names}
    assert (off["sign"], off["mesh_idx"], off["pos"], off["force"], off["oldforce"], off["r"], off["V"], off["Vf"],
            off["Vcs"], off["Vns"], off["phase"], off["cell_clock"]) == (0, 4, 8, 20, 32, 44, 48, 52, 56, 60, 64, 68)


def test_default_params_are_reference_definitions():
    p = rt.default_params()
    assert (p.N, p.dx, p.capacity, p.n_substrates) == (100, 25.0, 2_000_000, 1)
    assert (p.dt_bio, np.float32(p.dt_mech), np.float32(p.dt_diff)) == (6.0, np.float32(0.1), np.float32(0.01))
    assert (p.sub[0].boundary, p.sub[0].D, np.float32(p.sub[0].decay), p.sub[0].uptake,
            np.float32(p.sub[0].dead_factor)) == (38.0, 1e5, np.float32(0.1), -10.0, np.float32(0.1))
