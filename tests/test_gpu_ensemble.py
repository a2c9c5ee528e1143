"""Ensembles (gell_ensemble_*): every replicate of an ensemble, advanced in
graph mode (one captured graph per biology step, counts device-resident) or
stream mode, is bit-identical to the same replicate run alone through the
per-context calls -- cells, fields and counts after biology steps with
births and deaths. Replicates differ in N, n, substrates and seed, and one
is empty; the benchmark's own ensemble (first 16 replicates, full size) is
checked the same way."""
import numpy as np
import pytest

from gell_amd import runtime as rt
from gell_amd import spheroid
from helpers import log, params

pytestmark = pytest.mark.gpu

# (N, n, substrates)
SPECS = [(32, 300, 1), (48, 1500, 2), (64, 4000, 1), (32, 900, 1), (64, 200, 2), (48, 0, 1)]
BIO_STEPS = 3


def replicate(r, N, n, n_sub):
    P = params(N, capacity=2 * n + 64, n_sub=n_sub, seed=1000 + r)
    radius = 0.3 * N * 10.0
    cells = spheroid.make_cells(n, N, 10.0, radius, 50 + r, capacity=P.capacity)
    if n:
        spheroid.radial_profile(cells, n, (N * 5.0,) * 3, radius, 70 + r)
    return P, cells


def solo(P, cells, n):
    ctx = rt.Context(P)
    buf = rt.DeviceBuffer(P.capacity * 72)
    buf.upload(cells)
    mech = rt.float_loop_count(P.dt_bio, P.dt_mech)
    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
    for b in range(BIO_STEPS):
        n = ctx.birth(buf, n, b)
        for _ in range(mech):
            ctx.fm_update(buf, n)
            for s in range(P.n_substrates):
                ctx.prelod(s, buf, n)
            for _ in range(diff):  # the reference's interleaved substrate order
                for s in range(P.n_substrates):
                    ctx.lod_substep(s)
        n = ctx.death(buf, n)
    out = buf.download(rt.CELL_DTYPE, P.capacity), [ctx.field(s) for s in range(P.n_substrates)], n
    buf.free()
    ctx.close()
    return out


@pytest.fixture(scope="module")
def solo_runs():
    reps = [replicate(r, *spec) for r, spec in enumerate(SPECS)]
    return reps, [solo(P, c, spec[1]) for (P, c), spec in zip(reps, SPECS)]


@pytest.mark.parametrize("mode", [rt.Ensemble.GRAPH, rt.Ensemble.STREAMS])
def test_ensemble_equals_solo_runs(mode, solo_runs):
    reps, want = solo_runs
    ens = rt.Ensemble([P for P, _ in reps])
    ens.set_mode(mode)
    bufs = []
    for P, c in reps:
        b = rt.DeviceBuffer(P.capacity * 72)
        b.upload(c)
        bufs.append(b)
    n = [spec[1] for spec in SPECS]
    for b in range(BIO_STEPS):
        n = ens.bio_step(bufs, n, b)
        log(f"[ensemble mode {mode}] bio step {b}: n = {n}")
    ens.sync()
    stats = ens.stats()
    births = 0
    for r, ((P, c0), (wc, wf, wn)) in enumerate(zip(reps, want)):
        assert n[r] == wn, f"replicate {r}: count {n[r]} != {wn}"
        got = bufs[r].download(rt.CELL_DTYPE, P.capacity)
        assert np.array_equal(got[:wn].view(np.uint8), wc[:wn].view(np.uint8)), f"replicate {r}: cells differ"
        for s in range(P.n_substrates):
            f = ens.members[r].field(s)
            assert np.array_equal(f.view(np.uint32), wf[s].view(np.uint32)), f"replicate {r} substrate {s}"
        births += max(0, wn - SPECS[r][1])
        bufs[r].free()
    ens.close()
    if mode == rt.Ensemble.GRAPH:
        # one graph launch per biology step; captured again only while the
        # host-side state it depends on settles (at most once more)
        assert stats["graph_launches"] == BIO_STEPS and 1 <= stats["graph_captures"] <= 2, stats
    else:
        assert stats["graph_launches"] == 0
    assert births > 0  # the profile has dividing cells, so the counts change between steps


def test_ensemble_graph_follows_knob_changes(solo_runs):
    """Changing a kernel variant on a member invalidates the captured graph."""
    reps, _ = solo_runs
    P, c = reps[2]
    ens = rt.Ensemble([P])
    buf = rt.DeviceBuffer(P.capacity * 72)
    buf.upload(c)
    ens.mech_steps([buf], [SPECS[2][1]], 3)
    ens.members[0].tune("fm_prefetch", 3)
    ens.mech_steps([buf], [SPECS[2][1]], 3)
    ens.sync()
    assert ens.stats() == {"graph_captures": 2, "graph_launches": 4}
    buf.free()
    ens.close()


def test_ensemble_rejects_bad_arguments():
    P0 = params(16)
    P1 = params(16)
    P1.dt_mech = 0.2
    with pytest.raises(rt.GellError):
        rt.Ensemble([P0, P1])
    ens = rt.Ensemble([P0])
    buf = rt.DeviceBuffer(P0.capacity * 72)
    with pytest.raises(rt.GellError):
        ens.mech_steps([buf], [P0.capacity + 1], 1)
    with pytest.raises(rt.GellError):
        ens.set_mode(7)
    buf.free()
    ens.close()


def _solo_steps(P, cells, n, steps, knobs):
    ctx = rt.Context(P)
    for k, v in knobs.items():  # the member's tuned variants (timing may pick others)
        ctx.tune(k, v)
    buf = rt.DeviceBuffer(P.capacity * 72)
    buf.upload(cells)
    mech = rt.float_loop_count(P.dt_bio, P.dt_mech)
    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
    for b in range(steps):
        n = ctx.birth(buf, n, b)
        for _ in range(mech):
            ctx.fm_update(buf, n)
            for s in range(P.n_substrates):
                ctx.prelod(s, buf, n)
            for s in range(P.n_substrates):
                for _ in range(diff):
                    ctx.lod_substep(s)
        n = ctx.death(buf, n)
    out = buf.download(rt.CELL_DTYPE, P.capacity), ctx.field(0), n
    buf.free()
    ctx.close()
    return out


@pytest.mark.timeout(600)
def test_benchmark_ensemble_first_16_replicates_equal_solo_runs():
    """The benchmark's batched workload (n log-uniform in [1k, 200k], grids of
    64-192^3, each replicate its own seed), its first 16 replicates at full
    size with the radial phase profile, two biology steps in graph mode:
    every replicate bit for bit its solo run."""
    import bench

    reps = bench.ensemble_workload(20240611)[:16]
    for r, (P, cells, n) in enumerate(reps):
        c = P.N * 10.0 / 2
        spheroid.radial_profile(cells, n, (c, c, c), 1200.0 * (n / 1e6) ** (1 / 3), 90 + r)
    ens = rt.Ensemble([P for P, _, _ in reps])
    ens.autotune()
    bufs = []
    for P, cells, _ in reps:
        b = rt.DeviceBuffer(P.capacity * 72)
        b.upload(cells)
        bufs.append(b)
    n = [k for _, _, k in reps]
    for b in range(2):
        n = ens.bio_step(bufs, n, b)
        log(f"[ensemble B16] bio step {b}: n = {n}")
    stats = ens.stats()
    assert stats["graph_launches"] == 2, stats
    for r, (P, cells, n0) in enumerate(reps):
        knobs = {k: ens.members[r].tune_get(k) for k in rt.Context.KNOBS}
        wc, wf, wn = _solo_steps(P, cells, n0, 2, knobs)
        assert n[r] == wn, f"replicate {r}: {n[r]} != {wn}"
        got = bufs[r].download(rt.CELL_DTYPE, P.capacity)
        assert np.array_equal(got[:wn].view(np.uint8), wc[:wn].view(np.uint8)), f"replicate {r}: cells"
        assert np.array_equal(ens.members[r].field(0).view(np.uint32), wf.view(np.uint32)), f"replicate {r}: field"
        bufs[r].free()
    log(f"[ensemble B16] {len(reps)} replicates equal their solo runs; graph stats {stats}")
    ens.close()
