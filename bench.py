#!/usr/bin/env python3
"""Throughput of the Gell per-step hot path on one MI355X.

Workload (default, "L"): 256^3 voxels at 10 um, 1,000,000 cells seeded
uniformly in a 1.2 mm-radius ball at the domain centre (radii 4.5-5.5 um),
two substrates (oxygen and glucose). One timed step is one biology step of the
reference's time loop (Main.cu:108-127):

    birth -> 61 x ( mechanics step, uptake map per substrate,
                    11 x LOD substep per substrate ) -> death

with the trip counts the reference's float loops produce. Inputs are resident
in HBM before the timed region; W untimed steps run first.

Output: one JSON line on stdout with agent-steps/s (cells x mechanics steps
per second) as `value`, voxel-updates/s (voxels x substrates x LOD substeps
per second) beside it, the roofline of the dominant kernel (the LOD substep,
timed live with HIP events on the library's stream) and the CPU oracle timed
on this host's cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "agent-steps/s & voxel-updates/s, 1M cells/256^3; surrogate vs LOD at 1e-3 L2"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per voxel-update of one LOD substep: three sweeps
# x 12 (p, p_rate in; p out) + y 8 + z 8 = 28. The roofline is priced on these
# 28 B whatever the variant; the fused substep (x sweep and y blocks in one
# pass, y coupling applied inside the z sweep) moves only 12 + 8 = 20 of them,
# which the counted traffic shows.
LOD_BYTES_PER_VOXEL = 28
LOD_BYTES_PER_VOXEL_FUSED = 20
# Algorithmic bytes per agent-step of the force + movement kernel: the sorted
# 72-byte record read once and written back once (neighbour reads are
# expected to hit the caches and are not charged).
FORCE_BYTES_PER_CELL = 144

# Substrate constants: oxygen as in the reference's Definitions.h:21-25;
# glucose with a smaller diffusivity, no decay and its own boundary value.
SUBSTRATES = [
    dict(boundary=38.0, D=1.0e5, decay=0.1, uptake=-10.0, dead_factor=0.1, initial=38.0),
    dict(boundary=5.0, D=4.0e4, decay=0.0, uptake=-1.0, dead_factor=0.1, initial=5.0),
]

CONFIGS = {
    "L": dict(N=256, dx=10.0, cells=1_000_000, radius=1200.0, n_sub=2, capacity=2_000_000),
    "S": dict(N=64, dx=10.0, cells=10_000, radius=250.0, n_sub=1, capacity=20_000),
    # heterogeneous spheroid: ~1 cell per voxel (26 neighbours in the stencil),
    # proliferating rim, quiescent shell, necrotic core
    "H": dict(N=512, dx=10.0, cells=10_000_000, radius=1336.0, n_sub=1, capacity=12_000_000, profile=True),
}


# Trained parameters of the LOD surrogate (K = 50 substeps, oxygen).
SURROGATE_WEIGHTS = ROOT / "gell_amd" / "weights" / "lod_surrogate_k50.npy"

# Batched ensemble: 64 replicates, n log-uniform in [1k, 200k], each in the
# smallest grid of {64, 96, 128, 192}^3 (10 um voxels) that holds its ball
# (density of the 1M-cell / 1.2 mm ball) with a 50 um margin; seed + r.
ENSEMBLE = dict(replicates=64, n_lo=1_000, n_hi=200_000, grids=(64, 96, 128, 192), dx=10.0)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_workload(cfg: dict, seed: int):
    from gell_amd import runtime as rt
    from gell_amd import spheroid

    P = rt.make_params(N=cfg["N"], dx=cfg["dx"], capacity=cfg["capacity"], seed=seed,
                       substrates=SUBSTRATES[: cfg["n_sub"]])
    cells = spheroid.make_cells(cfg["cells"], cfg["N"], cfg["dx"], cfg["radius"], seed, capacity=cfg["capacity"])
    if cfg.get("profile"):
        c = cfg["N"] * cfg["dx"] / 2.0
        spheroid.radial_profile(cells, cfg["cells"], (c, c, c), cfg["radius"], seed + 1)
    return P, cells


def read_traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, if present."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def gpu_leg(cfg, P, cells, steps, warmup, graph=False, knobs=None, fused=True):
    """Timed biology steps of the workload. With graph, the steps run through
    the library's graph path (a one-replicate ensemble); otherwise through the
    per-step calls, each mechanics step with its uptake maps as one
    gell_mech_step (fused) or as gell_fm_update + gell_prelod per substrate."""
    from gell_amd import runtime as rt

    n = cfg["cells"]
    ens = rt.Ensemble([P]) if graph else None
    ctx = ens.members[0] if graph else rt.Context(P)
    buf = rt.DeviceBuffer(P.capacity * rt.CELL_DTYPE.itemsize)
    buf.upload(cells)
    mech = rt.float_loop_count(P.dt_bio, P.dt_mech)
    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)

    tuned = ctx.autotune()
    log(f"[bench] autotuned: {tuned}")
    for k, v in (knobs or {}).items():  # --knob overrides the autotuner's choice
        ctx.tune(k, v)
        tuned[k] = v

    # The substrates do not interact inside a mechanics step (each substep
    # reads only its own field and uptake map), so running all substeps of
    # one substrate before the next gives the same fields as the reference's
    # interleaved order while keeping one field's working set in cache.
    def mech_and_maps(n):
        if fused:
            ctx.mech_step(buf, n)
        else:
            ctx.fm_update(buf, n)
            for s in range(P.n_substrates):
                ctx.prelod(s, buf, n)

    def bio_step(b, n):
        if graph:
            agents = n * mech  # counts are fixed between birth and death
            m = ens.bio_step([buf], [n], b)[0]
            return m, agents
        n = ctx.birth(buf, n, b)
        agents = 0
        for _ in range(mech):
            mech_and_maps(n)
            for s in range(P.n_substrates):
                for _ in range(diff):
                    ctx.lod_substep(s)
            agents += n
        n = ctx.death(buf, n)
        return n, agents

    for w in range(warmup):
        n, _ = bio_step(w, n)
        log(f"[bench] warmup step {w} done, n={n}")
    ctx.sync()
    agents = 0
    t0 = time.perf_counter()
    for k in range(steps):
        n, a = bio_step(warmup + k, n)
        agents += a
    ctx.sync()
    t1 = time.perf_counter()
    log(f"[bench] timed {steps} steps in {t1 - t0:.3f} s, n={n}")
    # Per-phase device times come from a separate, untimed pass with the
    # library's event profiling on (which also runs every substrate on the
    # context stream, so each substep is timed alone): mechanics steps of the
    # same state as timed, then one more biology step for birth and death;
    # then mechanics steps with separate uptake maps, which time the force +
    # movement kernel alone.
    ctx.profile(True)
    ctx.profile_read(reset=True)
    n_prof = n  # cells of each profiled mechanics step
    for _ in range(min(mech, 10)):
        mech_and_maps(n)
        for s in range(P.n_substrates):
            for _ in range(diff):
                ctx.lod_substep(s)
    n = ctx.death(buf, ctx.birth(buf, n, warmup + steps))
    prof = ctx.profile_read(reset=True)
    n_alone = n
    for _ in range(5):
        ctx.fm_update(buf, n)
        for s in range(P.n_substrates):
            ctx.prelod(s, buf, n)
    prof_alone = ctx.profile_read(reset=True)
    ctx.profile(False)
    out = dict(seconds=t1 - t0, agents=agents, n_final=n, mech=mech, diff=diff, prof=prof,
               device=rt.device_name(), tuned=tuned, n_prof=n_prof, prof_alone=prof_alone, n_alone=n_alone,
               fused=fused)
    buf.free()
    (ens if graph else ctx).close()
    return out


def cpu_leg(cfg, P, cells, budget_s):
    """The OpenMP oracle on this host's cores: whole mechanics steps (FM,
    uptake maps, 11 LOD substeps per substrate) of the same workload, until
    about budget_s seconds have been spent."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle

    orc = pyoracle.Oracle(omp=True)
    n = cfg["cells"]
    V = cfg["N"] ** 3
    fields = np.concatenate([np.full(V, P.sub[s].initial, np.float32) for s in range(P.n_substrates)])
    work = cells.copy()
    times = np.zeros(3, np.float64)
    t0 = time.perf_counter()
    orc.run(P, work, n, fields, 1, mech_limit=1, with_pd=False, times=times)
    t1 = time.perf_counter()
    per = t1 - t0
    more = max(0, int(budget_s / max(per, 1e-3)) - 1)
    if more:
        orc.run(P, work, n, fields, 1, mech_limit=more, with_pd=False, times=times)
    t2 = time.perf_counter()
    steps = 1 + more
    secs = t2 - t0
    from gell_amd import runtime as rt

    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
    return dict(
        value=n * steps / secs,
        voxel_updates_per_s=V * P.n_substrates * diff * steps / secs,
        seconds=secs,
        steps=steps,
        threads=orc.threads(),
        phase_ms=dict(fm=times[0], prelod=times[1], lod=times[2]),
    )


def only_leg(cfg, P, cells, what: str, steps: int, warmup: int, knobs: dict | None = None) -> dict:
    """One kernel group alone, for profiler runs: `lod` (LOD substeps of
    substrate 0 after one mechanics step + uptake map), `fm` (mechanics steps),
    `mech` (mechanics steps with the uptake maps accumulated by the force
    kernel, gell_mech_step), `mech_unfused` (mechanics step + one uptake map
    per substrate, the calls gell_mech_step stands for),
    `surrogate` (predictions of the committed LOD surrogate on that field and
    uptake map) or `copy` (the streaming copy kernel over 256 MiB, the
    byte-counter calibration)."""
    from gell_amd import runtime as rt

    n = cfg["cells"]
    ctx = rt.Context(P)
    for k, v in (knobs or {}).items():
        ctx.tune(k, v)
    buf = rt.DeviceBuffer(P.capacity * rt.CELL_DTYPE.itemsize)
    buf.upload(cells)
    ctx.fm_update(buf, n)
    for s in range(P.n_substrates):
        ctx.prelod(s, buf, n)
    ctx.sync()
    out = {"what": what, "steps": steps}
    if what == "surrogate":
        sur = rt.Surrogate(cfg["N"], 50)
        sur.set_params(np.load(SURROGATE_WEIGHTS, allow_pickle=False))
        pred = rt.DeviceBuffer(cfg["N"] ** 3 * 4)
        p0 = rt.DeviceBuffer(cfg["N"] ** 3 * 4)
        ctx.sync()
        ctx.lib.gell_copy_d2d(p0.ptr, ctx.field_ptr(0), cfg["N"] ** 3 * 4)
        ctx.lod_substep(0)
        rate_ptr = ctx.lib.gell_uptake_field(ctx.h, 0)
        for k in range(warmup + steps):
            if k == warmup:
                ctx.sync()
                t0 = time.perf_counter()
            sur.predict(p0, ctx.field_ptr(0), rate_ptr, pred)
        ctx.sync()
        out["predict_ms"] = (time.perf_counter() - t0) / max(steps, 1) * 1e3
        p0.free()
        pred.free()
        sur.close()
    elif what == "copy":
        nbytes = 256 << 20
        a, b = rt.DeviceBuffer(nbytes), rt.DeviceBuffer(nbytes)
        for k in range(warmup + steps):
            if k == warmup:
                ctx.lib.gell_device_sync()
                t0 = time.perf_counter()
            ctx.lib.gell_stream_copy(b.ptr, a.ptr, nbytes, None)
        ctx.lib.gell_device_sync()
        dt = (time.perf_counter() - t0) / max(steps, 1)
        out["bytes_per_launch_each_way"] = nbytes
        # host clock around back-to-back launches: the achievable streaming rate
        out["copy_us"] = dt * 1e6
        out["copy_GBps"] = 2 * nbytes / dt / 1e9
        a.free()
        b.free()
    else:
        for k in range(warmup + steps):
            if k == warmup:
                ctx.sync()
                ctx.profile(True)
                ctx.profile_read(reset=True)
            if what == "lod":
                ctx.lod_substep(0)
            elif what == "mech":
                ctx.mech_step(buf, n)
            elif what == "mech_unfused":
                ctx.fm_update(buf, n)
                for s in range(P.n_substrates):
                    ctx.prelod(s, buf, n)
            else:
                ctx.fm_update(buf, n)
        prof = ctx.profile_read(reset=True)
        out["phase_us"] = {k: v[0] / max(v[1], 1) * 1e3 for k, v in prof.items() if v[1]}
        out["phase_calls"] = {k: v[1] for k, v in prof.items() if v[1]}
    buf.free()
    ctx.close()
    return out


def ensemble_workload(seed: int):
    from gell_amd import runtime as rt
    from gell_amd import spheroid

    E = ENSEMBLE
    rng = np.random.default_rng(seed)
    reps = []
    for r in range(E["replicates"]):
        n = int(round(np.exp(rng.uniform(np.log(E["n_lo"]), np.log(E["n_hi"])))))
        radius = 1200.0 * (n / 1e6) ** (1 / 3)
        N = next(g for g in E["grids"] if g * E["dx"] >= 2 * radius + 100.0)
        P = rt.make_params(N=N, dx=E["dx"], capacity=n + n // 4 + 64, seed=seed + r, substrates=SUBSTRATES[:1])
        cells = spheroid.make_cells(n, N, E["dx"], radius, seed + r, capacity=P.capacity)
        reps.append((P, cells, n))
    return reps


def ensemble_main(args) -> int:
    """Config B: the whole ensemble per biology step, graph mode, against the
    same replicates run one after another through the per-context calls."""
    from gell_amd import runtime as rt

    reps = ensemble_workload(args.seed)
    mech = rt.float_loop_count(6.0, 0.1)
    diff = rt.float_loop_count(0.1, 0.01)
    ens = rt.Ensemble([P for P, _, _ in reps])
    ens.autotune()
    knobs = {k: int(v) for k, v in (kv.split("=", 1) for kv in args.knob)}
    for m in ens.members:  # --knob overrides the autotuner's choice
        for k, v in knobs.items():
            m.tune(k, v)
    bufs = []
    for P, cells, _ in reps:
        b = rt.DeviceBuffer(P.capacity * rt.CELL_DTYPE.itemsize)
        b.upload(cells)
        bufs.append(b)
    n0 = [n for _, _, n in reps]

    def run(mode, steps, warmup):
        ens.set_mode(mode)
        for b, (_, cells, _) in zip(bufs, reps):
            b.upload(cells)
        n = list(n0)
        for w in range(warmup):
            n = ens.bio_step(bufs, n, w)
        ens.sync()
        agents = 0
        t0 = time.perf_counter()
        for k in range(steps):
            agents += sum(n) * mech
            n = ens.bio_step(bufs, n, warmup + k)
        ens.sync()
        return agents, time.perf_counter() - t0, n

    # The graph of a biology step is captured on the first step and once more
    # on the second (the uptake maps' x-line spans exist from then on, which
    # changes the clear kernels the step records); from the third step on it
    # is replayed. Two warm-up steps keep both captures out of the timed steps.
    warm = max(args.warmup, 2)
    agents, secs, n_end = run(rt.Ensemble.GRAPH, args.steps, warm)
    log(f"[bench B] graph: {secs:.3f} s for {args.steps} steps after {warm} warm-up steps")
    a_st, s_st, _ = run(rt.Ensemble.STREAMS, 1, 0)
    log(f"[bench B] streams: {s_st:.3f} s for 1 step")
    stats = ens.stats()
    ens.close()

    # the same replicates one after another, each through the per-context calls
    seq_agents, seq_secs = 0, 0.0
    for (P, cells, n), b in zip(reps, bufs):
        ctx = rt.Context(P)
        ctx.autotune()
        for k, v in knobs.items():
            ctx.tune(k, v)
        b.upload(cells)
        ctx.sync()
        t0 = time.perf_counter()
        seq_agents += n * mech
        n = ctx.birth(b, n, 0)
        for _ in range(mech):
            ctx.fm_update(b, n)
            ctx.prelod(0, b, n)
            for _ in range(diff):
                ctx.lod_substep(0)
        n = ctx.death(b, n)
        ctx.sync()
        seq_secs += time.perf_counter() - t0
        ctx.close()
    for b in bufs:
        b.free()
    log(f"[bench B] sequential: {seq_secs:.3f} s for 1 step")
    grids = {}
    for P, _, _ in reps:
        grids[P.N] = grids.get(P.N, 0) + 1
    line = {
        "metric": METRIC,
        "value": agents / secs,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded spheroids)",
        "config": {
            "workload": (f"B: {len(reps)} replicates, n log-uniform in [{ENSEMBLE['n_lo']}, {ENSEMBLE['n_hi']}], "
                         f"grids {grids} at 10 um, oxygen; step = one biology step of every replicate"),
            "cells_total": sum(n0),
            "voxels_total": sum(P.N ** 3 for P, _, _ in reps),
        },
        "roofline": None,
        "cpu_baseline": None,
        "streams_value": a_st / s_st,
        "sequential_value": seq_agents / seq_secs,
        "speedup_vs_sequential": (agents / secs) / (seq_agents / seq_secs),
        "graph_stats": stats,
        "n_final_total": sum(n_end),
    }
    print(json.dumps(line), flush=True)
    return 0


def drift_main(args) -> int:
    """Drift of the default (autotuned) GPU kernels over `--drift-steps`
    mechanics steps of the workload (FM + uptake maps + 11 LOD substeps per
    substrate, no PD) against the FP32 oracle (the parity reference) and the
    FP64 oracle (the drift of FP32 arithmetic itself). All three runs start
    from the same seeded state. The oracles run with OpenMP on this host's
    cores, in chunks of 10 steps, printing their progress."""
    from gell_amd import runtime as rt

    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle

    cfg = CONFIGS[args.config]
    P, cells = make_workload(cfg, args.seed)
    n, N, K = cfg["cells"], cfg["N"], args.drift_steps
    V = N ** 3
    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
    ctx = rt.Context(P)
    tuned = ctx.autotune()
    buf = rt.DeviceBuffer(P.capacity * rt.CELL_DTYPE.itemsize)
    buf.upload(cells)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(K):
        ctx.fm_update(buf, n)
        for s in range(P.n_substrates):
            ctx.prelod(s, buf, n)
        for s in range(P.n_substrates):
            for _ in range(diff):
                ctx.lod_substep(s)
    ctx.sync()
    t_gpu = time.perf_counter() - t0
    g_cells = buf.download(rt.CELL_DTYPE, P.capacity)[:n]
    g_fields = [ctx.field(s).astype(np.float64) for s in range(P.n_substrates)]
    buf.free()
    ctx.close()
    log(f"[drift] GPU: {K} steps in {t_gpu:.2f} s")

    orc = pyoracle.Oracle(omp=True)
    runs = {}
    for name, dt in (("fp32", np.float32), ("fp64", np.float64)):
        work = cells.copy()
        f = np.concatenate([np.full(V, P.sub[s].initial, dt) for s in range(P.n_substrates)])
        t0 = time.perf_counter()
        done = 0
        while done < K:
            k = min(10, K - done)
            orc.run_mech(P, work, n, f, k)
            done += k
            log(f"[drift] oracle {name}: {done}/{K} steps, {time.perf_counter() - t0:.1f} s")
        runs[name] = (work[:n], [f[s * V:(s + 1) * V].astype(np.float64) for s in range(P.n_substrates)],
                      time.perf_counter() - t0)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    c32, f32, t32 = runs["fp32"]
    c64, f64, t64 = runs["fp64"]
    live = (c32["sign"] != 0) & (c64["sign"] != 0) & (g_cells["sign"] != 0)
    pos = lambda c: c["pos"][live].astype(np.float64)  # noqa: E731
    line = {
        "metric": f"drift over {K} mechanics steps vs the FP32 and FP64 oracles (relative L2)",
        "config": {"workload": f"{args.config}: {N}^3 voxels, {n} cells, {P.n_substrates} substrate(s); "
                               f"step = FM + uptake maps + {diff} LOD substeps per substrate, no PD",
                   "kernel_variants": tuned},
        "steps": K,
        "field_gpu_vs_fp32": [rel(g, w) for g, w in zip(g_fields, f32)],
        "field_gpu_vs_fp64": [rel(g, w) for g, w in zip(g_fields, f64)],
        "field_fp32_vs_fp64": [rel(a, b) for a, b in zip(f32, f64)],
        "pos_gpu_vs_fp32": rel(pos(g_cells), pos(c32)),
        "pos_gpu_vs_fp64": rel(pos(g_cells), pos(c64)),
        "pos_fp32_vs_fp64": rel(pos(c32), pos(c64)),
        "live_cells": {"gpu": int((g_cells["sign"] != 0).sum()), "fp32": int((c32["sign"] != 0).sum()),
                       "fp64": int((c64["sign"] != 0).sum())},
        "seconds": {"gpu": t_gpu, "oracle_fp32": t32, "oracle_fp64": t64},
        "oracle_threads": orc.threads(),
        "host_nproc": os.cpu_count(),
    }
    print(json.dumps(line), flush=True)
    return 0


# What the default run reports of the surrogate leg.
SURROGATE_SUMMARY_KEYS = ("value", "unit", "predict_ms", "kept_call_ms", "lod50_ms", "bio_step_ms", "speedup_bio_step",
                          "field_rel_l2_vs_plain", "births", "deaths", "event_delta_frac", "surrogate_calls",
                          "surrogate_accepted", "tolerance")


def surrogate_main(args) -> int:
    print(json.dumps(surrogate_leg(args)), flush=True)
    return 0


def surrogate_leg(args) -> dict:
    """Config N: the 256^3 / 1M-cell spheroid (oxygen), one biology step with
    plain LOD substeps against the same step with the surrogate standing in
    for each run of K = 50 substeps (uptake map of the mechanics step it
    starts in; gell_surrogate_advance keeps a prediction only within its
    error budget, else runs the 50 substeps). Both runs start from the same
    state after the warm-up steps. The per-call speed-up is the 50 substeps'
    time over the time of an advance call whose prediction is kept (the
    first substep, the increment's norms, the network, the output)."""
    from gell_amd import runtime as rt

    cfg = dict(CONFIGS["L"], n_sub=1)
    P, cells = make_workload(cfg, args.seed)
    n0 = cfg["cells"]
    # a proliferating rim (clocks spread over their windows) and a necrotic
    # core, so that the compared biology step has births and deaths to count
    from gell_amd import spheroid

    c = cfg["N"] * cfg["dx"] / 2.0
    spheroid.radial_profile(cells, n0, (c, c, c), cfg["radius"], args.seed + 1)
    rng = np.random.default_rng(args.seed + 2)
    ctx = rt.Context(P)
    ctx.autotune()
    buf = rt.DeviceBuffer(P.capacity * rt.CELL_DTYPE.itemsize)
    buf.upload(cells)
    K = 50
    sur = rt.Surrogate(cfg["N"], K)
    sur.set_params(np.load(args.weights, allow_pickle=False))
    mech = rt.float_loop_count(P.dt_bio, P.dt_mech)
    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
    n = n0
    for w in range(args.warmup):
        n = ctx.birth(buf, n, w)
        for _ in range(mech):
            ctx.fm_update(buf, n)
            ctx.prelod(0, buf, n)
            for _ in range(diff):
                ctx.lod_substep(0)
        n = ctx.death(buf, n)
    ctx.sync()
    snap_cells = buf.download(rt.CELL_DTYPE, P.capacity)
    snap_field = ctx.field(0)
    snap_n = n
    # 1 % of the cells dead at the start of the compared step (removed by its death)
    snap_cells["sign"][:snap_n][rng.random(snap_n) < 0.01] = 0

    def bio_step(use_sur: bool, step: int):
        stats = dict(calls=0, used=0, residual=[], lod_substeps=0)
        if use_sur:
            sur.reset_budget()  # the error budget is per biology step
        m = ctx.birth(buf, snap_n, step)
        born = m - snap_n
        owed = 0
        for _ in range(mech):
            ctx.fm_update(buf, m)
            ctx.prelod(0, buf, m)
            if use_sur:
                owed += diff
                while owed >= K:
                    used, res = sur.advance(ctx, 0, args.sur_tol)
                    stats["calls"] += 1
                    stats["used"] += int(used)
                    stats["residual"].append(res)
                    stats["lod_substeps"] += 0 if used else K
                    owed -= K
            else:
                for _ in range(diff):
                    ctx.lod_substep(0)
                stats["lod_substeps"] += diff
        for _ in range(owed):
            ctx.lod_substep(0)
        stats["lod_substeps"] += owed
        m2 = ctx.death(buf, m)
        ctx.sync()
        stats["births"], stats["deaths"], stats["n_end"] = born, m - m2, m2
        return stats

    out = {}
    for mode in ("plain", "surrogate"):
        buf.upload(snap_cells)
        ctx.set_field(0, snap_field)
        ctx.sync()
        t0 = time.perf_counter()
        st = bio_step(mode == "surrogate", args.warmup)
        st["seconds"] = time.perf_counter() - t0
        st["field"] = ctx.field(0)
        out[mode] = st
        log(f"[bench N] {mode}: {st['seconds'] * 1e3:.1f} ms, calls {st['calls']}, used {st['used']}")
    sur_gate = sur.gate_stats()
    # the surrogate alone against the 50 substeps it replaces: the network
    # given the first substep's field, and a whole kept advance call (a
    # second surrogate with an unbounded budget, after its calibrating check)
    ctx.set_field(0, snap_field)
    pred = rt.DeviceBuffer(cfg["N"] ** 3 * 4)
    p0 = rt.DeviceBuffer(cfg["N"] ** 3 * 4)
    p0.upload(snap_field)
    ctx.lod_substep(0)
    rate_ptr = ctx.lib.gell_uptake_field(ctx.h, 0)
    for _ in range(2):
        sur.predict(p0, ctx.field_ptr(0), rate_ptr, pred)
    ctx.sync()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        sur.predict(p0, ctx.field_ptr(0), rate_ptr, pred)
    ctx.sync()
    t_pred = (time.perf_counter() - t0) / reps
    ctx.set_field(0, snap_field)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(K):
        ctx.lod_substep(0)
    ctx.sync()
    t_lod = time.perf_counter() - t0
    free_run = rt.Surrogate(cfg["N"], K)
    free_run.set_params(np.load(args.weights, allow_pickle=False))
    ctx.set_field(0, snap_field)
    free_run.advance(ctx, 0, 1e30)  # the calibrating check
    call_s, kept = [], 0
    for _ in range(reps):
        ctx.set_field(0, snap_field)  # a host upload, outside the timed call
        ctx.sync()
        t1 = time.perf_counter()
        used, _ = free_run.advance(ctx, 0, 1e30)
        ctx.sync()  # a kept call returns with the network's launches queued
        call_s.append(time.perf_counter() - t1)
        kept += int(used)
    t_call = float(np.median(call_s))
    free_run.close()
    p0.free()
    pl, su = out["plain"], out["surrogate"]
    res = su["residual"]
    line = {
        "metric": "surrogate vs LOD at 1e-3 L2 (K = 50 substeps, 256^3, 1M cells)",
        "value": t_lod / t_call,
        "unit": "x (50 LOD substeps / one advance call whose prediction is kept)",
        "higher_is_better": True,
        "n_gpus": 1,
        "dtype": "bf16 (network) / f32 (field)",
        "data": "synthetic (seeded spheroid)",
        "config": {"workload": "N: 256^3 at 10 um, 1000000 cells in a 1200 um ball, oxygen; surrogate every 50 "
                               "substeps", "tolerance": args.sur_tol, "weights": str(args.weights)},
        "predict_ms": t_pred * 1e3,
        "kept_call_ms": t_call * 1e3,
        "kept_calls_timed": f"{kept} of {reps}",
        "lod50_ms": t_lod * 1e3,
        "bio_step_ms": {"plain": pl["seconds"] * 1e3, "surrogate": su["seconds"] * 1e3},
        "surrogate_calls": su["calls"],
        "surrogate_accepted": su["used"],
        "first_increment_median": float(np.median(res)) if res else None,
        "first_increment_max": float(np.max(res)) if res else None,
        "field_rel_l2_vs_plain": float(np.linalg.norm(su["field"].astype(np.float64) - pl["field"])
                                       / np.linalg.norm(pl["field"].astype(np.float64))),
        "births": {"plain": pl["births"], "surrogate": su["births"]},
        "deaths": {"plain": pl["deaths"], "surrogate": su["deaths"]},
        "lod_substeps": {"plain": pl["lod_substeps"], "surrogate": su["lod_substeps"]},
        "gate": sur_gate,
        "speedup_bio_step": pl["seconds"] / su["seconds"],
        "event_delta_frac": {
            k: (abs(su[k] - pl[k]) / pl[k] if pl[k] else float(su[k] != pl[k])) for k in ("births", "deaths")},
        "tolerance": args.sur_tol,
    }
    pred.free()
    buf.free()
    sur.close()
    ctx.close()
    return line


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=2, help="timed biology steps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed biology steps")
    ap.add_argument("--config", choices=sorted(CONFIGS) + ["B", "N"], default="L")
    ap.add_argument("--weights", default=str(SURROGATE_WEIGHTS), help="surrogate parameters (config N)")
    ap.add_argument("--graph", action="store_true", help="time through the graph path (one-replicate ensemble)")
    ap.add_argument("--sur-tol", type=float, default=1e-3, help="surrogate acceptance tolerance (config N)")
    ap.add_argument("--unfused", dest="fused", action="store_false",
                    help="mechanics step and uptake maps as separate calls (gell_fm_update + gell_prelod)")
    ap.add_argument("--no-surrogate", dest="surrogate", action="store_false",
                    help="config L: skip the surrogate-vs-LOD leg")
    ap.add_argument("--drift-steps", type=int, default=0,
                    help="instead of timing: drift of the GPU run over this many mechanics steps against the FP32 "
                         "and FP64 oracles (config L or S)")
    ap.add_argument("--seed", type=int, default=20240611)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU-oracle sample budget (0 = skip)")
    ap.add_argument("--only", choices=["lod", "fm", "mech", "mech_unfused", "copy", "surrogate"], default=None,
                    help="run one kernel group alone (profiler runs); prints its timings")
    ap.add_argument("--knob", action="append", default=[], metavar="NAME=VALUE",
                    help="kernel-variant knob (e.g. lod_parts=8); in a full run it overrides the autotuner's choice")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the library has no multi-GPU path",
              file=sys.stderr)
        return 2
    if int(os.environ.get("RANK", "0")) != 0:
        return 0

    if args.config == "B":
        return ensemble_main(args)
    if args.config == "N":
        return surrogate_main(args)
    if args.drift_steps > 0:
        return drift_main(args)
    cfg = CONFIGS[args.config]
    log(f"[bench] building workload {args.config}")
    P, cells = make_workload(cfg, args.seed)
    knobs = {k: int(v) for k, v in (kv.split("=", 1) for kv in args.knob)}
    if args.only:
        print(json.dumps(only_leg(cfg, P, cells, args.only, args.steps, args.warmup, knobs)), flush=True)
        return 0
    g = gpu_leg(cfg, P, cells, args.steps, args.warmup, graph=cfg.get("graph", False) or args.graph, knobs=knobs,
                fused=args.fused)

    secs = g["seconds"]
    V = cfg["N"] ** 3
    lod_subs = args.steps * g["mech"] * g["diff"] * P.n_substrates
    lod_ms, lod_calls = g["prof"]["lod"]
    lod_us = lod_ms / max(lod_calls, 1) * 1e3
    tuned = g["tuned"]
    fused = bool(tuned.get("lod_fused")) and tuned["lod_parts"] > 1 and cfg["N"] in (128, 256, 512)
    moved = LOD_BYTES_PER_VOXEL_FUSED if fused else LOD_BYTES_PER_VOXEL
    achieved = LOD_BYTES_PER_VOXEL * V / (lod_us * 1e-6) / 1e9
    traffic = read_traffic(f"lod_substep_{cfg['N']}_traffic.txt")
    phases = {k: round(v[0] / max(v[1], 1) * 1e3, 2) for k, v in g["prof"].items() if v[1]}
    force_ms, force_calls = g["prof_alone"]["fm_force_move"]
    force_us = force_ms / max(force_calls, 1) * 1e3
    force_achieved = FORCE_BYTES_PER_CELL * g["n_alone"] / (force_us * 1e-6) / 1e9
    force_traffic = read_traffic(f"fm_{cfg['N']}_traffic.txt")
    fv = tuned.get("fm_prefetch")
    if fv == 10:  # auto: the rank-compacted list on large sparse tissue
        n_a = g["n_alone"]
        fv = f"10 = auto, here {9 if n_a >= 524288 and 8 * n_a <= V else 7}"

    cpu = None
    if args.cpu_seconds > 0:
        log("[bench] CPU oracle leg")
        c = cpu_leg(cfg, P, cells, args.cpu_seconds)
        cpu = {
            "value": c["value"],
            "unit": "agent-steps/s",
            "cores": c["threads"],
            "kind": "port",
            "sample": (f"{c['steps']} mechanics steps (FM + uptake maps + {g['diff']} LOD substeps x "
                       f"{P.n_substrates} substrates, no PD) of the same workload, OpenMP oracle, "
                       f"{c['seconds']:.1f} s"),
            "voxel_updates_per_s": c["voxel_updates_per_s"],
            "host_nproc": os.cpu_count(),
        }

    line = {
        "metric": METRIC,
        "value": g["agents"] / secs,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded spheroid)",
        "config": {
            "workload": (f"{args.config}: {cfg['N']}^3 voxels at {cfg['dx']:g} um, {cfg['cells']} cells in a "
                         f"{cfg['radius']:g} um ball, {P.n_substrates} substrate(s); step = birth + "
                         f"{g['mech']} x (FM + uptake maps + {g['diff']} LOD substeps per substrate) + death"),
            "grid": cfg["N"],
            "cells": cfg["cells"],
            "path": ("graph (one-replicate ensemble)" if (cfg.get("graph", False) or args.graph) else
                     "per-step calls, " + ("gell_mech_step (uptake maps accumulated by the force kernel)" if
                                           args.fused else "gell_fm_update + gell_prelod")),
            "substrates": P.n_substrates,
            "mech_steps_per_step": g["mech"],
            "lod_substeps_per_mech_step": g["diff"],
        },
        "voxel_updates_per_s": V * lod_subs / secs,
        "roofline": {
            "bound": "hbm",
            "kernel": ("LOD substep, fused: x sweep + y blocks (with boundary + uptake), y reduce, "
                       "z sweep with y correction" if fused else
                       "LOD substep (x sweep with boundary + uptake, y sweep, z sweep)"),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": LOD_BYTES_PER_VOXEL * V,
            "bytes_per_voxel": LOD_BYTES_PER_VOXEL,
            "bytes_per_voxel_variant_moves": moved,
            "us_per_launch": lod_us,
        },
        # the second target kernel: force + movement, 144 B per agent-step
        "roofline_force": {
            "bound": "hbm",
            "kernel": f"force + movement (fm_force_move, variant {fv})",
            "achieved": force_achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": force_achieved / HBM_PEAK_GBS,
            "traffic": force_traffic,
            "bytes_per_launch": FORCE_BYTES_PER_CELL * g["n_alone"],
            "us_per_launch": force_us,
            "timed_in": "mechanics steps with separate uptake maps (the kernel alone)",
            "fused_us_per_launch": (phases.get("fm_force_move") if g["fused"] else None),
        },
        "cpu_baseline": cpu,
        "phase_us": phases,
        "kernel_variants": g["tuned"],
        "device": g["device"],
    }
    if args.config == "L" and args.surrogate:
        # the metric's third clause: the surrogate against plain LOD on the
        # same workload (oxygen only), one biology step each
        log("[bench] surrogate leg")
        sl = surrogate_leg(args)
        line["surrogate"] = {k: sl[k] for k in SURROGATE_SUMMARY_KEYS if k in sl}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
