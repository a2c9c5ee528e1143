"""Training and evaluation of the LOD surrogate (gell_surrogate_* of the C
ABI) on fields produced by the library's own kernels.

A sample is what gell_surrogate_advance sees in the loop: the field p0 right
after a mechanics step's uptake map, the field p1 one real substep later,
the uptake map, and the target, the field K substeps after p0 under that map.
Samples come from simulated spheroids (seeded cells with a radial phase
profile: a proliferating rim, a quiescent shell, a necrotic core) run through
biology steps of the library (birth, 61 mechanics steps of cell list,
forces, uptake map and 11 LOD substeps, death): a few from the first
biology step, where the oxygen field still relaxes from its initial value,
the rest from later, quasi-steady biology steps, among them the mechanics
step right after births. The trajectory goes on from the stored p0.

The network predicts the K-substep increment in units of the first
substep's increment, so every sample weighs the same in the loss whatever
its size. Evaluation reports, on held-out samples, the relative L2 error of
the predicted increment and of the predicted field, beside those of keeping
the field (p0) and of one substep (p1).

    python -m gell_amd.surrogate_train --N 256 --trajectories 32 --steps 4000 \\
        --out weights.npy
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import runtime as rt
from . import spheroid

OXYGEN = dict(boundary=38.0, D=1.0e5, decay=0.1, uptake=-10.0, dead_factor=0.1, initial=38.0)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


class Sample:
    def __init__(self, N: int):
        nb = N ** 3 * 4
        self.p0 = rt.DeviceBuffer(nb)
        self.p1 = rt.DeviceBuffer(nb)
        self.rate = rt.DeviceBuffer(nb)
        self.target = rt.DeviceBuffer(nb)
        self.tag = ""

    def free(self):
        for b in (self.p0, self.p1, self.rate, self.target):
            b.free()


def trajectory_samples(N: int, n_cells: int, radius: float, seed: int, K: int, picks: set[tuple[int, int]],
                       dx: float = 10.0) -> list[Sample]:
    """Samples at (biology step, mechanics step) pairs of one spheroid's run."""
    lib = rt.load_library()
    P = rt.make_params(N=N, dx=dx, capacity=2 * n_cells + 64, seed=seed, substrates=[OXYGEN])
    cells = spheroid.make_cells(n_cells, N, dx, radius, seed, capacity=P.capacity)
    c = N * dx / 2.0
    spheroid.radial_profile(cells, n_cells, (c, c, c), radius, seed + 1)
    ctx = rt.Context(P)
    buf = rt.DeviceBuffer(P.capacity * rt.CELL_DTYPE.itemsize)
    buf.upload(cells)
    mech = rt.float_loop_count(P.dt_bio, P.dt_mech)
    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
    nb = N ** 3 * 4
    field = ctx.field_ptr(0)
    rate = lib.gell_uptake_field(ctx.h, 0)
    out = []
    n = n_cells
    last_bio = max(b for b, _ in picks)
    for b in range(last_bio + 1):
        n = ctx.birth(buf, n, b + 1)
        for m in range(mech):
            ctx.fm_update(buf, n)
            ctx.prelod(0, buf, n)
            if (b, m) in picks:
                s = Sample(N)
                s.tag = f"bio {b} mech {m}"
                ctx.sync()
                lib.gell_copy_d2d(s.p0.ptr, field, nb)
                lib.gell_copy_d2d(s.rate.ptr, rate, nb)
                ctx.lod_substep(0)
                ctx.sync()
                lib.gell_copy_d2d(s.p1.ptr, field, nb)
                for _ in range(K - 1):
                    ctx.lod_substep(0)
                ctx.sync()
                lib.gell_copy_d2d(s.target.ptr, field, nb)
                lib.gell_copy_d2d(field, s.p0.ptr, nb)
                out.append(s)
            for _ in range(diff):
                ctx.lod_substep(0)
        n = ctx.death(buf, n)
    ctx.sync()
    buf.free()
    ctx.close()
    return out


def make_dataset(N: int, trajectories: int, K: int, seed: int, per_traj: int = 6, bio_steps: int = 3):
    """Spheroids of 100k..1M cells at 256^3 (scaled by (N/256)^3), radius
    giving 0.1..0.3 cells per voxel; per trajectory a third of the samples in
    the first biology step's first 20 mechanics steps, the rest in later
    biology steps, one of them at the mechanics step right after births."""
    rng = np.random.default_rng(seed)
    data = []
    for t in range(trajectories):
        n_cells = int(np.exp(rng.uniform(np.log(1e5), np.log(1e6))) * (N / 256) ** 3)
        density = rng.uniform(0.1, 0.3)  # cells per voxel (10 um voxels)
        radius = min((3 * n_cells / (4 * np.pi * density)) ** (1 / 3) * 10.0, 0.42 * N * 10.0)
        early = per_traj // 3
        picks = {(0, int(m)) for m in rng.choice(20, early, replace=False)}
        picks.add((int(rng.integers(1, bio_steps)), 0))
        while len(picks) < per_traj:
            picks.add((int(rng.integers(1, bio_steps)), int(rng.integers(0, 61))))
        t0 = time.perf_counter()
        data += trajectory_samples(N, n_cells, radius, int(seed + 1000 * t), K, picks)
        log(f"[train] trajectory {t}: {n_cells} cells, radius {radius:.0f} um, samples at {sorted(picks)} "
            f"({time.perf_counter() - t0:.1f} s)")
    return data


def sq_errors(sur: rt.Surrogate, s: Sample, N: int, out: rt.DeviceBuffer) -> dict:
    """Squared norms of one sample: prediction error, target, and the errors
    of keeping p0 and of taking p1."""
    sur.predict(s.p0, s.p1, s.rate, out)
    V = N ** 3
    pred = out.download(np.float32, V).astype(np.float64)
    p0 = s.p0.download(np.float32, V).astype(np.float64)
    p1 = s.p1.download(np.float32, V).astype(np.float64)
    tgt = s.target.download(np.float32, V).astype(np.float64)
    return {"pred": float(((pred - tgt) ** 2).sum()), "target": float((tgt ** 2).sum()),
            "keep": float(((tgt - p0) ** 2).sum()), "one": float(((tgt - p1) ** 2).sum())}


def summarize(errs: list[dict]) -> dict:
    e = {k: np.array([d[k] for d in errs]) for k in errs[0]}
    field = np.sqrt(e["pred"] / e["target"])
    incr = np.sqrt(e["pred"] / np.maximum(e["keep"], 1e-300))
    return {"field_rel_l2_mean": float(field.mean()), "field_rel_l2_max": float(field.max()),
            "increment_rel_l2_mean": float(incr.mean()), "increment_rel_l2_max": float(incr.max()),
            "keep_field_rel_l2_mean": float(np.sqrt(e["keep"] / e["target"]).mean()),
            "one_substep_field_rel_l2_mean": float(np.sqrt(e["one"] / e["target"]).mean()),
            "one_substep_increment_rel_l2_mean": float(np.sqrt(e["one"] / np.maximum(e["keep"], 1e-300)).mean())}


def train(sur: rt.Surrogate, data: list[Sample], steps: int, lr: float, seed: int, batch: int = 4,
          report_every: int = 250):
    """Adam steps on mini-batches of `batch` samples (gradients summed)."""
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    for it in range(steps):
        # linear warm-up over 200 steps, then cosine decay to lr / 20
        lr_t = lr * min(1.0, (it + 1) / 200) * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * it / max(steps, 1))))
        loss = 0.0
        for b in range(batch):
            s = data[int(rng.integers(len(data)))]
            loss += sur.gradients(s.p0, s.p1, s.rate, s.target, accumulate=b > 0) / batch
        sur.adam(float(lr_t))
        hist.append(loss)
        if (it + 1) % report_every == 0:
            log(f"[train] step {it + 1}: mean loss {np.mean(hist[-report_every:]):.5f} "
                f"({(time.perf_counter() - t0) / (it + 1) * 1e3:.2f} ms/step)")
    return hist, time.perf_counter() - t0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--K", type=int, default=50)
    ap.add_argument("--trajectories", type=int, default=32)
    ap.add_argument("--per-traj", type=int, default=6)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init", default=None, help="start from these parameters (.npy)")
    ap.add_argument("--out", required=True, help="where the trained parameters go (.npy)")
    args = ap.parse_args()

    sur = rt.Surrogate(args.N, args.K)
    if args.init:
        sur.set_params(np.load(args.init, allow_pickle=False))
    t0 = time.perf_counter()
    data = make_dataset(args.N, args.trajectories, args.K, args.seed, args.per_traj)
    held = data[-max(2, len(data) // 8):]
    train_set = data[:len(data) - len(held)]
    log(f"[train] {len(train_set)} training / {len(held)} held-out samples in {time.perf_counter() - t0:.1f} s")
    out = rt.DeviceBuffer(args.N ** 3 * 4)
    before = summarize([sq_errors(sur, s, args.N, out) for s in held])
    hist, secs = train(sur, train_set, args.steps, args.lr, args.seed, args.batch)
    after_each = [sq_errors(sur, s, args.N, out) for s in held]
    after = summarize(after_each)
    fit = summarize([sq_errors(sur, s, args.N, out) for s in train_set[:24]])
    np.save(args.out, sur.params())
    rep = {
        "N": args.N, "K": args.K, "train_samples": len(train_set), "held_out": len(held), "steps": args.steps,
        "train_seconds": secs, "final_loss": float(np.mean(hist[-100:])),
        "held_before_training": before, "held": after, "train_subset": fit,
        "held_per_sample": [{"tag": s.tag, "increment_rel_l2": float(np.sqrt(d["pred"] / max(d["keep"], 1e-300))),
                             "keep_field_rel_l2": float(np.sqrt(d["keep"] / d["target"]))}
                            for s, d in zip(held, after_each)],
    }
    print(json.dumps(rep), flush=True)
    out.free()
    for s in data:
        s.free()
    sur.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
