"""Training and evaluation of the LOD surrogate (gell_surrogate_* of the C
ABI) on fields produced by the library's own LOD kernels.

Samples come from simulated spheroids: cells are placed, mechanics steps
(cell list, forces, uptake map, 11 LOD substeps each) advance the oxygen
field from its initial value, and at chosen steps the pair (field, uptake
map) is stored with the field K substeps later under that map (the
trajectory then continues from the stored field). Training minimises the
squared interior error with Adam; evaluation reports the relative L2 error
of the predicted field and of the predicted update on held-out samples.

    python -m gell_amd.surrogate_train --N 128 --trajectories 24 --steps 3000 \
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
        self.rate = rt.DeviceBuffer(nb)
        self.target = rt.DeviceBuffer(nb)

    def free(self):
        for b in (self.p0, self.rate, self.target):
            b.free()


def trajectory_samples(N: int, n_cells: int, seed: int, K: int, sample_steps: list[int], dx: float = 10.0):
    """Samples along one spheroid trajectory (mechanics steps from a uniform
    field; a sample at each step in sample_steps)."""
    lib = rt.load_library()
    P = rt.make_params(N=N, dx=dx, capacity=n_cells + 64, seed=seed, substrates=[OXYGEN])
    radius = min(1200.0 * (n_cells / 1e6) ** (1 / 3), 0.42 * N * dx)
    cells = spheroid.make_cells(n_cells, N, dx, radius, seed, capacity=P.capacity)
    ctx = rt.Context(P)
    buf = rt.DeviceBuffer(P.capacity * rt.CELL_DTYPE.itemsize)
    buf.upload(cells)
    diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
    out = []
    nb = N ** 3 * 4
    field = ctx.field_ptr(0)
    rate = lib.gell_uptake_field(ctx.h, 0)
    for m in range(max(sample_steps) + 1):
        ctx.fm_update(buf, n_cells)
        ctx.prelod(0, buf, n_cells)
        if m in sample_steps:
            s = Sample(N)
            ctx.sync()
            lib.gell_copy_d2d(s.p0.ptr, field, nb)
            lib.gell_copy_d2d(s.rate.ptr, rate, nb)
            for _ in range(K):
                ctx.lod_substep(0)
            ctx.sync()
            lib.gell_copy_d2d(s.target.ptr, field, nb)
            lib.gell_copy_d2d(field, s.p0.ptr, nb)
            out.append(s)
        for _ in range(diff):
            ctx.lod_substep(0)
    ctx.sync()
    buf.free()
    ctx.close()
    return out


def make_dataset(N: int, trajectories: int, K: int, seed: int, per_traj: int = 6, max_step: int = 240,
                 early: int = 20):
    """Half of each trajectory's samples come from its first `early`
    mechanics steps, where the field still relaxes from its initial value and
    the K-substep update is large; the rest from the quasi-steady remainder."""
    rng = np.random.default_rng(seed)
    data = []
    for t in range(trajectories):
        n_cells = int(np.exp(rng.uniform(np.log(2e4), np.log(1.6e5))) * (N / 128) ** 3)
        h = per_traj // 2
        steps = sorted(set(int(x) for x in np.concatenate([rng.integers(0, early, h),
                                                             rng.integers(early, max_step, per_traj - h)])))
        data += trajectory_samples(N, n_cells, int(seed + 1000 * t), K, steps)
        log(f"[train] trajectory {t}: {n_cells} cells, samples at mechanics steps {steps}")
    return data


def sq_errors(sur: rt.Surrogate, s: Sample, N: int, out: rt.DeviceBuffer) -> tuple[float, float, float]:
    """(|pred - target|^2, |target|^2, |target - p0|^2) of one sample."""
    sur.predict(s.p0, s.rate, out)
    pred = out.download(np.float32, N ** 3).astype(np.float64)
    p0 = s.p0.download(np.float32, N ** 3).astype(np.float64)
    tgt = s.target.download(np.float32, N ** 3).astype(np.float64)
    return float(((pred - tgt) ** 2).sum()), float((tgt ** 2).sum()), float(((tgt - p0) ** 2).sum())


def summarize(errs) -> dict:
    e = np.array(errs)
    field = np.sqrt(e[:, 0] / e[:, 1])
    return {"field_rel_l2_mean": float(field.mean()), "field_rel_l2_max": float(field.max()),
            "update_rel_l2": float(np.sqrt(e[:, 0].sum() / max(e[:, 2].sum(), 1e-30))),
            "zero_update_field_rel_l2_mean": float(np.sqrt(e[:, 2] / e[:, 1]).mean())}


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
            loss += sur.gradients(s.p0, s.rate, s.target, accumulate=b > 0) / batch
        sur.adam(float(lr_t))
        hist.append(loss)
        if (it + 1) % report_every == 0:
            log(f"[train] step {it + 1}: mean loss {np.mean(hist[-report_every:]):.5f} "
                f"({(time.perf_counter() - t0) / (it + 1) * 1e3:.2f} ms/step)")
    return hist, time.perf_counter() - t0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--K", type=int, default=50)
    ap.add_argument("--trajectories", type=int, default=24)
    ap.add_argument("--per-traj", type=int, default=6)
    ap.add_argument("--steps", type=int, default=3000)
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
    after = summarize([sq_errors(sur, s, args.N, out) for s in held])
    fit = summarize([sq_errors(sur, s, args.N, out) for s in train_set[:24]])
    np.save(args.out, sur.params())
    rep = {
        "N": args.N, "K": args.K, "train_samples": len(train_set), "held_out": len(held), "steps": args.steps,
        "train_seconds": secs, "final_loss": float(np.mean(hist[-100:])),
        "held_before_training": before, "held": after, "train_subset": fit,
    }
    print(json.dumps(rep), flush=True)
    out.free()
    for s in data:
        s.free()
    sur.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
