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

With --stream F the training set is streamed instead: a pool of samples in
device memory with F of them replaced by fresh fields before every Adam
step, so a run of S steps trains on about S x F distinct fields:

    python -m gell_amd.surrogate_train --N 256 --stream 4 --steps 2500 \\
        --init weights.npy --out weights2.npy
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


def spheroid_size(N: int, rng) -> tuple[int, float]:
    """Cells (100k..1M at 256^3, scaled by (N/256)^3, log-uniform) and a
    radius giving 0.1..0.3 cells per 10 um voxel."""
    n_cells = int(np.exp(rng.uniform(np.log(1e5), np.log(1e6))) * (N / 256) ** 3)
    density = rng.uniform(0.1, 0.3)
    radius = min((3 * n_cells / (4 * np.pi * density)) ** (1 / 3) * 10.0, 0.42 * N * 10.0)
    return n_cells, radius


class Trajectory:
    """One spheroid (radial phase profile) run through biology steps of the
    library: birth, 61 mechanics steps (cell list, forces, uptake map, 11 LOD
    substeps), death. `map_step` runs the next mechanics step up to its uptake
    map, `sample_into` stores that state with the field K substeps later (the
    run then goes on from the stored field), `finish_step` runs the step's
    substeps."""

    def __init__(self, N: int, n_cells: int, radius: float, seed: int, K: int, dx: float = 10.0):
        self.lib = rt.load_library()
        P = rt.make_params(N=N, dx=dx, capacity=2 * n_cells + 64, seed=seed, substrates=[OXYGEN])
        cells = spheroid.make_cells(n_cells, N, dx, radius, seed, capacity=P.capacity)
        c = N * dx / 2.0
        spheroid.radial_profile(cells, n_cells, (c, c, c), radius, seed + 1)
        self.ctx = rt.Context(P)
        self.buf = rt.DeviceBuffer(P.capacity * rt.CELL_DTYPE.itemsize)
        self.buf.upload(cells)
        self.mech = rt.float_loop_count(P.dt_bio, P.dt_mech)
        self.diff = rt.float_loop_count(P.dt_mech, P.dt_diff)
        self.N, self.K, self.n = N, K, n_cells
        self.bio, self.step = 0, 0  # the next mechanics step

    def map_step(self) -> None:
        if self.step == 0:
            self.n = self.ctx.birth(self.buf, self.n, self.bio + 1)
        self.ctx.fm_update(self.buf, self.n)
        self.ctx.prelod(0, self.buf, self.n)

    def sample_into(self, s: "Sample") -> None:
        nb = self.N ** 3 * 4
        field = self.ctx.field_ptr(0)
        s.tag = f"bio {self.bio} mech {self.step}"
        self.ctx.sync()
        self.lib.gell_copy_d2d(s.p0.ptr, field, nb)
        self.lib.gell_copy_d2d(s.rate.ptr, self.lib.gell_uptake_field(self.ctx.h, 0), nb)
        self.ctx.lod_substep(0)
        self.ctx.sync()
        self.lib.gell_copy_d2d(s.p1.ptr, field, nb)
        for _ in range(self.K - 1):
            self.ctx.lod_substep(0)
        self.ctx.sync()
        self.lib.gell_copy_d2d(s.target.ptr, field, nb)
        self.lib.gell_copy_d2d(field, s.p0.ptr, nb)

    def finish_step(self) -> None:
        for _ in range(self.diff):
            self.ctx.lod_substep(0)
        self.step += 1
        if self.step == self.mech:
            self.n = self.ctx.death(self.buf, self.n)
            self.step = 0
            self.bio += 1

    def close(self) -> None:
        self.ctx.sync()
        self.buf.free()
        self.ctx.close()


def trajectory_samples(N: int, n_cells: int, radius: float, seed: int, K: int, picks: set[tuple[int, int]],
                       dx: float = 10.0) -> list[Sample]:
    """Samples at (biology step, mechanics step) pairs of one spheroid's run."""
    t = Trajectory(N, n_cells, radius, seed, K, dx)
    out = []
    last_bio = max(b for b, _ in picks)
    while t.bio <= last_bio:
        t.map_step()
        if (t.bio, t.step) in picks:
            s = Sample(N)
            t.sample_into(s)
            out.append(s)
        t.finish_step()
    t.close()
    return out


def make_dataset(N: int, trajectories: int, K: int, seed: int, per_traj: int = 6, bio_steps: int = 3):
    """Per spheroid (spheroid_size) a third of the samples in the first
    biology step's first 20 mechanics steps, the rest in later biology steps,
    one of them at the mechanics step right after births."""
    rng = np.random.default_rng(seed)
    data = []
    for t in range(trajectories):
        n_cells, radius = spheroid_size(N, rng)
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


class SampleStream:
    """Fresh samples on demand for streamed training: spheroids (spheroid_size)
    run for `bio_steps` biology steps each, one sample every 1..3 mechanics
    steps (so a third come from the first biology step, where the field still
    relaxes, and the rest from the quasi-steady steps after it, births and
    deaths included)."""

    def __init__(self, N: int, K: int, seed: int, bio_steps: int = 3):
        self.N, self.K, self.bio_steps = N, K, bio_steps
        self.rng = np.random.default_rng(seed)
        self.seed = seed
        self.traj = None
        self.spheroids = 0
        self.fields = 0

    def _trajectory(self) -> Trajectory:
        if self.traj is None or self.traj.bio >= self.bio_steps:
            if self.traj is not None:
                self.traj.close()
            n_cells, radius = spheroid_size(self.N, self.rng)
            self.traj = Trajectory(self.N, n_cells, radius, self.seed + 7919 * self.spheroids, self.K)
            self.spheroids += 1
        return self.traj

    def next_into(self, s: Sample) -> None:
        for _ in range(int(self.rng.integers(0, 3))):  # skip 0..2 mechanics steps
            t = self._trajectory()
            t.map_step()
            t.finish_step()
        t = self._trajectory()
        t.map_step()
        t.sample_into(s)
        t.finish_step()
        self.fields += 1

    def close(self) -> None:
        if self.traj is not None:
            self.traj.close()


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


def train_streamed(sur: rt.Surrogate, stream: SampleStream, pool: list[Sample], steps: int, fresh: int, lr: float,
                   seed: int, batch: int = 4, report_every: int = 250):
    """Adam steps on mini-batches drawn from a pool of samples in device
    memory, `fresh` pool entries replaced by new fields of the stream before
    each step: about steps x fresh distinct fields, each seen batch / fresh
    times on average."""
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    for it in range(steps):
        lr_t = lr * min(1.0, (it + 1) / 200) * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * it / max(steps, 1))))
        for _ in range(fresh):
            stream.next_into(pool[int(rng.integers(len(pool)))])
        loss = 0.0
        for b in range(batch):
            s = pool[int(rng.integers(len(pool)))]
            loss += sur.gradients(s.p0, s.p1, s.rate, s.target, accumulate=b > 0) / batch
        sur.adam(float(lr_t))
        hist.append(loss)
        if (it + 1) % report_every == 0:
            log(f"[train] step {it + 1}: mean loss {np.mean(hist[-report_every:]):.5f}, {stream.fields} fields from "
                f"{stream.spheroids} spheroids ({(time.perf_counter() - t0) / (it + 1) * 1e3:.2f} ms/step)")
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
    ap.add_argument("--stream", type=int, default=0, metavar="FRESH",
                    help="streamed training: FRESH new fields per step into a pool (0: a fixed dataset)")
    ap.add_argument("--pool", type=int, default=128, help="pool size of streamed training (samples)")
    ap.add_argument("--init", default=None, help="start from these parameters (.npy)")
    ap.add_argument("--out", required=True, help="where the trained parameters go (.npy)")
    args = ap.parse_args()

    sur = rt.Surrogate(args.N, args.K)
    if args.init:
        sur.set_params(np.load(args.init, allow_pickle=False))
    t0 = time.perf_counter()
    stream = None
    if args.stream:
        # held-out spheroids from their own seeds; the pool filled from the stream
        held = make_dataset(args.N, max(2, args.trajectories // 8), args.K, args.seed + 10 ** 6, args.per_traj)
        stream = SampleStream(args.N, args.K, args.seed)
        train_set = [Sample(args.N) for _ in range(args.pool)]
        for s in train_set:
            stream.next_into(s)
        data = train_set + held
    else:
        data = make_dataset(args.N, args.trajectories, args.K, args.seed, args.per_traj)
        held = data[-max(2, len(data) // 8):]
        train_set = data[:len(data) - len(held)]
    log(f"[train] {len(train_set)} training / {len(held)} held-out samples in {time.perf_counter() - t0:.1f} s")
    out = rt.DeviceBuffer(args.N ** 3 * 4)
    before = summarize([sq_errors(sur, s, args.N, out) for s in held])
    if stream:
        hist, secs = train_streamed(sur, stream, train_set, args.steps, args.stream, args.lr, args.seed, args.batch)
    else:
        hist, secs = train(sur, train_set, args.steps, args.lr, args.seed, args.batch)
    after_each = [sq_errors(sur, s, args.N, out) for s in held]
    after = summarize(after_each)
    fit = summarize([sq_errors(sur, s, args.N, out) for s in train_set[:24]])
    np.save(args.out, sur.params())
    rep = {
        "N": args.N, "K": args.K, "train_samples": stream.fields if stream else len(train_set),
        "spheroids": stream.spheroids if stream else args.trajectories, "held_out": len(held), "steps": args.steps,
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
    if stream:
        stream.close()
    sur.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
