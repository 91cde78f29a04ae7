"""Train the learned pressure guess on solver-generated pairs with the library's own
forward, backward and Adam kernels (fs_trainer_*, fluidsim.trainer.GuessTrainer).

Data: band-forced 2D turbulence at 256^2 (kf = 4 +- 1, the 2048^2 forcing
wavenumber 32 scaled to the grid, 24 modes, amplitude 0.1, drag 0.02, nu = 4e-4,
seed 1000 * run + step), one run per seed from the seeded random-phase initial
field, dt at CFL 0.5, every projection converged to an FP64 residual of 1e-9.
The pressure after each step is copied, device to device, into the trainer's
snapshot store (FP64); consecutive pressures (p_prev, p) form a pair. Since L p = div u* to the gate, the network
input is r = L p - L p_prev and the target is (p - p_prev) / (max|r| h^2).

    python -m fluidsim.train_guess --out fluid-sim_amd/weights/guess_unet_256.fsgn
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np


def stencil(e):
    """h^2 times the 5-point periodic Laplacian of e [.., n, n]."""
    return (np.roll(e, 1, -1) + np.roll(e, -1, -1) + np.roll(e, 1, -2) + np.roll(e, -1, -2)) - 4.0 * e


def generate(tr, n, runs, steps, warmup, log):
    """Pressure snapshots of every run into the trainer's store: run q, step k at slot
    q (steps + 1) + k."""
    import fluidsim
    from fluidsim import initial
    tr.reserve(runs * (steps + 1))
    for run in range(runs):
        u0, v0 = initial.decaying_turbulence(n, seed=run + 1)
        with fluidsim.GPUFluidSolver(0, n=n, nu=4e-4, advection="maccormack", pressureSolver="multigrid",
                                     pressureIterations=60, tolerance=1e-9, mgMinSize=4, guess="previous") as g:
            g.setForcing(dict(kf=4.0, width=1.0, amp=0.1, drag=0.02, seed=1000 * (run + 1)))
            g.uploadVelocity(u0, v0)
            for _ in range(warmup):
                g.stepCFL(0.5)
            for k in range(steps + 1):
                g.stepCFL(0.5)
                tr.store_pressure(run * (steps + 1) + k, g)
        if run % 4 == 0:
            log(f"data run {run + 1}/{runs}")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--runs", type=int, default=32)
    ap.add_argument("--steps", type=int, default=625)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--iters", type=int, default=12000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    from fluidsim.guess_net import init_blob
    from fluidsim.trainer import GuessTrainer

    def log(msg):
        print(f"[train {time.strftime('%H:%M:%S')}] {msg}", flush=True)

    import fluidsim
    fluidsim.createDevice()
    h = 2 * math.pi / a.n
    tr = GuessTrainer(0, n=a.n, batch=a.batch, h=h)
    tr.load(init_blob(a.seed))
    t0 = time.time()
    generate(tr, a.n, a.runs, a.steps, a.warmup, log)
    log(f"{a.runs * a.steps} pairs at {a.n}^2 in {time.time() - t0:.1f} s")
    S = a.steps + 1
    # pair k of run q starts at snapshot q S + k; the last two runs are held out
    train = np.array([q * S + k for q in range(a.runs - 2) for k in range(a.steps)], np.int64)
    val = np.array([q * S + k for q in range(a.runs - 2, a.runs) for k in range(a.steps)], np.int64)
    rng = np.random.default_rng(a.seed)
    t1 = time.time()
    loss = float("nan")
    for it in range(a.iters):
        idx = train[rng.integers(0, len(train), a.batch)]
        tr.batch_step(idx)
        # cosine annealing from lr to 0 over the run
        tr.adam(a.lr * 0.5 * (1 + math.cos(math.pi * it / a.iters)))
        if it % 500 == 0 or it == a.iters - 1:
            loss = tr.loss()
            log(f"iter {it} loss {loss:.4e} ({(time.time() - t1) / (it + 1) * 1e3:.2f} ms/iter)")
    train_s = time.time() - t1
    # held-out: how far the guess reduces the residual the previous pressure leaves,
    # max|L(p - p_guess)| / max|r| = max|h^2 L (o - t)|, through the library's forward pass
    ratios = []
    for c in range(0, len(val) - a.batch + 1, a.batch):
        tr.batch_step(val[c:c + a.batch])
        e = tr.error().astype(np.float64)
        ratios += list(np.max(np.abs(stencil(e)), axis=(-2, -1)))
    red = np.array(ratios)
    stats = {"pairs": a.runs * a.steps, "n": a.n, "iters": a.iters, "batch": a.batch, "lr": a.lr,
             "final_loss": loss, "train_seconds": round(train_s, 1),
             "trainer": "libfluidsim fs_trainer: own forward, backward and Adam kernels",
             "heldout_residual_ratio_median": float(np.median(red)),
             "heldout_residual_ratio_mean": float(np.mean(red))}
    log(json.dumps(stats))
    blob = tr.export()
    tr.close()
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "wb") as f:
        f.write(blob)
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(stats, f, indent=1)
    log(f"wrote {a.out} ({len(blob)} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
