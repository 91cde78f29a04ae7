"""Train the learned pressure guess on solver-generated pairs (GPU, torch autograd).

Data: band-forced 2D turbulence at 256^2 (kf = 4 +- 1, the 2048^2 forcing
wavenumber 32 scaled to the grid, 24 modes, amplitude 0.1, drag 0.02, nu = 4e-4,
seed 1000 * run + step), one run per seed from the seeded random-phase initial
field, dt at CFL 0.5, every projection converged to an FP64 residual of 1e-9.
The pressure after each step is recorded; consecutive pressures (p_prev, p) form
a pair. Since L p = div u* to the gate, the network input is r = L p - L p_prev
and the target is (p - p_prev) / (max|r| h^2).

    python -m fluidsim.train_guess --out fluid-sim_amd/weights/guess_unet.fsgn
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np


def laplacian_t(p, h):
    import torch
    return (torch.roll(p, 1, -1) + torch.roll(p, -1, -1) + torch.roll(p, 1, -2) + torch.roll(p, -1, -2) - 4 * p) / (h * h)


def generate(n, runs, steps, warmup, log):
    import torch
    import fluidsim
    from fluidsim import initial
    fluidsim.createDevice()
    P = torch.empty((runs, steps + 1, n, n), dtype=torch.float32, device="cuda")
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
                P[run, k] = torch.from_numpy(g.readField("pressure").astype(np.float32).reshape(n, n)).cuda()
        if run % 4 == 0:
            log(f"data run {run + 1}/{runs}")
    return P


def make_batch(P, idx, h):
    import torch
    runs, steps1 = P.shape[0], P.shape[1]
    r_i = idx // (steps1 - 1)
    k_i = idx % (steps1 - 1) + 1
    p = P[r_i, k_i].double()
    pp = P[r_i, k_i - 1].double()
    r = laplacian_t(p, h) - laplacian_t(pp, h)
    sr = r.abs().amax(dim=(-2, -1), keepdim=True).clamp_min(1e-300)
    sp = pp.abs().amax(dim=(-2, -1), keepdim=True).clamp_min(1e-300)
    x = torch.stack([(r / sr).float(), (pp / sp).float()], 1)
    t = ((p - pp) / (sr * h * h)).float().unsqueeze(1)
    return x, t, p, pp, sr


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--runs", type=int, default=32)
    ap.add_argument("--steps", type=int, default=625)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    import torch
    from fluidsim.guess_net import build_model, export_blob

    def log(msg):
        print(f"[train {time.strftime('%H:%M:%S')}] {msg}", flush=True)

    torch.manual_seed(a.seed)
    h = 2 * math.pi / a.n
    t0 = time.time()
    P = generate(a.n, a.runs, a.steps, a.warmup, log)
    log(f"{a.runs * a.steps} pairs at {a.n}^2 in {time.time() - t0:.1f} s")
    npairs = a.runs * a.steps
    hold = torch.arange(npairs, device="cuda")
    val = hold[hold // a.steps >= a.runs - 2]          # the last two runs are held out
    train = hold[hold // a.steps < a.runs - 2]
    model = build_model().cuda()
    opt = torch.optim.Adam(model.parameters(), lr=a.lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, a.iters)
    for it in range(a.iters):
        idx = train[torch.randint(len(train), (a.batch,), device="cuda")]
        x, t, _, _, _ = make_batch(P, idx, h)
        loss = torch.mean((model(x) - t) ** 2)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        if it % 250 == 0 or it == a.iters - 1:
            log(f"iter {it} loss {loss.item():.4e}")
    # held-out check: how far the guess reduces the residual the previous pressure leaves
    model.eval()
    red, red0 = [], []
    with torch.no_grad():
        for c in range(0, len(val), 32):
            idx = val[c:c + 32]
            x, t, p, pp, sr = make_batch(P, idx, h)
            o = model(x).double().squeeze(1)
            pg = pp + sr * h * h * o
            rg = (laplacian_t(p, h) - laplacian_t(pg, h)).abs().amax(dim=(-2, -1))
            red.append((rg / sr.squeeze(-1).squeeze(-1)).cpu())
    red = torch.cat(red)
    stats = {"pairs": npairs, "n": a.n, "iters": a.iters, "final_loss": float(loss.item()),
             "heldout_residual_ratio_median": float(red.median()), "heldout_residual_ratio_mean": float(red.mean())}
    log(json.dumps(stats))
    blob = export_blob(model)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "wb") as f:
        f.write(blob)
    with open(a.out + ".json", "w") as f:
        json.dump(stats, f, indent=1)
    log(f"wrote {a.out} ({len(blob)} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
