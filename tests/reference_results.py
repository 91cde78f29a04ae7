"""The reference's two result harnesses, tools/gputest.html (10 blocks, 16 named
checks) and tools/bench.html (same algorithm, equal quality, interactive ceiling),
restated over this library's Python binding, with the f64 oracle as the CPU side.
Writes gpu-verify.json and bench.json with the reference's keys and value meanings
(results/gpu-verify.json, results/bench.json), so reporting built for those files
reads MI355X runs unchanged.

    python tests/reference_results.py --out results-mi355x [--quick]

This is test infrastructure (it runs the oracle); tests/test_gpu_reference_results.py
runs it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TWO_PI = 2 * math.pi


def _paths():
    for p in (os.path.join(ROOT, "fluid-sim_amd"), os.path.join(ROOT, "oracle")):
        if p not in sys.path:
            sys.path.insert(0, p)


def rel_diff(a, b):
    """relDiff of tools/gputest.html:53-62: relative L2 difference."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    num = float(np.sum((a - b) ** 2))
    den = float(np.sum(b * b))
    return math.sqrt(num / (den or 1.0))


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def _device_info(fs):
    d = fs.createDevice()
    return {"vendor": "amd", "architecture": "gfx950", "device": d["info"]["description"],
            "description": d["info"]["description"]}


# --------------------------------------------------------------- gpu-verify.json

def gpu_verify(fs, O):
    """tools/gputest.html:64-449 block by block; R has the same keys."""
    from fluidsim.solid import RigidDisk
    R = {"checks": [], "errors": []}

    def check(name, ok, detail):
        R["checks"].append({"name": name, "pass": bool(ok), "detail": detail})

    R["adapter"] = _device_info(fs)
    R["hasTimestamps"] = True
    R["limits"] = {"maxTextureDimension2D": 16384}
    N, dt = 64, 0.01

    # 1. advection (gputest :72-107): one SL step of Taylor-Green, no projection
    u0, v0 = O.taylor_green(N, N)
    with fs.GPUFluidSolver(0, n=N, nu=0.0, pressureIterations=0, advection="semi-lagrangian") as g:
        g.uploadVelocity(u0, v0)
        g.runKernel("advect-sl", dt)
        out = g.readVelocity()
    ou, ov = O.advect_velocity("semi-lagrangian", N, N, f32(u0), f32(v0), dt)
    du, dv = rel_diff(out["u"], ou), rel_diff(out["v"], ov)
    check("advection kernel matches CPU", max(du, dv) < 2e-5,
          f"relL2 u={du:.2e} v={dv:.2e} (f32 floor ~1e-6)")
    R["advection"] = {"du": du, "dv": dv}

    # 2. divergence (:110-136), seed 991
    u0, v0 = O.js_random_velocity(N, 991)
    with fs.GPUFluidSolver(0, n=N, nu=0.0) as g:
        g.uploadVelocity(u0, v0)
        g.runKernel("divergence")
        dg = g.readField("div")
    dd = rel_diff(dg, O.divergence(N, N, f32(u0), f32(v0)))
    check("divergence kernel matches CPU", dd < 1e-5, f"relL2={dd:.2e}")
    R["divergence"] = dd

    # 3. projection reduces divergence, equal budgets (:138-186), seed 4242
    u0, v0 = O.js_random_velocity(N, 4242)
    res = {}
    for name in ("jacobi", "red-black"):
        with fs.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver=name, pressureIterations=200) as g:
            g.uploadVelocity(u0, v0)
            before = g.maxDivergence()
            g.step(dt)
            after = g.maxDivergence()
        res[name] = {"before": before, "after": after, "reduction": before / after}
    R["projection"] = res
    rb, jc = res["red-black"], res["jacobi"]
    check("GPU projection substantially reduces divergence", rb["reduction"] > 100,
          f"red-black: {rb['before']:.2e} -> {rb['after']:.2e} ({rb['reduction']:.0f}x)")
    check("red-black clears divergence better than Jacobi at equal cost", rb["after"] < jc["after"],
          f"after 200 sweeps: RB {rb['after']:.2e} vs J {jc['after']:.2e} ({jc['after'] / rb['after']:.1f}x better)")

    # 4. convergence curves (:188-228), seed 777
    curves = {}
    u0, v0 = O.js_random_velocity(N, 777)
    for name in ("jacobi", "red-black", "multigrid"):
        budget = [0, 1, 2, 3, 4, 5, 6, 8, 10, 12] if name == "multigrid" else [0, 1, 2, 5, 10, 20, 40, 80, 160, 320]
        with fs.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver=name) as g:
            g.uploadVelocity(u0, v0)
            curves[name] = [{"iterations": it,
                             "equivalentSweeps": it * (4 / 3) * 4 if name == "multigrid" else it,
                             "residual": g.poissonResidualAfter(it)} for it in budget]
    R["poissonCurves"] = curves
    j, r = curves["jacobi"], curves["red-black"]
    check("red-black Gauss-Seidel converges faster than Jacobi", r[-1]["residual"] < j[-1]["residual"],
          f"at {j[-1]['iterations']} sweeps: RB {r[-1]['residual']:.2e} vs J {j[-1]['residual']:.2e}")
    check("Jacobi residual decreases monotonically",
          all(i == 0 or p["residual"] <= j[i - 1]["residual"] * 1.001 for i, p in enumerate(j)), "")

    # 5. full steps against the CPU (exact FFT projection) (:230-264)
    u0, v0 = O.taylor_green(N, N)
    steps = 10
    with fs.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver="red-black", pressureIterations=400,
                           advection="semi-lagrangian") as g:
        g.uploadVelocity(u0, v0)
        for _ in range(steps):
            g.step(dt)
        out = g.readVelocity()
        gdiv = g.maxDivergence()
    s = O.Solver(n=N, nu=0.0, advection="semi-lagrangian", projection="fft")
    s.u[:], s.v[:] = f32(u0), f32(v0)
    for _ in range(steps):
        s.step(dt)
    du, dv = rel_diff(out["u"], s.u), rel_diff(out["v"], s.v)
    cdiv = O.max_divergence(N, N, s.u, s.v)
    check(f"GPU matches CPU after {steps} full steps", max(du, dv) < 5e-3, f"relL2 u={du:.2e} v={dv:.2e}")
    check("GPU keeps the field divergence-free", gdiv < 1e-3, f"max|div|={gdiv:.2e} (CPU {cdiv:.2e})")
    R["fullStep"] = {"steps": steps, "du": du, "dv": dv, "gpuMaxDiv": gdiv, "cpuMaxDiv": cdiv}

    # 6. MacCormack against the CPU (:267-299)
    with fs.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver="multigrid", pressureIterations=4,
                           advection="maccormack") as g:
        g.uploadVelocity(u0, v0)
        for _ in range(steps):
            g.step(dt)
        out = g.readVelocity()
    s = O.Solver(n=N, nu=0.0, advection="maccormack", projection="fft")
    s.u[:], s.v[:] = f32(u0), f32(v0)
    for _ in range(steps):
        s.step(dt)
    du, dv = rel_diff(out["u"], s.u), rel_diff(out["v"], s.v)
    check(f"GPU MacCormack matches CPU after {steps} steps", max(du, dv) < 5e-3, f"relL2 u={du:.2e} v={dv:.2e}")
    R["macCormack"] = {"steps": steps, "du": du, "dv": dv}

    # 7. MacCormack keeps more energy than SL at nu = 0 (:301-328)
    ke = {}
    for scheme in ("semi-lagrangian", "maccormack"):
        with fs.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver="multigrid", pressureIterations=3,
                               advection=scheme) as g:
            g.uploadVelocity(u0, v0)
            for _ in range(200):
                g.step(0.01)
            out = g.readVelocity()
        ke[scheme] = 0.5 * float(np.mean(out["u"].astype(np.float64) ** 2 + out["v"].astype(np.float64) ** 2))
    check("GPU MacCormack retains more energy than semi-Lagrangian", ke["maccormack"] > ke["semi-lagrangian"],
          f"after 200 steps at nu=0 (exact conserves 0.25): SL {ke['semi-lagrangian']:.4f}, MC {ke['maccormack']:.4f}")
    R["schemeEnergy"] = {**ke, "exact": 0.25}

    # 8. the solid mask on the GPU (:330-374): a pinned disk in a uniform +x stream; the
    # per-cell impulse is the coupling's change of the field (the impulse texture)
    M = 128
    ones, zeros = np.ones(M * M), np.zeros(M * M)
    with fs.GPUFluidSolver(0, n=M, nu=0.0, pressureSolver="multigrid", pressureIterations=3,
                           advection="maccormack") as g:
        g.setSolid(RigidDisk(x=math.pi, y=math.pi, r=0.6, density=8, fixed=True))
        g.uploadVelocity(ones, zeros)
        g.runKernel("solid", dt)
        cell = g.readVelocity()["u"].astype(np.float64) - 1.0
    finite = np.isfinite(cell)
    non_finite = int(np.count_nonzero(~finite))
    non_zero = int(np.count_nonzero(np.abs(cell[finite]) > 1e-9))
    impulse = float(np.sum(cell[finite])) * (TWO_PI / M) ** 2
    check("GPU solid mask is finite everywhere", non_finite == 0, f"{non_finite} non-finite cells of {M * M}")
    check("GPU solid mask covers a plausible area", 100 < non_zero < M * M * 0.5, f"{non_zero} cells carry impulse")
    check("a pinned disk slows the fluid", impulse < 0, f"net fluid impulse {impulse:.2e} (negative expected)")
    R["gpuSolid"] = {"impulse": impulse, "nonFinite": non_finite, "nonZero": non_zero, "n": M}

    # 9. the coupled step and the body force against the CPU (:376-414)
    disk_cfg = dict(x=math.pi, y=math.pi, r=0.6)
    s = O.Solver(n=M, nu=0.0, advection="maccormack", projection="fft", disk=disk_cfg, passes=1)
    s.u[:], s.v[:] = ones, zeros
    s.step(dt)
    # the uniform stream is a fixed point of advection, so the CPU coupling acts on it as is
    f_cpu = -O.apply_coupling(M, M, ones.copy(), zeros.copy(), disk_cfg, passes=1)[0] / dt
    gdisk = RigidDisk(x=math.pi, y=math.pi, r=0.6, density=8, fixed=True)
    with fs.GPUFluidSolver(0, n=M, nu=0.0, pressureSolver="multigrid", pressureIterations=4,
                           advection="maccormack") as g:
        g.setSolid(gdisk)
        g.uploadVelocity(ones, zeros)
        g.step(dt)
        g.integrateSolid(dt)
        out = g.readVelocity()
    du, dv = rel_diff(out["u"], s.u), rel_diff(out["v"], s.v)
    f_gpu = gdisk.force[0]
    f_rel = abs(f_gpu - f_cpu) / (abs(f_cpu) or 1.0)
    check("GPU solid coupling matches the CPU field", max(du, dv) < 5e-3, f"relL2 u={du:.2e} v={dv:.2e}")
    check("GPU and CPU agree on the force the body feels", f_rel < 5e-2,
          f"CPU {f_cpu:.3e} vs GPU {f_gpu:.3e} ({f_rel * 100:.1f}%)")
    R["gpuSolidAgreement"] = {"du": du, "dv": dv, "fCpu": f_cpu, "fGpu": f_gpu, "fRel": f_rel}

    # 10. a free disk is carried downstream (:416-436)
    disk = RigidDisk(x=math.pi, y=math.pi, r=0.55, density=8)
    with fs.GPUFluidSolver(0, n=M, nu=0.0, pressureSolver="multigrid", pressureIterations=3,
                           advection="maccormack") as g:
        g.setSolid(disk)
        g.uploadVelocity(ones, zeros)
        x0 = disk.x
        for _ in range(60):
            g.step(dt)
            g.integrateSolid(dt)
    ok = math.isfinite(disk.vx) and 0.02 < disk.vx < 1.05
    check("a free disk is carried downstream by the flow", ok, f"vx = {disk.vx:.4f}, moved {disk.x - x0:.4f}")
    R["gpuSolidFree"] = {"vx": disk.vx, "vy": disk.vy, "moved": disk.x - x0}
    R["done"] = True
    return R


# -------------------------------------------------------------------- bench.json

def _median(a):
    s = sorted(a)
    return s[len(s) // 2]


def _time_gpu(g, dt, steps):
    """timeGPU (tools/bench.html:59-80): 12 warm-up steps, then each step synchronised."""
    for k in range(12):
        g.step(dt)
        if k % 4 == 3:
            g.sync()
    g.sync()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        g.step(dt)
        g.sync()
        times.append((time.perf_counter() - t0) * 1e3)
    return _median(times)


def _time_cpu(s, dt, steps):
    """timeCPU (tools/bench.html:82-91) on the oracle: 2 warm-up steps."""
    for _ in range(2):
        s.step(dt)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        s.step(dt)
        times.append((time.perf_counter() - t0) * 1e3)
    return _median(times)


def bench_results(fs, O, quick=False):
    """tools/bench.html:93-258; the CPU side is the f64 oracle on one host thread."""
    R = {"runs": [], "errors": [], "adapter": _device_info(fs), "hasTimestamps": True,
         "userAgent": "fluidsim (MI355X, HIP) / f64 oracle", "hardwareConcurrency": os.cpu_count()}
    dt, nu, iters = 0.005, 0.0, 40

    def tg_solver(n, **kw):
        s = O.Solver(n=n, nu=nu, advection="semi-lagrangian", **kw)
        u0, v0 = O.taylor_green(n, n)
        s.u[:], s.v[:] = u0, v0
        return s

    rows = []
    for n in ([32, 64] if quick else [32, 64, 128, 256, 512]):
        s = tg_solver(n, projection="jacobi", iterations=iters)
        steps = 3 if n >= 256 else 6 if n >= 128 else 15
        c = _time_cpu(s, dt, steps)
        with fs.GPUFluidSolver(0, n=n, nu=nu, pressureSolver="jacobi", pressureIterations=iters) as g:
            g.uploadVelocity(s.u, s.v)
            gm = _time_gpu(g, dt, max(steps, 12))
            gdiv = g.maxDivergence()
        rows.append({"n": n, "cells": n * n, "cpuMs": c, "gpuMs": gm, "speedup": c / gm,
                     "cpuMaxDiv": O.max_divergence(n, n, s.u, s.v), "gpuMaxDiv": gdiv})
    R["sameAlgorithm"] = {"pressureIterations": iters, "rows": rows}

    target = 1e-3
    best = []
    for n in ([64] if quick else [64, 128, 256, 512]):
        s = tg_solver(n, projection="fft")
        steps = 4 if n >= 256 else 10
        c = _time_cpu(s, dt, steps)
        res = {}
        for name, budgets in (("red-black", [10, 20, 40, 80, 160, 320, 640, 1280]),
                              ("multigrid", [1, 2, 3, 4, 5, 6, 8])):
            for trial in budgets:
                g = fs.GPUFluidSolver(0, n=n, nu=nu, pressureSolver=name, pressureIterations=trial)
                g.uploadVelocity(s.u, s.v)
                g.step(dt)
                gdiv = g.maxDivergence()
                if gdiv <= target or trial == budgets[-1]:
                    break
                g.destroy()
            ms = _time_gpu(g, dt, max(steps, 10))
            g.destroy()
            res[name] = {"ms": ms, "budget": trial, "div": gdiv, "reachedTarget": gdiv <= target}
        best.append({"n": n, "cpuFftMs": c, "cpuMaxDiv": O.max_divergence(n, n, s.u, s.v),
                     "redBlack": res["red-black"], "multigrid": res["multigrid"],
                     "ratioRedBlack": c / res["red-black"]["ms"], "ratioMultigrid": c / res["multigrid"]["ms"]})
    R["equalQuality"] = {"targetDivergence": target, "rows": best}

    frame = 1000.0 / 60.0
    inter = {"gpu": None, "cpu": None, "gpuRows": [], "cpuRows": []}
    for n in ([128, 256] if quick else [128, 256, 512, 1024, 2048]):
        with fs.GPUFluidSolver(0, n=n, nu=nu, pressureSolver="red-black", pressureIterations=20) as g:
            u0, v0 = O.taylor_green(n, n)
            g.uploadVelocity(u0, v0)
            t = _time_gpu(g, dt, 24)
        inter["gpuRows"].append({"n": n, "ms": t})
        if t <= frame:
            inter["gpu"] = n
        if t > 4 * frame:
            break
    for n in ([32] if quick else [32, 64, 128, 256]):
        s = tg_solver(n, projection="fft")
        t = _time_cpu(s, dt, 4 if n >= 256 else 16)
        inter["cpuRows"].append({"n": n, "ms": t})
        if t <= frame:
            inter["cpu"] = n
        if t > 4 * frame:
            break
    R["interactive"] = inter
    R["done"] = True
    return R


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True, help="directory for gpu-verify.json and bench.json")
    ap.add_argument("--quick", action="store_true", help="the smallest sizes of each bench section only")
    a = ap.parse_args(argv)
    _paths()
    import fluidsim
    import oracle as O
    os.makedirs(a.out, exist_ok=True)
    v = gpu_verify(fluidsim, O)
    with open(os.path.join(a.out, "gpu-verify.json"), "w") as f:
        json.dump(v, f, indent=1)
    passed = sum(c["pass"] for c in v["checks"])
    print(f"gpu-verify: {passed}/{len(v['checks'])} checks passed", flush=True)
    b = bench_results(fluidsim, O, quick=a.quick)
    with open(os.path.join(a.out, "bench.json"), "w") as f:
        json.dump(b, f, indent=1)
    print(f"bench: interactive ceiling GPU {b['interactive']['gpu']}^2, CPU {b['interactive']['cpu']}^2", flush=True)
    return 0 if passed == len(v["checks"]) else 1


if __name__ == "__main__":
    sys.exit(main())
