#!/usr/bin/env python3
"""Throughput of the stable-fluids step on one MI355X.

Workload (the configuration the headline metric is quoted on): decaying 2D
turbulence on a 1024^2 periodic MAC grid, FP32 fields, nu = 1e-4, MacCormack
advection + explicit diffusion, dt at CFL 0.5 evaluated on the device each step,
projection by multigrid V(2,2) red-black iterated until the FP64 residual
max|div u* - L p| <= 1e-6 after every projection. The initial state is seeded
random-phase (synthetic). The V-cycle runs over 1024^2 .. 64^2 and solves the 64^2
level exactly (--coarse exact, the default); --coarse sweeps runs the reference's
9 levels to minSize 4 with 40 coarse sweeps instead (both are reported, the other
in the ablation).

A step is one full advect/diffuse/project pass. `value` is cell-steps/s over the
timed steps (inputs already resident on the device). The roofline object is
measured with HIP events on the solver's own stream over a second, profiled run
of the same number of steps; the CPU baseline is the f64 oracle timed on this
host on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fluid-sim_amd"))

METRIC = "cell-steps/s at 1024² to max|∇·u|≤1e-6, learned-MG vs plain MG; HBM GB/s"
HBM_PEAK_GBS = 8000.0

# Kernel classes timed for the roofline. The library counts each launch's
# algorithmic bytes: every distinct input element read once and every output
# written once at its storage precision (FP32 velocity and correction, FP64
# pressure; the coarse right-hand side is a quarter cell per cell).
KIND_TEXT = {
    "advect_sl": "semi-Lagrangian forward trace (k_advect_sl), 16 B/cell",
    "advect_mc": "MacCormack back-trace + limiter (k_advect_mc), 24 B/cell",
    "diffuse": "explicit diffusion (k_diffuse), 16 B/cell",
    "down0": "level-0 residual in FP64 + pre-smoothing + restriction (k_mgt_down0), 37-45 B/cell",
    "down": "level>0 pre-smoothing + restriction (k_mgt_down), 9 B/cell of the level",
    "coarse": "coarse levels in one workgroup (k_mg_coarse_ml), 8 B/cell of its top level",
    "up": "prolongation + post-smoothing (k_mgt_up, k_mgt_up2: two levels), 13 B/cell of the level",
    "up_coarse": "two levels up with the exact 64^2 coarse solve in every workgroup (k_mgt_up2, folded)",
    "gradient": "FP64 pressure-gradient subtraction (k_gradient64), 24 B/cell",
    "residual": "FP64 residual only (k_residual64), 20-32 B/cell",
    "rb": "red-black colour sweep (k_mg_rb), 6 B/cell",
    "guess": "learned-guess 3x3 convolutions (k_conv3x3, BF16 MFMA), BF16 activations in + out",
    "advect_fused": "fused MacCormack + diffusion (k_advect_mc_diffuse), 16 B/cell: u, v in + out",
    "forcing": "band forcing + drag (k_forcing_sep), 16 B/cell: u, v in + out",
}

# The kernels of each class as the profiler names them, for the HBM counter passes.
KIND_KERNELS = {
    "advect_sl": r"k_advect_sl\b", "advect_mc": r"k_advect_mc\b", "diffuse": r"k_diffuse\b",
    "down0": r"k_mgt_down0q?<", "down": r"k_mgt_down2?<", "coarse": r"k_mg_coarse_(ml|exact)<",
    "up": r"k_mgt_up(<|2q?<[^(]*false>)", "up_coarse": r"k_mgt_up2q?<[^(]*true>",
    "gradient": r"k_gradient64\b", "residual": r"k_residual64<", "rb": r"k_(mg_rb|rb_colour)<",
    "advect_fused": r"k_advect_mc_diffuse\b", "forcing": r"k_forcing_(tables|sep)\b",
}
CAL_BYTES = 256 << 20   # one plain device copy of the counter calibration
CAL_COPIES = 5


# Workloads: the headline one (default) and the forced, mixed-precision one at 2048^2
# (band forcing kf = 32 +- 2, 24 modes, amplitude 0.1, seed 1000 + step, drag 0.02;
# BF16-stored smoother, FP32 arithmetic, FP64 refinement residual).
WORKLOADS = {
    "decaying1024": dict(n=1024, nu=1e-4, forcing=None, precision="fp32"),
    "forced2048": dict(n=2048, nu=5e-5, forcing=dict(kf=32.0, width=2.0, amp=0.1, drag=0.02, seed=1000),
                       precision="bf16"),
    # 256 lid-driven cavities (sizes 64..512, lid U ~ U[0.5, 2], Re ~ U[100, 5000], a pinned
    # disk each), every stage one launch over the whole batch
    "cavities256": dict(n=None, nu=None, forcing=None, precision="fp32"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="decaying1024", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default=None, choices=["fp32", "bf16"],
                    help="multigrid storage precision (default: the workload's)")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--n", type=int, default=None, help="grid size (default: the workload's)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--coarse", default="exact", choices=["sweeps", "exact"],
                    help="coarsest level: the reference's V-cycle down to minSize 4 with 40 sweeps, or the hierarchy "
                         "stopped at 64^2 and that level solved exactly")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--guess", default="zero",
                    choices=["zero", "previous", "extrapolated", "learned", "learned-coarse", "learned-quarter"])
    ap.add_argument("--no-target", action="store_true",
                    help="skip the forced-1024^2 comparison of learned-guess BF16 multigrid with plain FP32 multigrid")
    ap.add_argument("--no-ablation", action="store_true",
                    help="skip the second timed run with the other pressure guess")
    ap.add_argument("--report", action="store_true",
                    help="forced2048: time-averaged E(k) (device spectrum, Parseval) over a second run and the "
                         "field error against the f64 oracle over --validate-steps steps")
    ap.add_argument("--validate-steps", type=int, default=20)
    ap.add_argument("--display", type=int, default=0, metavar="CALLS",
                    help="after the timed steps, time CALLS render() calls with index.html's parameters, then "
                         "run 20 steps with vorticity confinement on (for a kernel trace of k_confine)")
    ap.add_argument("--no-copy-ceiling", action="store_true",
                    help="skip timing plain 1 GiB device copies for the roofline's copy_ceiling_gbs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device-to-device copies of 256 MiB (counter calibration) and exit")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) behind roofline.traffic")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run --warmup steps, then a profiled pass of 10 steps, print its launch counts and "
                         "exit (the program the counter passes run)")
    return ap.parse_args()


def copy_calibration():
    """Five plain device copies of 256 MiB through the HIP runtime, for scaling HBM counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 0, C.c_size_t(nbytes))
    for _ in range(5):
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"copy calibration: 5 copies of {nbytes} bytes", flush=True)


def copy_ceiling():
    """GB/s (read + write) of plain device-to-device copies of 1 GiB through the HIP
    runtime, timed with HIP events: what a streaming kernel can reach on this GPU."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = 1 << 30
    a, b, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        hip.hipMemset(a, 0, C.c_size_t(nbytes))
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
        reps = 10
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3):
                raise RuntimeError("hipMemcpy failed")
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def _counter_pass(counter, argv, timeout_s=90):
    """Run `python3 bench.py <argv>` under `rocprofv3 --pmc <counter>` in a child process
    group of its own; returns (its stdout, [(dispatch id, kernel name, value)]) or None."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    out = tempfile.mkdtemp(prefix="fs_pmc_", dir="/tmp")
    cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "pmc", "--",
           sys.executable, os.path.join(ROOT, "bench.py")] + argv
    env = dict(os.environ, TMPDIR="/tmp")
    p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         start_new_session=True)
    try:
        text, _ = p.communicate(timeout=timeout_s)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.communicate()
        print(f"# counter pass {counter} {' '.join(argv)}: no result in {timeout_s} s", file=sys.stderr, flush=True)
        shutil.rmtree(out, ignore_errors=True)
        return None
    rows = []
    for f in glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    rows.append((int(r.get("Dispatch_Id") or len(rows)), r.get("Kernel_Name", ""),
                                 float(r["Counter_Value"])))
    shutil.rmtree(out, ignore_errors=True)
    if p.returncode != 0 or not rows:
        print(f"# counter pass {counter} {' '.join(argv)}: exit {p.returncode}, {len(rows)} rows",
              file=sys.stderr, flush=True)
        return None
    rows.sort()
    return text, rows


def hbm_traffic(a, kind):
    """HBM bytes per launch of one kernel class, measured in two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE) over a profiled pass of the same workload run by this
    script in a child process. Each counter is scaled by what it reads for plain
    device-to-device copies of 256 MiB (--copy-calibration), as bytes read and written by
    the copy over the counter's KiB. The class's launches are the last dispatches of its
    kernels (the probe ends with its profiled pass), so the autotuner's trial launches at
    create are left out. None if rocprofv3 is missing or a pass fails."""
    import re
    pat = KIND_KERNELS.get(kind)
    if pat is None:
        return None
    probe = ["--traffic-probe", "--workload", a.workload, "--warmup", str(a.warmup), "--tol", str(a.tol),
             "--guess", a.guess, "--coarse", a.coarse] + (["--n", str(a.n)] if a.n else []) + (["--precision", a.precision]
                                                                        if a.precision else [])
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = _counter_pass(counter, ["--copy-calibration"])
        run = _counter_pass(counter, probe)
        if cal is None or run is None:
            return None
        cal_kib = sum(v for _, name, v in cal[1] if "copyBuffer" in name)
        if cal_kib <= 0:
            return None
        factor = CAL_COPIES * CAL_BYTES / (cal_kib * 1024.0)
        m = re.search(r"^# probe launches (\{.*\})$", run[0], re.M)
        launches = json.loads(m.group(1)).get(kind, 0) if m else 0
        mine = [v for _, name, v in run[1] if re.search(pat, name)]
        if launches <= 0 or len(mine) < launches:
            return None
        res[counter] = (sum(mine[-launches:]) * 1024.0 * factor / launches, factor, launches)
    return {"bytes": res["FETCH_SIZE"][0] + res["WRITE_SIZE"][0], "read_bytes": res["FETCH_SIZE"][0],
            "write_bytes": res["WRITE_SIZE"][0], "read_factor": res["FETCH_SIZE"][1],
            "write_factor": res["WRITE_SIZE"][1], "launches": res["FETCH_SIZE"][2],
            "counters": "rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE, scaled by plain copies of 256 MiB"}


GUESS_WEIGHTS = os.path.join(ROOT, "fluid-sim_amd", "weights", "guess_unet_256.fsgn")


def make_solver(fluidsim, wl, n, tol, guess="zero", coarse="sweeps"):
    g = fluidsim.GPUFluidSolver(0, n=n, nu=wl["nu"], advection="maccormack", pressureSolver="multigrid",
                                pressureIterations=60, tolerance=tol, mgMinSize=4, mgPre=2, mgPost=2,
                                mgCoarse=40, guess="zero" if guess.startswith("learned") else guess,
                                mgPrecision=wl["precision"], mgCoarseSolve=coarse)
    if guess.startswith("learned"):
        g.loadGuess(GUESS_WEIGHTS)
        g.setGuess(guess)
    if wl["forcing"]:
        g.setForcing(wl["forcing"])
    return g


def cpu_baseline(wl, n, seconds):
    """f64 oracle (single thread, then all host cores) on the same workload, exact FFT projection."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    from fluidsim import initial
    u0, v0 = initial.decaying_turbulence(n, seed=1)
    h = 2 * math.pi / n
    out = {}
    for label, threads in (("single", 1), ("omp", min(16, os.cpu_count() or 1))):
        O.set_threads(threads)
        f = wl["forcing"]
        s = O.Solver(n=n, nu=wl["nu"], advection="maccormack", explicit_diffusion=True, projection="fft",
                     forcing=f, drag=f["drag"] if f else 0.0)
        s.u[:] = u0.astype(np.float32)
        s.v[:] = v0.astype(np.float32)
        budget = seconds if label == "single" else max(2.0, seconds / 4)
        steps = 0
        t0 = time.perf_counter()
        while True:
            dt = 0.5 / (np.max(np.abs(s.u)) / h + np.max(np.abs(s.v)) / h)
            s.step(dt, force_seed=(f["seed"] + steps) if f else 0)
            steps += 1
            el = time.perf_counter() - t0
            if el >= budget or steps >= 200:
                break
        out[label] = (n * n * steps / el, threads, steps, el)
    O.set_threads(1)
    v, c, st, el = out["single"]
    vo, co, sto, elo = out["omp"]
    return {"value": v, "unit": "cell-steps/s", "cores": c, "kind": "port",
            "sample": f"f64 oracle (oracle/fsoracle.c), same {n}^2 workload with exact FFT projection, "
                      f"{st} steps in {el:.1f} s on 1 thread",
            "omp": {"value": vo, "cores": co, "sample": f"{sto} steps in {elo:.1f} s"}}


def forced_report(a, fluidsim, wl, n, u0, v0):
    """Statistics of the workload: E(k) averaged over samples every 50 steps of a run as
    long as the timed one (device spectrum, FP64), its Parseval shortfall and the
    log-log slope over k in [2 kf, n/8]; and the FP32 field error against the f64
    oracle (exact FFT projection, same forcing seeds and dt) over the first steps."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    g = make_solver(fluidsim, wl, n, a.tol, "previous")
    g.uploadVelocity(u0, v0)
    f = wl["forcing"]
    # field error over the first steps
    O.set_threads(min(16, os.cpu_count() or 1))
    s = O.Solver(n=n, nu=wl["nu"], advection="maccormack", explicit_diffusion=True, projection="fft",
                 forcing=f, drag=f["drag"] if f else 0.0)
    s.u[:] = u0.astype(np.float32)
    s.v[:] = v0.astype(np.float32)
    errs = []
    t0 = time.perf_counter()
    for k in range(a.validate_steps):
        g.stepCFL(0.5)
        dt = g.stats()["last_dt"]
        s.step(dt, force_seed=(f["seed"] + k) if f else 0)
        if k in (0, a.validate_steps // 2, a.validate_steps - 1):
            vel = g.readVelocity()
            den = math.sqrt(np.sum(s.u ** 2) + np.sum(s.v ** 2))
            errs.append({"step": k + 1, "relL2": math.sqrt(np.sum((vel["u"] - s.u) ** 2) +
                                                            np.sum((vel["v"] - s.v) ** 2)) / den})
    O.set_threads(1)
    print(f"# field error vs f64 oracle: {errs} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    # time-averaged spectrum
    Es = []
    for k in range(a.steps):
        g.stepCFL(0.5)
        if (k + 1) % 50 == 0:
            _, E, _ = g.energySpectrum()
            Es.append(E)
    d = g.diagnostics()
    _, Elast, cnt = g.energySpectrum()
    g.destroy()
    E = np.mean(Es, axis=0)
    kf = f["kf"] if f else 8.0
    k = np.arange(len(E))
    sel = (k >= 2 * kf) & (k <= n // 8) & (E > 0)
    slope = float(np.polyfit(np.log(k[sel]), np.log(E[sel]), 1)[0]) if sel.sum() > 3 else None
    parseval = (d["kinetic_energy"] - float(np.sum(Elast))) / d["kinetic_energy"]
    shells = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024]
    rep = {"samples": len(Es), "every": 50, "E_mean": {str(q): float(E[q]) for q in shells if q < len(E)},
           "slope_2kf_to_n8": slope, "parseval_shortfall_last": parseval, "field_error_vs_f64": errs}
    print(f"# spectrum: {len(Es)} samples, slope {slope}, Parseval shortfall {parseval:.3g}", file=sys.stderr,
          flush=True)
    return rep


def bench_cavities(a, fluidsim, dev):
    """256 batched cavities: cell-steps/s over the whole batch (sum of n^2 x steps / s)."""
    cavs = fluidsim.random_cavities(256, seed=1)
    guess = a.guess if a.guess in ("zero", "previous") else "zero"
    b = fluidsim.CavityBatch(cavs, tolerance=a.tol, guess=guess, coarseSolve=a.coarse)
    cells = b.stats()["cells"]
    print(f"# device: {dev['info']['description']}; {len(cavs)} cavities, {cells} cells", file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        b.stepCFL(0.5)
    b.sync()
    s0 = b.stats()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        b.stepCFL(0.5)
    b.sync()
    el = time.perf_counter() - t0
    s1 = b.stats()
    cycles = (s1["total_vcycles"] - s0["total_vcycles"]) / max(a.steps, 1)
    value = cells * a.steps / el
    print(f"# timed {a.steps} steps in {el:.3f} s: {value:.4g} cell-steps/s, {cycles:.2f} V-cycles/step, "
          f"last residual {s1['last_residual']:.3g}", file=sys.stderr, flush=True)
    b.destroy()
    line = {"metric": "cell-steps/s, 256 batched lid-driven cavities to max|div u* - L p| <= tol",
            "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": el / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "256 lid-driven cavities, sizes {64,128,256,512}^2 (splitmix64 seed 1), lid "
                                   "U ~ U[0.5,2], Re ~ U[100,5000], pinned disk r ~ U[0.05,0.2] each, MacCormack + "
                                   "explicit diffusion, dt per cavity at CFL 0.5, Neumann multigrid V(2,2) to the "
                                   f"FP64 gate {a.tol:g} over all cavities, {guess} initial guess, each cavity's "
                                   + ("64^2 level solved exactly" if a.coarse == "exact" else
                                      "levels 64^2 .. 4^2 in one workgroup"),
                       "cells": cells, "vcycles_per_step": cycles, "guess": guess, "coarse_solve": a.coarse}}
    print(json.dumps(line), flush=True)
    return 0


# The learned-guess target: band-forced turbulence at 1024^2 (the 2048^2 forcing band
# kf = 32 +- 2 scaled to the grid: kf = 16 +- 1; 24 modes, amplitude 0.1, drag 0.02,
# seed 1000 + step), nu = 1e-4, the same multigrid and FP64 gate as the headline line.
TARGET_WORKLOAD = dict(n=1024, nu=1e-4, forcing=dict(kf=16.0, width=1.0, amp=0.1, drag=0.02, seed=1000),
                       precision="fp32")
TARGET_RUNS = [("zero", "fp32"), ("learned", "bf16"), ("learned", "fp32"), ("previous", "fp32"),
               ("previous", "bf16"), ("zero", "bf16"), ("learned-coarse", "bf16"), ("learned-coarse", "fp32"),
               ("learned-quarter", "bf16"), ("learned-quarter", "fp32"), ("extrapolated", "fp32"),
               ("extrapolated", "bf16")]


def target_comparison(a, fluidsim, initial):
    """Learned-guess mixed-precision multigrid against plain FP32 multigrid (zero guess)
    on forced 1024^2 turbulence, same steps and timing as the headline line, every run
    with the headline's coarse solve; plus plain FP32 multigrid with the reference's
    V-cycle down to 4^2 (40 coarse sweeps), when the headline uses the exact one."""
    n = TARGET_WORKLOAD["n"]
    if not os.path.exists(GUESS_WEIGHTS):
        return None
    u0, v0 = initial.decaying_turbulence(n, seed=1)
    runs = []
    plan = [(g, p, a.coarse) for g, p in TARGET_RUNS]
    if a.coarse != "sweeps":
        plan.append(("zero", "fp32", "sweeps"))
    for guess, prec, coarse in plan:
        wl = dict(TARGET_WORKLOAD, precision=prec)
        g = make_solver(fluidsim, wl, n, a.tol, guess, coarse)
        g.uploadVelocity(u0, v0)
        for _ in range(a.warmup):
            g.stepCFL(0.5)
        g.sync()
        s0 = g.stats()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            g.stepCFL(0.5)
        g.sync()
        el = time.perf_counter() - t0
        s1 = g.stats()
        runs.append({"guess": guess, "mg_storage": prec, "coarse": coarse, "value": n * n * a.steps / el,
                     "ms_per_step": el / a.steps * 1e3,
                     "vcycles_per_step": (s1["total_vcycles"] - s0["total_vcycles"]) / max(a.steps, 1),
                     "last_residual": s1["last_residual"]})
        g.destroy()
        r = runs[-1]
        print(f"# target {guess:15s} {prec} {coarse:6s}: {r['value']:.4g} cell-steps/s, "
              f"{r['vcycles_per_step']:.2f} V-cycles/step, {r['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)

    def val(guess, prec, coarse=a.coarse):
        return next(r["value"] for r in runs if (r["guess"], r["mg_storage"], r["coarse"]) == (guess, prec, coarse))
    plain = val("zero", "fp32")
    learned = {k: val(k, "bf16") for k in ("learned", "learned-coarse", "learned-quarter")}
    out = {"workload": "band-forced 2D turbulence 1024^2 (kf 16 +- 1, 24 modes, amp 0.1, drag 0.02), nu=1e-4, "
                       f"MacCormack, multigrid V(2,2) ({a.coarse} coarse solve) to FP64 max|div u* - L p| <= tol",
           "target": "learned-guess BF16-storage multigrid >= 2x the cell-steps/s of plain FP32 multigrid",
           "ratio_learned_bf16_to_plain_fp32": learned["learned"] / plain,
           "ratio_learned_coarse_bf16_to_plain_fp32": learned["learned-coarse"] / plain,
           "ratio_learned_quarter_bf16_to_plain_fp32": learned["learned-quarter"] / plain,
           "met": max(learned.values()) / plain >= 2.0, "runs": runs}
    if a.coarse != "sweeps":
        ref = val("zero", "fp32", "sweeps")
        out["ratio_best_learned_bf16_to_reference_vcycle_fp32"] = max(learned.values()) / ref
        out["ratio_plain_fp32_to_reference_vcycle_fp32"] = plain / ref
    return out


def display_timing(g, n, calls):
    """render() as index.html:250-251 calls it (vorticity ramp, bloom threshold 0.62, two
    blur rounds), `calls` times back to back; then 20 steps with vorticity confinement on."""
    g.sync()
    t0 = time.perf_counter()
    for _ in range(calls):
        g.render(mode=0, scale=1.5, bloomThreshold=0.62, bloomPasses=2)
    g.sync()
    per = (time.perf_counter() - t0) / calls
    # algorithmic bytes: the render reads u, v and writes an rgba half pixel (16 B per cell,
    # no dye here); the bright pass reads 4 and writes 1 half pixel per half-resolution
    # pixel (40 B); four blur passes read and write one each (16 B)
    hb = (n // 2) * (n // 2)
    nbytes = 16 * n * n + 40 * hb + 4 * 16 * hb
    g.confinement = 1.0
    for _ in range(20):
        g.stepCFL(0.5)
    g.sync()
    g.confinement = 0.0
    return {"render_ms": per * 1e3, "gbs": nbytes / per / 1e9, "calls": calls,
            "algorithmic_bytes": nbytes, "launches": 6,
            "params": "mode 0, scale 1.5, bloomThreshold 0.62, bloomPasses 2 (index.html:250-251)"}


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return 0
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); this build has no multi-GPU path. "
              "Run with --gpus 1.", file=sys.stderr)
        return 2
    import fluidsim
    from fluidsim import initial
    dev = fluidsim.createDevice()
    if a.workload == "cavities256":
        return bench_cavities(a, fluidsim, dev)
    wl = dict(WORKLOADS[a.workload])
    if a.precision:
        wl["precision"] = a.precision
    n = a.n or wl["n"]
    u0, v0 = initial.decaying_turbulence(n, seed=1)
    g = make_solver(fluidsim, wl, n, a.tol, a.guess, a.coarse)
    g.uploadVelocity(u0, v0)
    print(f"# device: {dev['info']['description']}; lib {fluidsim.library_path()}", file=sys.stderr, flush=True)
    tiles = g.tileReport()
    for ln in tiles.strip().splitlines():
        print(f"# tiles {ln}", file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        g.stepCFL(0.5)
    g.sync()
    if a.traffic_probe:
        g.profileEnable()
        for _ in range(10):
            g.stepCFL(0.5)
        g.sync()
        counts = {k: v[1] for k, v in g.profileRead().items()}
        g.destroy()
        print("# probe launches " + json.dumps(counts), flush=True)
        return 0
    st0 = g.stats()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        g.stepCFL(0.5)
    g.sync()
    el = time.perf_counter() - t0
    st1 = g.stats()
    cycles = (st1["total_vcycles"] - st0["total_vcycles"]) / max(a.steps, 1)
    value = n * n * a.steps / el
    d = g.diagnostics()
    print(f"# timed {a.steps} steps in {el:.3f} s: {value:.4g} cell-steps/s, {cycles:.2f} V-cycles/step, "
          f"last residual {st1['last_residual']:.3g}, stored-field max|div| {d['max_divergence']:.3g}",
          file=sys.stderr, flush=True)

    roof = None
    if not a.no_roofline:
        psteps = max(10, min(a.steps, 200))
        g.profileEnable()          # every kernel class, one launch-timed pass
        for _ in range(psteps):
            g.stepCFL(0.5)
        g.sync()
        prof = g.profileRead()
        g.profileEnable([])
        step_ms = el / a.steps * 1e3
        for k, (ms_k, l_k, b_k) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
            if l_k:
                print(f"# {k:10s} {ms_k / l_k * 1e3:8.2f} us/launch {l_k / psteps:6.2f} launches/step "
                      f"{ms_k / psteps:7.4f} ms/step ({ms_k / psteps / step_ms:5.1%}) "
                      f"{b_k / (ms_k / 1e3) / 1e9:7.0f} GB/s algorithmic", file=sys.stderr, flush=True)
        # dominant kernel = the class with the largest share of the step's time
        dom = max(prof, key=lambda k: prof[k][0])
        ms, launches, nbytes = prof[dom]
        avg_s = ms / 1e3 / max(launches, 1)
        achieved = nbytes / (ms / 1e3) / 1e9
        detail = None if a.no_traffic else hbm_traffic(a, dom)
        traffic = detail["bytes"] if detail else None
        if detail:
            print(f"# {dom} HBM traffic {traffic / 1e6:.2f} MB per launch (read {detail['read_bytes'] / 1e6:.2f}, "
                  f"write {detail['write_bytes'] / 1e6:.2f}) against {nbytes / max(launches, 1) / 1e6:.2f} MB "
                  "algorithmic", file=sys.stderr, flush=True)
        roof = {"bound": "hbm", "kernel": KIND_TEXT[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_detail": detail,
                "avg_launch_us": avg_s * 1e6, "launches_per_step": launches / psteps,
                "algorithmic_bytes_per_launch": nbytes / max(launches, 1),
                "share_of_step": ms / psteps / step_ms,
                # every timed kernel class summed per step (the profiled pass launches one by
                # one with events; its kernels are the timed steps' kernels), against ms_per_step
                "kernels_ms_per_step": sum(v[0] for v in prof.values()) / psteps,
                # plain 1 GiB device copies on the same GPU: the streaming ceiling in practice
                "copy_ceiling_gbs": None if a.no_copy_ceiling else copy_ceiling(),
                "all_kernels": {k: {"us_per_launch": v[0] / max(v[1], 1) * 1e3, "launches_per_step": v[1] / psteps,
                                    "gbs": v[2] / (v[0] / 1e3) / 1e9 if v[0] else 0.0}
                                for k, v in prof.items() if v[1]}}
    display = None
    if a.display > 0:
        display = display_timing(g, n, a.display)
        print(f"# render + bloom: {display['render_ms'] * 1e3:.1f} us per call, {display['gbs']:.0f} GB/s "
              "algorithmic", file=sys.stderr, flush=True)
    g.destroy()

    # ablation: the same workload and timing with the other initial pressure guesses
    ablation = None
    if not a.no_ablation:
        ablation = []
        # the other initial guesses with this coarse solve, then this guess with the other one
        other_coarse = "sweeps" if a.coarse == "exact" else "exact"
        others = [(x, a.coarse) for x in ("zero", "previous", "extrapolated", "learned", "learned-coarse",
                                          "learned-quarter") if x != a.guess] + [(a.guess, other_coarse)]
        for other, coarse in others:
            if other.startswith("learned") and (n % {"learned": 128, "learned-coarse": 256}.get(other, 512)
                                                or not os.path.exists(GUESS_WEIGHTS)):
                continue
            g2 = make_solver(fluidsim, wl, n, a.tol, other, coarse)
            g2.uploadVelocity(u0, v0)
            for _ in range(a.warmup):
                g2.stepCFL(0.5)
            g2.sync()
            s0 = g2.stats()
            t0 = time.perf_counter()
            for _ in range(a.steps):
                g2.stepCFL(0.5)
            g2.sync()
            el2 = time.perf_counter() - t0
            s1 = g2.stats()
            c2 = (s1["total_vcycles"] - s0["total_vcycles"]) / max(a.steps, 1)
            ablation.append({"guess": other, "coarse": coarse, "value": n * n * a.steps / el2,
                             "ms_per_step": el2 / a.steps * 1e3, "vcycles_per_step": c2,
                             "last_residual": s1["last_residual"]})
            g2.destroy()
            print(f"# {other} guess, {coarse} coarse: {ablation[-1]['value']:.4g} cell-steps/s, {c2:.2f} V-cycles/step, "
                  f"{ablation[-1]['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)

    target = None
    if not a.no_target and a.workload == "decaying1024" and not a.n:
        target = target_comparison(a, fluidsim, initial)

    report = None
    if a.report:
        report = forced_report(a, fluidsim, wl, n, u0, v0)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(wl, n, a.cpu_seconds)
        print(f"# cpu baseline: {cpu['value']:.4g} cell-steps/s (1 thread), "
              f"{cpu['omp']['value']:.4g} ({cpu['omp']['cores']} threads)", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"decaying 2D turbulence {n}^2" if not wl["forcing"] else
                         f"band-forced 2D turbulence {n}^2 (kf 32 +- 2, 24 modes, amp 0.1, seed 1000 + step, "
                         "drag 0.02)") +
                        f" periodic MAC grid, nu={wl['nu']:g}, MacCormack + explicit diffusion, CFL 0.5, plain "
                        + ("multigrid V(2,2) red-black over 9 levels to minSize 4 (40 coarse sweeps)"
                           if a.coarse == "sweeps" else
                           "multigrid V(2,2) red-black over 5 levels, the 64^2 level solved exactly (real Fourier "
                           "basis on the matrix cores) in place of the 4 levels below it") +
                        f" to FP64 max|div u* - L p| <= {a.tol:g} every step, {wl['precision']} smoother storage, "
                        f"{a.guess} initial guess",
            "n": n,
            "coarse_solve": a.coarse,
            "mg_storage": wl["precision"],
            "residual_tolerance": a.tol,
            "residual_dtype": "f64",
            "vcycles_per_step": cycles,
            "guess": a.guess,
            "stored_field_max_div": d["max_divergence"],
            # timed steps whose projection ran gated on the device (no host read-back), those
            # whose predicted V-cycles fell short of the tolerance, and step-graph replays
            "gated_steps": st1["gated_steps"] - st0["gated_steps"],
            "gate_misses": st1["gate_misses"] - st0["gate_misses"],
            "graph_launches": st1["graph_launches"] - st0["graph_launches"],
            "graph_captures": st1["graph_captures"] - st0["graph_captures"],
            "tiles": [ln.split(" | ")[0] for ln in tiles.strip().splitlines()],
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "ablation": ablation,
        "target": target,
        "report": report,
    }
    if display:
        line["display"] = display
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
