"""The Node drop-in (fluid-sim_amd/node/solver.mjs -> fluidsim.node -> libfluidsim.so)
runs the same step as the Python binding and produces the same bits."""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_node_shim_matches_python_binding_bitwise():
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed on this machine")
    out = subprocess.run([node, os.path.join(ROOT, "fluid-sim_amd", "node", "smoke.mjs")],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    js = json.loads(out.stdout.strip().splitlines()[-1])
    import fluidsim
    n = 64
    h = 2 * math.pi / n
    i = np.arange(n)[None, :]
    j = np.arange(n)[:, None]
    u = (np.sin(i * h) * np.cos((j + 0.5) * h)).reshape(-1)
    v = (-np.cos((i + 0.5) * h) * np.sin(j * h)).reshape(-1)
    with fluidsim.GPUFluidSolver(0, n=n, nu=0.01, advection="maccormack", pressureSolver="multigrid",
                                 pressureIterations=20, tolerance=1e-6) as g:
        g.uploadVelocity(u, v)
        for _ in range(5):
            g.step(0.01)
        vel = g.readVelocity()
        st = g.stats()
    k = np.arange(n * n)
    su = float(np.sum(vel["u"].astype(np.float64) * (k % 97 + 1)))
    sv = float(np.sum(vel["v"].astype(np.float64) * (k % 89 + 1)))
    # same library, same inputs: identical velocities; the JS and numpy sums of them
    # differ only in summation order
    assert abs(js["checksumU"] - su) <= 1e-9 * max(1.0, abs(su))
    assert abs(js["checksumV"] - sv) <= 1e-9 * max(1.0, abs(sv))
    assert js["steps"] == 5 and abs(js["t"] - 0.05) < 1e-15
    assert js["lastResidual"] == st["last_residual"] and js["lastResidual"] <= 1e-6
    assert js["unknownSelectorThrows"] is True
    # the disk path: same library, the disk state pushed and integrated by each binding
    from fluidsim.solid import RigidDisk
    disk = RigidDisk(x=2.0, y=3.5, r=0.6, density=2.5, vx=0.1)
    with fluidsim.GPUFluidSolver(0, n=n, nu=0.01, advection="maccormack", pressureSolver="multigrid",
                                 pressureIterations=20, tolerance=1e-6) as g:
        g.uploadVelocity(u, v)
        g.setSolid(disk, passes=2)
        for _ in range(5):
            g.step(0.01)
            g.integrateSolid(0.01)
        vel = g.readVelocity()
    du = float(np.sum(vel["u"].astype(np.float64) * (k % 97 + 1)))
    assert abs(js["disk"]["checksumU"] - du) <= 1e-9 * max(1.0, abs(du))
    for key in ("x", "y", "vx", "vy", "omega"):
        assert abs(js["disk"][key] - getattr(disk, key)) <= 1e-12 * max(1.0, abs(getattr(disk, key))), key


@pytest.mark.parametrize("mode", ["learned-quarter", "learned"])
def test_node_shim_runs_the_learned_guess_like_the_python_binding(mode):
    # loadGuess / setGuess through the addon: the same weights, field and steps as the
    # Python binding, the same bits
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed on this machine")
    out = subprocess.run([node, os.path.join(ROOT, "fluid-sim_amd", "node", "guess_smoke.mjs"), mode],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    js = json.loads(out.stdout.strip().splitlines()[-1])
    import fluidsim
    n = 512
    i = np.arange(n)[None, :]
    j = np.arange(n)[:, None]
    u = (((7 * i + 13 * j) % 31) / 31 - 0.5).reshape(-1)
    v = (((11 * i + 5 * j) % 29) / 29 - 0.5).reshape(-1)
    with fluidsim.GPUFluidSolver(0, n=n, nu=1e-4, advection="maccormack", pressureSolver="multigrid",
                                 pressureIterations=60, tolerance=1e-6, mgMinSize=4) as g:
        g.loadGuess(os.path.join(ROOT, "fluid-sim_amd", "weights", "guess_unet_256.fsgn"))
        g.setGuess(mode)
        g.uploadVelocity(u, v)
        for _ in range(5):
            g.stepCFL(0.5)
        vel = g.readVelocity()
        st = g.stats()
    k = np.arange(n * n)
    su = float(np.sum(vel["u"].astype(np.float64) * (k % 97 + 1)))
    sv = float(np.sum(vel["v"].astype(np.float64) * (k % 89 + 1)))
    assert st["last_residual"] <= 1e-6 and js["lastResidual"] <= 1e-6
    assert js["totalVCycles"] == st["total_vcycles"]
    assert math.isclose(js["checksumU"], su, rel_tol=1e-10) and math.isclose(js["checksumV"], sv, rel_tol=1e-10)


def _load(d, name, dtype):
    return np.fromfile(os.path.join(d, name), dtype=dtype)


def test_node_gputest_checks_equal_the_python_binding(tmp_path):
    # tools/gputest.html checks 1-7 run through solver.mjs alone (async readbacks,
    # readTexture of the named textures, runKernel, the Poisson curves, the spectrum);
    # the same calls through the Python binding on the inputs the JS engine generated
    # give the same bits
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed on this machine")
    d = str(tmp_path)
    out = subprocess.run([node, os.path.join(ROOT, "tests", "node", "gputest.mjs"), d], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    with open(os.path.join(d, "results.json")) as f:
        R = json.load(f)
    import fluidsim
    N, dt = 64, 0.01
    tu, tv = _load(d, "ic_tg_u.f64", np.float64), _load(d, "ic_tg_v.f64", np.float64)

    def same(a, b):
        a, b = np.asarray(a), np.asarray(b)
        return a.shape == b.shape and np.array_equal(a, b)

    with fluidsim.GPUFluidSolver(0, n=N, nu=0.0, pressureIterations=0, advection="semi-lagrangian") as g:
        g.uploadVelocity(tu, tv)
        g.runKernel("advect-sl", dt)
        vel = g.readVelocity()
    assert same(vel["u"], _load(d, "adv_u.f32", np.float32)) and same(vel["v"], _load(d, "adv_v.f32", np.float32))
    u0, v0 = _load(d, "ic_991_u.f64", np.float64), _load(d, "ic_991_v.f64", np.float64)
    with fluidsim.GPUFluidSolver(0, n=N, nu=0.0) as g:
        g.uploadVelocity(u0, v0)
        g.runKernel("divergence")
        assert same(g.readField("div"), _load(d, "div.f32", np.float32))
    assert R["readTextureRejectsHalfFloat"] is True
    u0, v0 = _load(d, "ic_4242_u.f64", np.float64), _load(d, "ic_4242_v.f64", np.float64)
    for name in ("jacobi", "red-black"):
        with fluidsim.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver=name, pressureIterations=200) as g:
            g.uploadVelocity(u0, v0)
            before = g.maxDivergence()
            g.step(0.01)
            after = g.maxDivergence()
        assert R["projection"][name] == {"before": before, "after": after}
    u0, v0 = _load(d, "ic_777_u.f64", np.float64), _load(d, "ic_777_v.f64", np.float64)
    for name in ("jacobi", "red-black", "multigrid"):
        budget = [0, 1, 2, 3, 4, 5, 6, 8, 10, 12] if name == "multigrid" else [0, 1, 2, 5, 10, 20, 40, 80, 160, 320]
        with fluidsim.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver=name) as g:
            g.uploadVelocity(u0, v0)
            assert R["curves"][name] == [g.poissonResidualAfter(it) for it in budget]
    with fluidsim.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver="red-black", pressureIterations=400,
                                 advection="semi-lagrangian") as g:
        g.uploadVelocity(tu, tv)
        for _ in range(10):
            g.step(dt)
        vel = g.readVelocity()
        assert R["fullStepMaxDiv"] == g.maxDivergence()
    assert same(vel["u"], _load(d, "full_u.f32", np.float32)) and same(vel["v"], _load(d, "full_v.f32", np.float32))
    with fluidsim.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver="multigrid", pressureIterations=4,
                                 advection="maccormack") as g:
        g.uploadVelocity(tu, tv)
        for _ in range(10):
            g.step(dt)
        vel = g.readVelocity()
    assert same(vel["u"], _load(d, "mc_u.f32", np.float32)) and same(vel["v"], _load(d, "mc_v.f32", np.float32))
    for scheme in ("semi-lagrangian", "maccormack"):
        with fluidsim.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver="multigrid", pressureIterations=3,
                                     advection=scheme) as g:
            g.uploadVelocity(tu, tv)
            for _ in range(200):
                g.step(0.01)
            vel = g.readVelocity()
        e = 0.5 * float(np.mean(vel["u"].astype(np.float64) ** 2 + vel["v"].astype(np.float64) ** 2))
        assert abs(R["schemeEnergy"][scheme] - e) <= 1e-12 * e
    assert R["schemeEnergy"]["maccormack"] > R["schemeEnergy"]["semi-lagrangian"]
    with fluidsim.GPUFluidSolver(0, n=N, nu=0.01, pressureSolver="multigrid", pressureIterations=20,
                                 tolerance=1e-6) as g:
        g.uploadVelocity(tu, tv)
        g.step(dt)
        assert same(g.readField("pressure"), _load(d, "tg_p.f64", np.float64))
        _, E, counts = g.energySpectrum()
    assert same(E, _load(d, "tg_E.f64", np.float64)) and R["spectrumModes"] == int(np.sum(counts))
    rgba = np.zeros(N * N * 4, np.float32)
    k = np.arange(N * N)
    rgba[0::4] = (k % 7).astype(np.float32) / np.float32(7)
    rgba[2::4] = 0.5
    splats = [dict(x=1.0, y=2.0, dx=3.0, dy=-1.0, radius=0.3),
              dict(x=6.1, y=0.1, r=0.9, g=0.2, b=0.5, radius=0.5, isDye=True)]
    with fluidsim.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver="multigrid", pressureIterations=3,
                                 dyeFade=0.02) as g:
        g.uploadVelocity(tu, tv)
        g.uploadDye(rgba)
        for _ in range(3):
            g.step(dt, splats=splats)
        assert same(g.readField("dye"), _load(d, "dye.f32", np.float32))
        vel = g.readVelocity()
        g.render(mode=0, scale=1.5, bloomThreshold=0.62, bloomPasses=2)
        img = g.readField("output")
        bloom = g.readField("bloom")
    assert same(vel["u"], _load(d, "splat_u.f32", np.float32)) and same(vel["v"], _load(d, "splat_v.f32", np.float32))
    assert R["renderOutputRejected"] is True
    assert same(img.view(np.uint16).reshape(-1), _load(d, "render.u16", np.uint16))
    assert same(bloom.view(np.uint16).reshape(-1), _load(d, "bloom.u16", np.uint16))
    with fluidsim.GPUFluidSolver(0, n=N, nu=0.0, pressureSolver="multigrid", pressureIterations=3,
                                 confinement=1.2) as g:
        g.uploadVelocity(tu, tv)
        for _ in range(3):
            g.step(dt)
        g.confinement = 2.5
        for _ in range(2):
            g.step(dt)
        vel = g.readVelocity()
    assert R["confinement"] == 2.5
    assert same(vel["u"], _load(d, "conf_u.f32", np.float32)) and same(vel["v"], _load(d, "conf_v.f32", np.float32))
