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
