"""Dye and splats (the physics of the reference's display surface) on the GPU against
the f64 oracle's restatement of ADVECT_SCALAR and SPLAT (src/gpu/shaders.js:234-266,
:474-511), in the reference's step order: forcing, splats, projection, then the dye
through the projected velocity."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TWO_PI = 2 * math.pi


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return math.sqrt(np.sum((a - b) ** 2) / max(np.sum(b * b), 1e-300))


@pytest.fixture(scope="module")
def fs():
    import fluidsim
    fluidsim.createDevice()
    return fluidsim


SPLATS = [dict(x=1.0, y=2.0, dx=3.0, dy=-1.0, radius=0.3),
          dict(x=6.1, y=0.1, r=0.9, g=0.2, b=0.5, radius=0.5, isDye=True),      # across both seams
          dict(x=3.0, y=4.5, dx=-2.0, dy=0.5, radius=0.2),
          dict(x=3.2, y=4.4, r=0.1, g=0.8, b=0.3, radius=0.25, isDye=True)]


def test_splats_then_dye_match_oracle(fs, O):
    # from rest with no projection sweeps the step is: splats in order (each with its own
    # parameters), then the dye advected through the splatted velocity
    n, dt = 128, 0.01
    z = np.zeros(n * n)
    with fs.GPUFluidSolver(0, n=n, nu=0.0, pressureIterations=0) as g:
        g.uploadVelocity(z, z)
        g.step(dt, splats=SPLATS)
        vel = g.readVelocity()
        dye = g.readField("dye")
    u, v, d = z, z, np.zeros(4 * n * n)
    for sp in SPLATS:
        u, v, d = O.splat(n, n, u, v, d, sp)
    d2 = O.advect_dye(n, n, np.float32(u).astype(np.float64), np.float32(v).astype(np.float64), d, dt)
    assert max(rel(vel["u"], u), rel(vel["v"], v)) < 1e-6
    assert rel(dye, d2) < 1e-5
    assert np.all(dye >= 0) and np.all(dye <= 1)


def test_dye_rides_the_projected_flow_and_fades(fs, O):
    # Taylor-Green carrying an uploaded dye pattern for ten steps through the gated
    # multigrid projection (and its captured graphs); the oracle advects the same dye
    # through its own exactly projected velocity after each step
    n, dt, fade, steps = 128, 0.02, 0.03, 10
    h = TWO_PI / n
    u0, v0 = O.taylor_green(n, n)
    X, Y = np.meshgrid((np.arange(n) + 0.5) * h, (np.arange(n) + 0.5) * h)
    rgba = np.stack([0.5 + 0.4 * np.sin(X), 0.5 + 0.4 * np.cos(2 * Y), 0.9 * np.exp(-((X - 3) ** 2 + (Y - 3) ** 2)),
                     np.zeros_like(X)], axis=-1).reshape(-1)
    # the dye pass inside the captured step graph
    with fs.GPUFluidSolver(0, n=n, nu=0.01, advection="maccormack", pressureSolver="multigrid",
                           pressureIterations=40, tolerance=1e-9, mgMinSize=4, dyeFade=fade, graph=True) as g:
        g.uploadVelocity(u0, v0)
        g.uploadDye(rgba)
        for _ in range(steps):
            g.step(dt)
        st = g.stats()
        dye = g.readField("dye")
    assert st["gate_misses"] == 0 and st["graph_launches"] > 0
    s = O.Solver(n=n, nu=0.01, advection="maccormack", explicit_diffusion=True, projection="fft")
    s.u[:], s.v[:] = np.float32(u0), np.float32(v0)
    d = np.float32(rgba).astype(np.float64)
    for _ in range(steps):
        s.step(dt)
        d = O.advect_dye(n, n, s.u, s.v, d, dt, fade)
    assert rel(dye, d) < 1e-5
    with fs.GPUFluidSolver(0, n=16) as g:
        with pytest.raises(ValueError, match="no dye"):
            g.readField("dye")
