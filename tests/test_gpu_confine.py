"""Vorticity confinement (VORTICITY_CONFINEMENT, src/gpu/shaders.js:437-467) on the GPU:
the kernel against the f64 oracle (oracle/confine_oracle.py), its place in the step
(after the splats, before the projection, src/gpu/solver.js:558-568), the gated and
graph-replayed step with it, and the setter index.html:471 uses."""
import math

import numpy as np
import pytest

import confine_oracle as C

pytestmark = pytest.mark.gpu
TWO_PI = 2 * math.pi


@pytest.fixture(scope="module")
def fs():
    import fluidsim
    fluidsim.createDevice()
    return fluidsim


def smooth_field(nx, ny, seed):
    from fluidsim import initial
    if nx == ny:
        u, v = initial.decaying_turbulence(nx, seed=seed)
        return np.asarray(u, np.float64), np.asarray(v, np.float64)
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    u = sum(rng.standard_normal() * np.sin(TWO_PI * (k * y[:, None] + m * x[None, :]) + rng.uniform(0, 6))
            for k in range(1, 4) for m in range(0, 4))
    v = sum(rng.standard_normal() * np.cos(TWO_PI * (k * y[:, None] + m * x[None, :]) + rng.uniform(0, 6))
            for k in range(0, 4) for m in range(1, 4))
    return u.reshape(-1), v.reshape(-1)


@pytest.mark.parametrize("nx,ny,lx,ly", [(128, 128, TWO_PI, TWO_PI), (96, 64, TWO_PI, 1.5 * math.pi)])
def test_confinement_kernel_matches_oracle(fs, nx, ny, lx, ly):
    eps, dt = 2.0, 0.02
    u0, v0 = smooth_field(nx, ny, 11)
    with fs.GPUFluidSolver(0, nx=nx, ny=ny, lx=lx, ly=ly, confinement=eps) as g:
        g.uploadVelocity(u0, v0)
        vel0 = g.readVelocity()
        g.runKernel("confine", dt)
        vel = g.readVelocity()
    u = vel0["u"].reshape(ny, nx)
    v = vel0["v"].reshape(ny, nx)
    dx, dy = lx / nx, ly / ny
    ru, rv = C.confine(u, v, dx, dy, eps, dt)
    fu, fv = ru - u, rv - v
    gu = vel["u"].reshape(ny, nx).astype(np.float64) - u
    gv = vel["v"].reshape(ny, nx).astype(np.float64) - v
    # the unit direction n is ill-conditioned where grad|w| nearly vanishes: compare the
    # force where |grad|w|| is above 1e-3 of its maximum, and bound it everywhere
    w = C.corner_vorticity(u, v, dx, dy)
    a = np.abs(w)
    g2 = np.hypot((np.roll(a, -1, 1) - np.roll(a, 1, 1)) / (2 * dx), (np.roll(a, -1, 0) - np.roll(a, 1, 0)) / (2 * dy))
    ok = g2 > 1e-3 * g2.max()
    assert np.count_nonzero(ok) > 0.9 * ok.size
    scale = max(np.max(np.abs(fu)), np.max(np.abs(fv)))
    # FP32 arithmetic (the direction n from differences of |w|, up to ~3e-4 of the force
    # at the 1e-3 cut), and the rounding of u + f to FP32
    tol = 1e-3 * scale + 2 * np.spacing(np.float32(np.max(np.abs(u)) + np.max(np.abs(v))))
    assert np.max(np.abs(gu - fu)[ok]) <= tol and np.max(np.abs(gv - fv)[ok]) <= tol
    bound = np.float32(eps) * np.float32(dx) * np.float32(dt) * a * 1.0001 + tol
    assert np.all(np.hypot(gu, gv) <= bound)
    assert scale > 1e-4


def test_step_applies_confinement_after_advection_and_before_projection(fs):
    # with no diffusion, forcing or splats and no projection sweeps, step(dt) is the
    # advection followed by the confinement
    n, eps, dt = 128, 1.5, 0.01
    u0, v0 = smooth_field(n, n, 12)
    kw = dict(n=n, nu=0.0, advection="semi-lagrangian", pressureIterations=0, confinement=eps)
    with fs.GPUFluidSolver(0, **kw) as a, fs.GPUFluidSolver(0, **kw) as b:
        a.uploadVelocity(u0, v0)
        a.step(dt)
        va = a.readVelocity()
        b.uploadVelocity(u0, v0)
        b.runKernel("advect-sl", dt)
        b.runKernel("confine", dt)
        vb = b.readVelocity()
    assert np.array_equal(va["u"], vb["u"]) and np.array_equal(va["v"], vb["v"])


def test_gated_steps_with_confinement_meet_the_gate_and_graphs_replay_them(fs):
    n, steps = 256, 6
    u0, v0 = smooth_field(n, n, 13)
    out = {}
    for graph in (False, True):
        with fs.GPUFluidSolver(0, n=n, nu=1e-4, pressureSolver="multigrid", pressureIterations=30,
                               tolerance=1e-6, confinement=0.5, graph=graph, mgCoarseSolve="exact") as g:
            g.uploadVelocity(u0, v0)
            for k in range(steps):
                if k == 3:
                    g.confinement = 2.0  # index.html:471; drops bodies captured with 0.5
                g.stepCFL(0.5)
            vel = g.readVelocity()
            out[graph] = (vel, g.stats())
    (va, sa), (vb, sb) = out[False], out[True]
    assert sa["gate_misses"] == 0 and sb["gate_misses"] == 0
    assert sa["last_residual"] <= 1e-6 and sb["last_residual"] <= 1e-6
    assert sb["graph_launches"] > 0
    assert np.array_equal(va["u"], vb["u"]) and np.array_equal(va["v"], vb["v"])
    with fs.GPUFluidSolver(0, n=n, nu=1e-4, pressureSolver="multigrid", pressureIterations=30,
                           tolerance=1e-6, mgCoarseSolve="exact") as g:
        g.uploadVelocity(u0, v0)
        for _ in range(steps):
            g.stepCFL(0.5)
        plain = g.readVelocity()
    assert not np.array_equal(plain["u"], va["u"])  # the force acted


def test_confinement_off_is_not_dispatched(fs):
    with fs.GPUFluidSolver(0, n=64) as g:
        with pytest.raises(Exception):
            g.runKernel("confine", 0.01)
        g.confinement = -1.0  # the reference's `> 0` test: off
        with pytest.raises(Exception):
            g.runKernel("confine", 0.01)
