"""GPU parity: the HIP path through the C ABI against the f64 oracle on the same
seeded inputs, restating tools/gputest.html checks 1-7 plus the tolerance-gated
projection. FP32 kernels are compared with the tolerance written in each test;
the FP64 refinement residual is checked against its gate exactly.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 64
TWO_PI = 2 * math.pi


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return math.sqrt(np.sum((a - b) ** 2) / max(np.sum(b * b), 1e-300))


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def fs():
    import fluidsim
    fluidsim.createDevice()
    return fluidsim


def solver(fs, **kw):
    return fs.GPUFluidSolver(**kw)


# ------------------------------------------------------------- single kernels

def test_advection_sl_kernel_matches_oracle(fs, O):
    # gputest check 1: one SL step of TG at 64^2, dt = 0.01 (reference relL2 2.3e-7)
    u0, v0 = O.taylor_green(N, N)
    with solver(fs, n=N, advection="semi-lagrangian", pressureIterations=0) as g:
        g.uploadVelocity(u0, v0)
        g.runKernel("advect-sl", 0.01)
        out = g.readVelocity()
    ou, ov = O.advect_velocity("semi-lagrangian", N, N, f32(u0), f32(v0), 0.01)
    assert max(rel(out["u"], ou), rel(out["v"], ov)) < 1e-6


@pytest.mark.parametrize("seed", [991, 31337])
def test_advection_mc_kernel_matches_oracle_on_rough_field(fs, O, seed):
    # random field: the limiter is active at most faces
    u0, v0 = O.js_random_velocity(N, seed)
    dt = 0.05
    with solver(fs, n=N, advection="maccormack") as g:
        g.uploadVelocity(u0, v0)
        g.runKernel("advect-mc", dt)
        out = g.readVelocity()
    ou, ov = O.advect_velocity("maccormack", N, N, f32(u0), f32(v0), dt)
    assert max(rel(out["u"], ou), rel(out["v"], ov)) < 1e-5


def test_divergence_kernel_matches_oracle(fs, O):
    # gputest check 2 (seed 991; reference relL2 5.8e-8)
    u0, v0 = O.js_random_velocity(N, 991)
    with solver(fs, n=N) as g:
        g.uploadVelocity(u0, v0)
        g.runKernel("divergence")
        d = g.readField("div")
    ref = O.divergence(N, N, f32(u0), f32(v0))
    assert rel(d, ref) < 1e-6


def test_diffusion_kernel_matches_oracle(fs, O):
    u0, v0 = O.js_random_velocity(N, 4242)
    nu, dt = 0.01, 0.02
    with solver(fs, n=N, nu=nu) as g:
        g.uploadVelocity(u0, v0)
        g.runKernel("diffuse", dt)
        out = g.readVelocity()
    ou = O.diffuse_explicit(N, N, f32(u0), nu, dt)
    ov = O.diffuse_explicit(N, N, f32(v0), nu, dt)
    assert max(rel(out["u"], ou), rel(out["v"], ov)) < 1e-6


def test_max_divergence_reduction(fs, O, golden):
    u0, v0 = O.js_random_velocity(N, 4242)
    with solver(fs, n=N) as g:
        g.uploadVelocity(u0, v0)
        md = g.maxDivergence()
        d = g.diagnostics()
    ref = O.max_divergence(N, N, f32(u0), f32(v0))
    assert abs(md - ref) <= 1e-12 * ref
    assert abs(md - golden["projection"]["red-black"]["before"]) / md < 2e-7
    assert abs(d["kinetic_energy"] - O.kinetic_energy(N, N, f32(u0), f32(v0))) < 1e-12
    assert abs(d["momentum_x"] - f32(u0).sum() * (TWO_PI / N) ** 2) < 1e-9
    ens = O.enstrophy(N, N, f32(u0), f32(v0))
    assert abs(d["enstrophy"] - ens) <= 1e-12 * ens


@pytest.mark.parametrize("n", [64, 1024])
def test_enstrophy_reduction_matches_oracle(fs, O, n):
    # src/core/grid.js:297-303 on the device in FP64: Taylor-Green (its closed form is
    # pinned on the oracle) and decaying turbulence
    from fluidsim import initial
    for u0, v0 in (O.taylor_green(n, n), initial.decaying_turbulence(n, seed=1)):
        with solver(fs, n=n) as g:
            g.uploadVelocity(u0, v0)
            d = g.diagnostics()
        ens = O.enstrophy(n, n, f32(u0), f32(v0))
        assert abs(d["enstrophy"] - ens) <= 1e-11 * ens


# ---------------------------------------------------- projection and solvers

@pytest.mark.parametrize("solver_name", ["jacobi", "red-black"])
def test_projection_after_200_sweeps(fs, O, golden, solver_name):
    # gputest check 3 (seed 4242, MacCormack default, 200 sweeps)
    u0, v0 = O.js_random_velocity(N, 4242)
    with solver(fs, n=N, pressureSolver=solver_name, pressureIterations=200) as g:
        g.uploadVelocity(u0, v0)
        before = g.maxDivergence()
        g.step(0.01)
        after = g.maxDivergence()
    s = O.Solver(n=N, nu=0.0, advection="maccormack", projection=solver_name, iterations=200)
    s.u[:], s.v[:] = f32(u0), f32(v0)
    s.step(0.01)
    ref = O.max_divergence(N, N, s.u, s.v)
    assert abs(after - ref) / ref < 1e-3
    assert abs(after - golden["projection"][solver_name]["after"]) / ref < 1e-3
    assert before / after > (1000 if solver_name == "red-black" else 25)


@pytest.mark.parametrize("solver_name", ["jacobi", "red-black", "multigrid"])
def test_poisson_curves(fs, O, golden, solver_name):
    # gputest check 4 (seed 777): residual after k iterations from a zeroed pressure
    u0, v0 = O.js_random_velocity(N, 777)
    d = O.divergence(N, N, f32(u0), f32(v0))
    with solver(fs, n=N, pressureSolver=solver_name) as g:
        g.uploadVelocity(u0, v0)
        for it, ref in golden["poissonCurves"][solver_name]:
            r = g.poissonResidualAfter(it)
            if it == 0:
                assert abs(r - 1.0) < 1e-12
                continue
            if solver_name == "jacobi":
                p = O.jacobi(N, N, d, it)
            elif solver_name == "red-black":
                p = O.redblack(N, N, d, it)
            else:
                p, _ = O.mg_vcycles(N, d, it)
            ro = O.poisson_residual_rel(N, N, p, d)
            if ro > 1e-5:
                assert abs(r - ro) / ro < 1e-3, (it, r, ro)
                assert abs(r - ref) / ref < 1e-3, (it, r, ref)
            else:
                # below the FP32 floor of the reference's plateau (~9e-8)
                assert r < 2e-6, (it, r)


# ------------------------------------------------------------- full steps

@pytest.mark.parametrize("scheme,solver_name,iters,steps,tol", [
    ("semi-lagrangian", "red-black", 400, 10, 1e-5),   # gputest check 5 (ref 3.3e-6 vs FFT)
    ("maccormack", "multigrid", 4, 10, 1e-5),          # gputest check 6 (ref 3.5e-7 vs FFT)
])
def test_full_steps_match_oracle(fs, O, scheme, solver_name, iters, steps, tol):
    u0, v0 = O.taylor_green(N, N)
    with solver(fs, n=N, advection=scheme, pressureSolver=solver_name, pressureIterations=iters) as g:
        g.uploadVelocity(u0, v0)
        for _ in range(steps):
            g.step(0.01)
        out = g.readVelocity()
    s = O.Solver(n=N, nu=0.0, advection=scheme, projection=solver_name, iterations=iters)
    s.u[:], s.v[:] = f32(u0), f32(v0)
    for _ in range(steps):
        s.step(0.01)
    assert max(rel(out["u"], s.u), rel(out["v"], s.v)) < tol
    sf = O.Solver(n=N, nu=0.0, advection=scheme, projection="fft")
    sf.u[:], sf.v[:] = f32(u0), f32(v0)
    for _ in range(steps):
        sf.step(0.01)
    assert max(rel(out["u"], sf.u), rel(out["v"], sf.v)) < 5e-3


@pytest.mark.parametrize("scheme", ["semi-lagrangian", "maccormack"])
def test_scheme_energy_after_200_steps(fs, O, golden, scheme):
    # gputest check 7: MacCormack keeps more energy than SL at nu = 0
    u0, v0 = O.taylor_green(N, N)
    with solver(fs, n=N, advection=scheme, pressureSolver="multigrid", pressureIterations=3) as g:
        g.uploadVelocity(u0, v0)
        for _ in range(200):
            g.step(0.01)
        out = g.readVelocity()
    ke = 0.5 * np.mean(out["u"].astype(np.float64) ** 2 + out["v"].astype(np.float64) ** 2)
    assert abs(ke - golden["schemeEnergy"][scheme]) / ke < 2e-5


def test_taylor_green_200_steps_fp32_against_f64(fs, O):
    # Taylor-Green 64^2, nu = 0.01, MacCormack + explicit diffusion, projection to
    # an FP64 residual of 1e-9, dt at CFL 0.5; FP32 GPU against the f64 oracle (exact FFT)
    n, nu = 64, 0.01
    u0, v0 = O.taylor_green(n, n)
    h = TWO_PI / n
    dt = 0.5 / (np.max(np.abs(u0)) / h + np.max(np.abs(v0)) / h)
    with solver(fs, n=n, nu=nu, advection="maccormack", pressureSolver="multigrid",
                pressureIterations=30, tolerance=1e-9) as g:
        g.uploadVelocity(u0, v0)
        for _ in range(200):
            g.step(dt)
            assert g.stats()["last_residual"] <= 1e-9
        out = g.readVelocity()
    s = O.Solver(n=n, nu=nu, advection="maccormack", explicit_diffusion=True, projection="fft")
    s.u[:], s.v[:] = f32(u0), f32(v0)
    for _ in range(200):
        s.step(dt)
    err = max(rel(out["u"], s.u), rel(out["v"], s.v))
    assert err < 1e-4, err
    T = 200 * dt
    ke0 = O.kinetic_energy(n, n, f32(u0), f32(v0))
    ke = O.kinetic_energy(n, n, out["u"].astype(np.float64), out["v"].astype(np.float64))
    rate_gpu = -math.log(ke / ke0) / T
    ke_o = O.kinetic_energy(n, n, s.u, s.v)
    rate_o = -math.log(ke_o / ke0) / T
    assert abs(rate_gpu - rate_o) / rate_o < 1e-3
    assert 4 * nu < rate_gpu < 1.35 * 4 * nu


def test_tolerance_gate_on_turbulence(fs, O):
    # decaying turbulence at 256^2: every projection meets max|div u* - L p| <= 1e-6 in FP64, and the
    # velocity tracks the oracle's exact projection
    from fluidsim import initial
    n, nu = 256, 1e-4
    u0, v0 = initial.decaying_turbulence(n, seed=1)
    with solver(fs, n=n, nu=nu, advection="maccormack", pressureSolver="multigrid",
                pressureIterations=40, tolerance=1e-6, mgMinSize=4) as g:
        g.uploadVelocity(u0, v0)
        dts = []
        for _ in range(5):
            g.stepCFL(0.5)
            st = g.stats()
            dts.append(st["last_dt"])
            assert st["last_residual"] <= 1e-6
            assert st["last_vcycles"] >= 1
        out = g.readVelocity()
        d = g.diagnostics()
    s = O.Solver(n=n, nu=nu, advection="maccormack", explicit_diffusion=True, projection="fft")
    s.u[:], s.v[:] = f32(u0), f32(v0)
    h = TWO_PI / n
    for k in range(5):
        dt = 0.5 / (np.max(np.abs(s.u)) / h + np.max(np.abs(s.v)) / h)
        assert abs(dt - dts[k]) / dt < 1e-5
        s.step(dts[k])
    assert max(rel(out["u"], s.u), rel(out["v"], s.v)) < 1e-5
    # the stored FP32 field's own divergence sits at the FP32 rounding floor
    assert d["max_divergence"] < 5e-4


MODES = [dict(kx=3, ky=-2, phase=0.4), dict(kx=0, ky=5, phase=2.0), dict(kx=0, ky=0, phase=1.0),
         dict(kx=-4, ky=1, phase=5.5), dict(kx=6, ky=4, phase=3.3)]


@pytest.mark.parametrize("n", [128, 96])
def test_mode_list_forcing_kernel_matches_oracle(fs, O, n):
    # setForcing(modes, {amplitude, drag, phaseDrift}) (src/gpu/solver.js:370-414) applied
    # once at t = 0 against the oracle's FORCING (src/gpu/shaders.js:532-566); 128: the
    # field as a product on the matrix cores, 96 (not a multiple of 64): the VALU sum
    dt = 0.02
    from fluidsim import initial
    u0, v0 = initial.decaying_turbulence(n, seed=2)
    with solver(fs, n=n) as g:
        g.uploadVelocity(u0, v0)
        g.setForcing(MODES, amplitude=1.9, drag=0.42, phaseDrift=0.7)
        g.runKernel("forcing", dt)
        out = g.readVelocity()
    ou, ov = O.mode_forcing(n, n, f32(u0), f32(v0), MODES, 1.9, 0.42, 0.7, 0.0, dt)
    assert max(rel(out["u"], ou), rel(out["v"], ov)) < 1e-6


@pytest.mark.parametrize("graph", [False, True])
def test_mode_list_forced_steps_match_oracle(fs, O, graph):
    # twenty steps whose forcing phases drift with the simulated time, through the gated
    # projection (and its captured graphs), against the oracle with the exact projection
    n, nu, dt, steps = 128, 0.01, 0.01, 20
    u0, v0 = O.taylor_green(n, n)
    with solver(fs, n=n, nu=nu, advection="maccormack", pressureSolver="multigrid", pressureIterations=40,
                tolerance=1e-9, mgMinSize=4, graph=graph) as g:
        g.setForcing(MODES, amplitude=1.9, drag=0.42, phaseDrift=0.7)
        g.uploadVelocity(u0, v0)
        for _ in range(steps):
            g.step(dt)
        st = g.stats()
        assert st["last_residual"] <= 1e-9 and st["gate_misses"] == 0
        assert abs(st["t"] - steps * dt) < 1e-12
        out = g.readVelocity()
    s = O.Solver(n=n, nu=nu, advection="maccormack", explicit_diffusion=True, projection="fft")
    s.set_modes(MODES, amplitude=1.9, drag=0.42, phaseDrift=0.7)
    s.u[:], s.v[:] = f32(u0), f32(v0)
    for _ in range(steps):
        s.step(dt)
    assert max(rel(out["u"], s.u), rel(out["v"], s.v)) < 1e-5


@pytest.mark.parametrize("n,kf,steps", [(64, 8, 20), (256, 16, 5)])
def test_forced_steps_match_oracle(fs, O, n, kf, steps):
    # band forcing + linear drag (src/cpu/solver.js:179-237, order :354-358): step k
    # uses LCG seed 1000 + k; FP32 GPU with an FP64-gated projection against the f64
    # oracle with the exact FFT projection
    nu, dt = 0.01, 0.01
    forcing = dict(kf=kf, width=2.0, amp=0.5, drag=0.05, seed=1000)
    u0, v0 = O.taylor_green(n, n)
    with solver(fs, n=n, nu=nu, advection="maccormack", pressureSolver="multigrid",
                pressureIterations=40, tolerance=1e-9, mgMinSize=4) as g:
        g.setForcing(forcing)
        g.uploadVelocity(u0, v0)
        for _ in range(steps):
            g.step(dt)
            assert g.stats()["last_residual"] <= 1e-9
        out = g.readVelocity()
    s = O.Solver(n=n, nu=nu, advection="maccormack", explicit_diffusion=True, projection="fft",
                 forcing=dict(kf=kf, width=2.0, amp=0.5), drag=0.05)
    s.u[:], s.v[:] = f32(u0), f32(v0)
    for k in range(steps):
        s.step(dt, force_seed=1000 + k)
    err = max(rel(out["u"], s.u), rel(out["v"], s.v))
    assert err < 2e-5, err
    # not vacuous: the unforced run is far from the forced one
    s0 = O.Solver(n=n, nu=nu, advection="maccormack", explicit_diffusion=True, projection="fft")
    s0.u[:], s0.v[:] = f32(u0), f32(v0)
    for _ in range(steps):
        s0.step(dt)
    assert rel(s0.u, s.u) > 100 * err


def test_runs_are_bitwise_deterministic(fs, O):
    # test/solid.test.js:339-358 style: two identical runs, identical bits
    u0, v0 = O.js_random_velocity(N, 31337)
    outs = []
    for _ in range(2):
        with solver(fs, n=N, nu=1e-3, advection="maccormack", pressureSolver="multigrid",
                    pressureIterations=20, tolerance=1e-7) as g:
            g.uploadVelocity(u0 * 0.1, v0 * 0.1)
            for _ in range(5):
                g.step(0.01)
            outs.append(g.readVelocity())
    assert (outs[0]["u"] == outs[1]["u"]).all() and (outs[0]["v"] == outs[1]["v"]).all()


def test_nonfinite_velocity_is_reported(fs):
    import fluidsim
    u = np.zeros(N * N)
    v = np.zeros(N * N)
    u[123] = np.nan
    with solver(fs, n=N) as g:
        g.uploadVelocity(u, v)
        with pytest.raises(fluidsim.FluidSimError, match="non-finite"):
            g.diagnostics()


# ------------------------------------------- fused V-cycle == per-operator V-cycle

@pytest.mark.parametrize("n,pre,post,min_size", [(64, 2, 2, 8), (256, 2, 2, 4), (256, 1, 3, 8),
                                                 (512, 3, 1, 4), (512, 1, 3, 4), (192, 2, 2, 4), (256, 2, 2, 2),
                                                 (256, 2, 2, 16), (1024, 2, 2, 4)])
def test_fused_vcycle_bitwise_equals_per_operator_path(fs, n, pre, post, min_size):
    from fluidsim import initial
    u0, v0 = initial.decaying_turbulence(n, seed=3)
    outs = []
    for fused in (True, False):
        with solver(fs, n=n, nu=1e-4, advection="maccormack", pressureSolver="multigrid", pressureIterations=3,
                    mgMinSize=min_size, mgPre=pre, mgPost=post, mgFused=fused) as g:
            g.uploadVelocity(u0, v0)
            for _ in range(3):
                g.step(0.004)
            outs.append((g.readVelocity(), g.readField("pressure"), g.stats()["last_residual"]))
    (a, pa, ra), (b, pb, rb) = outs
    assert (a["u"] == b["u"]).all() and (a["v"] == b["v"]).all()
    assert (pa == pb).all()
    assert ra == rb


def test_fused_tolerance_mode_matches_per_operator_counts(fs):
    from fluidsim import initial
    n = 256
    u0, v0 = initial.decaying_turbulence(n, seed=5)
    res = []
    for fused in (True, False):
        with solver(fs, n=n, nu=1e-4, pressureSolver="multigrid", pressureIterations=40, tolerance=1e-6,
                    mgMinSize=4, mgFused=fused) as g:
            g.uploadVelocity(u0, v0)
            cyc = []
            for _ in range(4):
                g.stepCFL(0.5)
                cyc.append(g.stats()["last_vcycles"])
            res.append((cyc, g.readVelocity()))
    assert res[0][0] == res[1][0]
    assert (res[0][1]["u"] == res[1][1]["u"]).all()


@pytest.mark.parametrize("n,min_size", [(96, 4), (192, 4), (256, 4), (128, 2), (256, 8), (100, 4)])
def test_vcycle_residual_curve_matches_oracle(fs, O, n, min_size):
    # non-power-of-two levels (96 -> 6, 192 -> 6, 100 stops at 50), the direct coarsest
    # solve (6x6, 4x4, 2x2), and 8x8 swept: residual after k V-cycles vs the f64 oracle
    u0, v0 = O.js_random_velocity(n, 777)
    d = O.divergence(n, n, f32(u0), f32(v0))
    with solver(fs, n=n, pressureSolver="multigrid", mgMinSize=min_size) as g:
        g.uploadVelocity(u0, v0)
        for it in (1, 2, 3):
            r = g.poissonResidualAfter(it)
            _, hist = O.mg_vcycles(n, d, it, min_size=min_size)
            ro = hist[-1]
            assert abs(r - ro) / ro < 2e-3, (n, it, r, ro)


@pytest.mark.parametrize("adv_fused", [True, False])
def test_profile_counts_every_launch_and_its_bytes(fs, O, adv_fused):
    # one fused V(2,2) step at 256^2, minSize 4: levels 256..4, tile kernels on 256
    # and 128, the one-workgroup kernel from 64 down; a fixed 3 V-cycles
    n = 256
    u0, v0 = O.js_random_velocity(n, 7)
    with solver(fs, n=n, advection="maccormack", pressureSolver="multigrid", pressureIterations=3,
                mgMinSize=4, advFused=adv_fused) as g:
        g.uploadVelocity(u0, v0)
        g.profileEnable()
        g.step(1e-3)
        g.sync()
        p = g.profileRead()
        g.profileEnable([])
        g.step(1e-3)
        assert g.profileRead()["advect_sl"][1] == 0 and g.profileRead()["advect_fused"][1] == 0
    if adv_fused:
        ms, launches, nbytes = p["advect_fused"]
        assert launches == 1 and ms > 0 and nbytes == 16 * n * n
        assert p["advect_sl"][1] == 0 and p["advect_mc"][1] == 0
    else:
        ms, launches, nbytes = p["advect_sl"]
        assert launches == 1 and ms > 0 and nbytes == 16 * n * n
        assert p["advect_mc"][1] == 1 and p["advect_mc"][2] == 24 * n * n
        assert p["advect_fused"][1] == 0
    assert p["gradient"][1] == 1 and p["gradient"][2] == 24 * n * n
    # three V-cycles: one level-0 down-sweep each, plus the final residual-only pass;
    # levels 1 and 0 come back up in one launch
    assert p["down0"][1] + p["residual"][1] == 4
    assert p["up"][1] == 3 and p["down"][1] == 3 and p["coarse"][1] == 3
    for k, (ms_k, l_k, b_k) in p.items():
        assert (ms_k > 0) == (l_k > 0) and (b_k > 0) == (l_k > 0), k


# ------------------------------------------------------------- immersed disk

@pytest.mark.parametrize("passes", [1, 2])
def test_solid_kernel_matches_oracle(fs, O, passes):
    # one coupling pass over a rough field, moving and rotating disk across the seam
    from fluidsim.solid import RigidDisk
    n = 64
    u0, v0 = O.js_random_velocity(n, 4242)
    disk = RigidDisk(x=0.3, y=3.0, r=0.9, vx=-0.4, vy=0.25, omega=1.3)
    with solver(fs, n=n) as g:
        g.uploadVelocity(u0, v0)
        g.setSolid(disk, passes=passes)
        g.runKernel("solid", 0.0)
        out = g.readVelocity()
        imp = g.solidImpulse()
    u, v = f32(u0), f32(v0)
    ref = O.apply_coupling(n, n, u, v, disk.as_dict(), passes=passes)
    assert np.max(np.abs(out["u"] - u)) < 4e-7 and np.max(np.abs(out["v"] - v)) < 4e-7
    for a, b in zip(imp, ref):
        assert abs(a - b) <= 1e-9 * max(abs(b), 1e-3), (imp, ref)


def test_pinned_disk_steps_match_oracle(fs, O):
    # fixed disk in a Taylor-Green flow: advect, diffuse, couple (2 passes), project
    from fluidsim.solid import RigidDisk
    n, nu, dt, steps = 64, 0.01, 0.01, 20
    u0, v0 = O.taylor_green(n, n)
    disk = RigidDisk(x=2.0, y=3.5, r=0.6, fixed=True)
    with solver(fs, n=n, nu=nu, advection="maccormack", pressureSolver="multigrid",
                pressureIterations=40, tolerance=1e-9, mgMinSize=4) as g:
        g.uploadVelocity(u0, v0)
        g.setSolid(disk, passes=2)
        for _ in range(steps):
            g.step(dt)
            g.integrateSolid(dt)
        out = g.readVelocity()
        d = g.diagnostics()
    assert disk.x == 2.0 and disk.vx == 0.0
    s = O.Solver(n=n, nu=nu, advection="maccormack", explicit_diffusion=True, projection="fft",
                 disk=disk.as_dict(), passes=2)
    s.u[:], s.v[:] = f32(u0), f32(v0)
    for _ in range(steps):
        s.step(dt)
    err = max(rel(out["u"], s.u), rel(out["v"], s.v))
    assert err < 2e-5, err
    s0 = O.Solver(n=n, nu=nu, advection="maccormack", explicit_diffusion=True, projection="fft")
    s0.u[:], s0.v[:] = f32(u0), f32(v0)
    for _ in range(steps):
        s0.step(dt)
    assert rel(s0.u, s.u) > 100 * err
    assert d["max_divergence"] < 1e-4


def test_free_disk_exchanges_momentum_with_the_fluid(fs, O):
    # Newton's third law through the device reduction: the fluid's momentum change
    # over a coupling pass equals the reported impulse, and the disk takes the opposite
    from fluidsim.solid import RigidDisk
    n, dt = 128, 0.01
    u0 = np.full(n * n, 1.0)
    v0 = np.full(n * n, -0.5)
    disk = RigidDisk(x=3.0, y=3.0, r=0.7, density=2.5, vx=-0.3, vy=0.2)
    with solver(fs, n=n) as g:
        g.uploadVelocity(u0, v0)
        p0 = g.diagnostics()
        g.setSolid(disk)
        g.runKernel("solid", dt)
        p1 = g.diagnostics()
        imp = g.solidImpulse()
        m0 = disk.momentum()
        disk.apply_reaction(imp, dt)
        disk.integrate(dt)
        m1 = disk.momentum()
    dpx = p1["momentum_x"] - p0["momentum_x"]
    dpy = p1["momentum_y"] - p0["momentum_y"]
    # the stored field is FP32: its momentum change carries the rounding of each face
    assert abs(dpx - imp[0]) < 1e-5 * abs(imp[0]) and abs(dpy - imp[1]) < 1e-5 * abs(imp[1])
    assert abs((m1[0] - m0[0]) + imp[0]) < 1e-12 and abs((m1[1] - m0[1]) + imp[1]) < 1e-12


@pytest.mark.parametrize("n", [256, 512])
def test_autotuned_tiles_give_the_same_bits(fs, O, n):
    # the tile shape changes only the schedule: autotuned and fixed-rule solvers agree
    # bit for bit, and so do both with the per-operator V-cycle
    u0, v0 = O.js_random_velocity(n, 11)
    outs = []
    for kw in (dict(autotune=True), dict(autotune=False), dict(mgFused=False)):
        with solver(fs, n=n, advection="maccormack", pressureSolver="multigrid", pressureIterations=3,
                    mgMinSize=4, **kw) as g:
            g.uploadVelocity(u0, v0)
            for _ in range(2):
                g.step(2e-3)
            outs.append(g.readVelocity())
            if kw.get("autotune"):
                rep = g.tileReport()
                assert "level 0" in rep and "down0" in rep and "us" in rep
    for o in outs[1:]:
        assert np.array_equal(outs[0]["u"], o["u"]) and np.array_equal(outs[0]["v"], o["v"])


@pytest.mark.parametrize("n,prec", [(256, "fp32"), (1024, "fp32"), (512, "bf16"), (192, "fp32")])
def test_every_forced_tile_shape_gives_the_same_bits(fs, O, n, prec):
    # each shape forced on every grid-wide level (64 x 64 also for levels 1 and 0 going up
    # together when the grid is whole 64 x 64 tiles; 192 is not, and takes 32 x 32 there)
    u0, v0 = O.js_random_velocity(n, 13)
    outs = []
    for tile in ("32x32", "64x32", "16x16", "8x8", "64x64", "32x32/1024", "16x16/256"):
        print(f"forced tile {tile} n={n} {prec}", flush=True)
        with solver(fs, n=n, advection="maccormack", pressureSolver="multigrid", pressureIterations=3,
                    mgMinSize=4, mgPrecision=prec, tile=tile, mgCoarseSolve="exact") as g:
            g.uploadVelocity(u0, v0)
            for _ in range(2):
                g.step(2e-3)
            outs.append((g.readVelocity(), g.readField("pressure")))
            assert f": {tile}" in g.tileReport()
    for o, p in outs[1:]:
        assert np.array_equal(outs[0][0]["u"], o["u"]) and np.array_equal(outs[0][0]["v"], o["v"])
        assert np.array_equal(outs[0][1], p)


# ------------------------------------------------------ BF16 multigrid storage

@pytest.mark.parametrize("n", [128, 256, 512])  # 512: the two-level down/up launches run
def test_bf16_storage_residual_curve_matches_oracle(fs, O, n):
    # BF16-stored correction and right-hand sides, FP32 arithmetic, FP64 refinement
    # residual: the residual after k V-cycles against the oracle's emulation of the
    # same storage points; rounding decisions of FP32 vs f64 arithmetic may differ
    # at a few cells, so the curves agree to a few percent, not bit for bit
    u0, v0 = O.js_random_velocity(n, 777)
    d = O.divergence(n, n, f32(u0), f32(v0))
    with solver(fs, n=n, pressureSolver="multigrid", mgMinSize=4, mgPrecision="bf16") as g:
        g.uploadVelocity(u0, v0)
        for it in (1, 2, 3, 4):
            r = g.poissonResidualAfter(it)
            p, _ = O.mg_refine(n, d, it, min_size=4, storage="bf16")
            ro = O.poisson_residual_rel(n, n, p, d)
            p32, _ = O.mg_refine(n, d, it, min_size=4, storage="fp32")
            r32 = O.poisson_residual_rel(n, n, p32, d)
            assert abs(r - ro) / ro < 0.05, (n, it, r, ro, r32)
            # and the storage matters: after the first V-cycle the BF16 curve is a few
            # percent off the FP32 one (the corrections passed up stay FP32)
            if it == 1:
                assert abs(ro - r32) / r32 > 1e-2


def test_bf16_storage_meets_the_tolerance_gate(fs, O):
    # decaying turbulence at 256^2 with BF16 smoother storage: every projection meets
    # the FP64 gate, within a V-cycle or so of the FP32-storage solver
    from fluidsim import initial
    n = 256
    u0, v0 = initial.decaying_turbulence(n, seed=1)
    cyc = {}
    for prec in ("fp32", "bf16"):
        with solver(fs, n=n, nu=1e-4, advection="maccormack", pressureSolver="multigrid",
                    pressureIterations=40, tolerance=1e-6, mgMinSize=4, mgPrecision=prec) as g:
            g.uploadVelocity(u0, v0)
            total = 0
            for _ in range(5):
                g.stepCFL(0.5)
                st = g.stats()
                assert st["last_residual"] <= 1e-6
                total += st["last_vcycles"]
            cyc[prec] = total
            out = g.readVelocity()
    assert cyc["bf16"] <= cyc["fp32"] + 5
    s = O.Solver(n=n, nu=1e-4, advection="maccormack", explicit_diffusion=True, projection="fft")
    s.u[:], s.v[:] = f32(u0), f32(v0)
    h = TWO_PI / n
    for _ in range(5):
        s.step(0.5 / (np.max(np.abs(s.u)) / h + np.max(np.abs(s.v)) / h))
    assert max(rel(out["u"], s.u), rel(out["v"], s.v)) < 1e-5


# ------------------------------------------------------------ learned guess

def _random_guess_blob(seed):
    import torch
    from guess_torch import build_model, export_blob
    torch.manual_seed(seed)
    return export_blob(build_model())


@pytest.mark.parametrize("seed", [0, 1])
def test_learned_guess_network_matches_torch(fs, O, seed):
    # the HIP U-Net (BF16 MFMA convolutions, FP32 accumulation) against the same
    # network in float64 torch with the library's BF16 roundings of weights and
    # activations; zero iterations leave the guess itself in the pressure
    from fluidsim import initial
    from guess_torch import reference_guess
    n = 256
    blob = _random_guess_blob(seed)
    u0, v0 = initial.decaying_turbulence(n, seed=3)
    with solver(fs, n=n, pressureSolver="multigrid", pressureIterations=0, mgMinSize=4) as g:
        g.uploadVelocity(u0, v0)
        g.loadGuess(blob)
        g.setGuess("learned")
        g.runKernel("project", 0.0)
        p = g.readField("pressure")
    r = O.divergence(n, n, f32(u0), f32(v0))
    ref = reference_guess(blob, r, np.zeros(n * n), n, TWO_PI / n)
    assert np.max(np.abs(ref)) > 0
    err = rel(p, ref)
    assert err < 3e-2, err


def _restrict(a, n, k):
    f = 1 << k
    a = np.asarray(a, np.float64).reshape(n // f, f, n // f, f)
    return a.mean(axis=(1, 3)).reshape(-1)


def _prolong(dc, n, k):
    # cell-centred bilinear, periodic: fine centre i + 1/2 at coarse coordinate
    # (i + 1/2) / 2^k - 1/2
    m = n >> k
    c = np.asarray(dc, np.float64).reshape(m, m)
    g = (np.arange(n) + 0.5) / (1 << k) - 0.5
    i0 = np.floor(g).astype(int)
    f = g - i0
    i0 %= m
    i1 = (i0 + 1) % m
    rows = c[i0, :] * (1 - f)[:, None] + c[i1, :] * f[:, None]
    return (rows[:, i0] * (1 - f)[None, :] + rows[:, i1] * f[None, :]).reshape(-1)


@pytest.mark.parametrize("k", [1, 2])
def test_prolong_restatement_is_bilinear(k):
    # pins the test's own interpolation: a constant stays constant, a linear ramp is
    # reproduced exactly away from the periodic seam, and for k = 1 the weights are
    # the multigrid's 3/4, 1/4
    n = 32
    m = n >> k
    assert np.allclose(_prolong(np.full(m * m, 3.0), n, k), 3.0)
    x = np.arange(m, dtype=np.float64)
    fine = _prolong(np.tile(x, (m, 1)).reshape(-1), n, k).reshape(n, n)
    xf = (np.arange(n) + 0.5) / (1 << k) - 0.5
    inner = (xf >= 0) & (xf <= m - 1)
    assert np.allclose(fine[:, inner], np.tile(xf[inner], (n, 1)))
    if k == 1:
        e = np.zeros(m * m)
        e[0] = 1.0
        assert np.isclose(_prolong(e, n, 1).reshape(n, n)[0, 0], 0.75 * 0.75)
        assert np.isclose(_prolong(e, n, 1).reshape(n, n)[0, 2], 0.75 * 0.25)


@pytest.mark.parametrize("n,prec,mode", [(256, "fp32", "learned-coarse"), (512, "bf16", "learned-coarse"),
                                         (512, "fp32", "learned-quarter")])
def test_learned_coarse_guess_matches_torch(fs, O, n, prec, mode):
    # the network at half (quarter) resolution: the 2x2 (4x4) averages of r and p_prev
    # in, the increment interpolated bilinearly onto the grid; against the float64
    # torch restatement on the same averages (the residual as stored: BF16 under BF16
    # storage)
    from fluidsim import initial
    from guess_torch import bf16_round, reference_guess
    k = 1 if mode == "learned-coarse" else 2
    blob = _random_guess_blob(3)
    u0, v0 = initial.decaying_turbulence(n, seed=7)
    with solver(fs, n=n, pressureSolver="multigrid", pressureIterations=0, mgMinSize=4, mgPrecision=prec) as g:
        g.uploadVelocity(u0, v0)
        g.loadGuess(blob)
        g.setGuess(mode)
        g.runKernel("project", 0.0)
        p = g.readField("pressure")
    r = np.asarray(O.divergence(n, n, f32(u0), f32(v0)), np.float32)
    if prec == "bf16":
        r = bf16_round(r)
    m = n >> k
    rc = _restrict(r.astype(np.float64), n, k).astype(np.float32).astype(np.float64)
    dc = reference_guess(blob, rc, np.zeros(m * m), m, (1 << k) * TWO_PI / n)
    ref = _prolong(dc, n, k)
    assert np.max(np.abs(ref)) > 0
    err = rel(p, ref)
    assert err < 3e-2, err


def test_learned_coarse_guess_keeps_the_gate_and_rejects_odd_sizes(fs, O):
    from fluidsim import initial
    n = 256
    blob = _random_guess_blob(4)
    u0, v0 = initial.decaying_turbulence(n, seed=1)
    with solver(fs, n=n, nu=1e-4, advection="maccormack", pressureSolver="multigrid",
                pressureIterations=40, tolerance=1e-6, mgMinSize=4) as g:
        g.loadGuess(blob)
        g.setGuess("learned-coarse")
        g.uploadVelocity(u0, v0)
        for _ in range(4):
            g.stepCFL(0.5)
            assert g.stats()["last_residual"] <= 1e-6
        out = g.readVelocity()
    s = O.Solver(n=n, nu=1e-4, advection="maccormack", explicit_diffusion=True, projection="fft")
    s.u[:], s.v[:] = f32(u0), f32(v0)
    h = TWO_PI / n
    for _ in range(4):
        s.step(0.5 / (np.max(np.abs(s.u)) / h + np.max(np.abs(s.v)) / h))
    assert max(rel(out["u"], s.u), rel(out["v"], s.v)) < 1e-5
    with solver(fs, n=128, pressureSolver="multigrid", mgMinSize=4) as g:
        g.loadGuess(blob)
        with pytest.raises(Exception, match="multiple of 256"):
            g.setGuess("learned-coarse")
    with solver(fs, n=256, pressureSolver="multigrid", mgMinSize=4) as g:
        g.loadGuess(blob)
        with pytest.raises(Exception, match="multiple of 512"):
            g.setGuess("learned-quarter")


def test_learned_guess_keeps_the_gate(fs, O):
    # whatever the weights, the refinement after the guess meets the FP64 gate and the
    # velocity tracks the exact projection
    from fluidsim import initial
    n = 256
    blob = _random_guess_blob(5)
    u0, v0 = initial.decaying_turbulence(n, seed=1)
    with solver(fs, n=n, nu=1e-4, advection="maccormack", pressureSolver="multigrid",
                pressureIterations=40, tolerance=1e-6, mgMinSize=4) as g:
        g.loadGuess(blob)
        g.setGuess("learned")
        g.uploadVelocity(u0, v0)
        for _ in range(4):
            g.stepCFL(0.5)
            assert g.stats()["last_residual"] <= 1e-6
        out = g.readVelocity()
    s = O.Solver(n=n, nu=1e-4, advection="maccormack", explicit_diffusion=True, projection="fft")
    s.u[:], s.v[:] = f32(u0), f32(v0)
    h = TWO_PI / n
    for _ in range(4):
        s.step(0.5 / (np.max(np.abs(s.u)) / h + np.max(np.abs(s.v)) / h))
    assert max(rel(out["u"], s.u), rel(out["v"], s.v)) < 1e-5


def test_failed_weights_load_falls_back_to_the_previous_pressure(fs):
    # a bad blob after a good one: the load fails, the solver keeps stepping with the
    # previous-pressure guess instead of a network without weights; after a good
    # reload the learned mode can be selected again
    from fluidsim import initial
    n = 256
    blob = _random_guess_blob(8)
    u0, v0 = initial.decaying_turbulence(n, seed=2)
    with solver(fs, n=n, nu=1e-4, advection="maccormack", pressureSolver="multigrid",
                pressureIterations=40, tolerance=1e-6, mgMinSize=4) as g:
        g.loadGuess(blob)
        g.setGuess("learned-coarse")
        g.uploadVelocity(u0, v0)
        g.stepCFL(0.5)
        bad = blob[:12] + blob[12:40] + b"\x00" * 8 + blob[48:]
        with pytest.raises(Exception):
            g.loadGuess(bad[:len(bad) - 100])
        for _ in range(2):
            g.stepCFL(0.5)
            assert g.stats()["last_residual"] <= 1e-6
        g.loadGuess(blob)
        g.setGuess("learned-coarse")
        g.stepCFL(0.5)
        assert g.stats()["last_residual"] <= 1e-6


def test_learned_guess_reads_the_bf16_stored_residual(fs, O):
    # under BF16 multigrid storage the residual the network reads is stored in BF16:
    # the same restatement, fed that rounded residual
    from fluidsim import initial
    from guess_torch import bf16_round, reference_guess
    n = 256
    blob = _random_guess_blob(2)
    u0, v0 = initial.decaying_turbulence(n, seed=5)
    with solver(fs, n=n, pressureSolver="multigrid", pressureIterations=0, mgMinSize=4, mgPrecision="bf16") as g:
        g.uploadVelocity(u0, v0)
        g.loadGuess(blob)
        g.setGuess("learned")
        g.runKernel("project", 0.0)
        p = g.readField("pressure")
    r = O.divergence(n, n, f32(u0), f32(v0))
    ref = reference_guess(blob, bf16_round(np.asarray(r, np.float32)).astype(np.float64), np.zeros(n * n), n,
                          TWO_PI / n)
    assert rel(p, ref) < 3e-2


def test_learned_guess_with_bf16_storage_keeps_the_gate(fs, O):
    from fluidsim import initial
    n = 256
    blob = _random_guess_blob(6)
    u0, v0 = initial.decaying_turbulence(n, seed=2)
    with solver(fs, n=n, nu=1e-4, advection="maccormack", pressureSolver="multigrid", pressureIterations=40,
                tolerance=1e-6, mgMinSize=4, mgPrecision="bf16") as g:
        g.loadGuess(blob)
        g.setGuess("learned")
        g.uploadVelocity(u0, v0)
        for _ in range(4):
            g.stepCFL(0.5)
            assert g.stats()["last_residual"] <= 1e-6
        out = g.readVelocity()
    s = O.Solver(n=n, nu=1e-4, advection="maccormack", explicit_diffusion=True, projection="fft")
    s.u[:], s.v[:] = f32(u0), f32(v0)
    h = TWO_PI / n
    for _ in range(4):
        s.step(0.5 / (np.max(np.abs(s.u)) / h + np.max(np.abs(s.v)) / h))
    assert max(rel(out["u"], s.u), rel(out["v"], s.v)) < 1e-5


# ------------------------------------------------------------ energy spectrum

@pytest.mark.parametrize("n", [64, 256])
def test_energy_spectrum_matches_oracle(fs, O, n):
    # energySpectrum (src/core/fft.js:196-234) on the device against the oracle's FFT of
    # the same centre velocities; Parseval against the physical-space energy
    from fluidsim import initial
    u0, v0 = initial.decaying_turbulence(n, seed=4)
    with solver(fs, n=n) as g:
        g.uploadVelocity(u0, v0)
        k, E, cnt = g.energySpectrum()
        vel = g.readVelocity()
    cu, cv = O.velocity_at_centers(n, n, f32(vel["u"]), f32(vel["v"]))
    Eo, co = O.energy_spectrum(n, n, cu, cv)
    assert np.array_equal(cnt, co)
    tot = np.sum(Eo)
    assert np.max(np.abs(E - Eo)) <= 1e-12 * tot
    phys = 0.5 * np.mean(cu * cu + cv * cv)
    corners = phys - tot                   # modes beyond n/2 are dropped, as in the reference
    assert -1e-12 * phys <= corners <= 1e-3 * phys
    # a single Taylor-Green mode sits in shell 1 (sqrt 2 rounds to 1)
    tu, tv = O.taylor_green(n, n)
    with solver(fs, n=n) as g:
        g.uploadVelocity(tu, tv)
        _, Et, _ = g.energySpectrum()
    assert Et[1] > 0.99 * np.sum(Et)


@pytest.mark.parametrize("nx,ny,nu,dt_fixed,cfl", [(256, 256, 1e-3, None, 0.5), (80, 48, 5e-4, None, 0.5),
                                                   (64, 64, 0.0, None, 0.5), (128, 128, 1e-3, 0.08, None),
                                                   (256, 256, 1e-3, None, 1.9)])
def test_fused_advection_diffusion_gives_the_same_bits(fs, nx, ny, nu, dt_fixed, cfl):
    # the one-launch MacCormack + diffusion tile kernel against the three launches;
    # a fixed dt of 0.08 moves samples several cells, outside the LDS windows; CFL 1.9
    # (the sum of both components' moves) puts the largest moves just under one cell,
    # the edge of the windows' rings
    from fluidsim import initial
    n = max(nx, ny)
    u0, v0 = initial.decaying_turbulence(n, seed=11)
    u0 = np.asarray(u0).reshape(n, n)[:ny, :nx].reshape(-1)
    v0 = np.asarray(v0).reshape(n, n)[:ny, :nx].reshape(-1)
    outs = []
    for fused in (True, False):
        # multigrid is square-only: the non-square grid projects with red-black sweeps
        proj = dict(pressureSolver="multigrid", pressureIterations=30, tolerance=1e-7, mgMinSize=4) \
            if nx == ny else dict(pressureSolver="red-black", pressureIterations=60)
        with solver(fs, nx=nx, ny=ny, nu=nu, advection="maccormack", advFused=fused, **proj) as g:
            g.uploadVelocity(u0, v0)
            for _ in range(3):
                if dt_fixed:
                    g.step(dt_fixed)
                else:
                    g.stepCFL(cfl)
            outs.append(g.readVelocity())
    assert np.array_equal(outs[0]["u"], outs[1]["u"]) and np.array_equal(outs[0]["v"], outs[1]["v"])
    assert np.max(np.abs(outs[0]["u"])) > 0
