"""Pins of the f64 CPU oracle: reference-held records, the reference's own
known-answer tests, and closed forms chosen so a plausible mistake (wrong
offset, sign, factor or colour order) fails one of them. CPU only.
"""
import math

import numpy as np
import pytest

N = 64
TWO_PI = 2 * math.pi


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


# ------------------------------------------------ reference-held records (GPU harness)

def test_lcg_and_divergence_reproduce_reference_max_div(O, golden):
    # tools/gputest.html:155-167 + results/gpu-verify.json projection.before
    u, v = O.js_random_velocity(N, 4242)
    md = O.max_divergence(N, N, f32(u), f32(v))
    ref = golden["projection"]["red-black"]["before"]
    assert abs(md - ref) / ref < 2e-7


@pytest.mark.parametrize("solver", ["jacobi", "red-black"])
def test_projection_after_200_sweeps_matches_reference(O, golden, solver):
    # gputest block 3: maccormack advection (the default), nu=0, 200 sweeps, dt=0.01
    u, v = O.js_random_velocity(N, 4242)
    s = O.Solver(n=N, nu=0.0, advection="maccormack", projection=solver, iterations=200)
    s.u[:] = f32(u)
    s.v[:] = f32(v)
    s.step(0.01)
    after = O.max_divergence(N, N, s.u, s.v)
    ref = golden["projection"][solver]["after"]
    assert abs(after - ref) / ref < 2e-5


@pytest.mark.parametrize("solver", ["jacobi", "red-black", "multigrid"])
def test_poisson_curves_match_reference(O, golden, solver):
    # gputest block 4 (tools/gputest.html:205-246): seed 777, residual after k iterations
    u, v = O.js_random_velocity(N, 777)
    d = O.divergence(N, N, f32(u), f32(v))
    for it, ref in golden["poissonCurves"][solver]:
        if solver == "jacobi":
            p = O.jacobi(N, N, d, it)
        elif solver == "red-black":
            p = O.redblack(N, N, d, it)
        else:
            p, _ = O.mg_vcycles(N, d, it, min_size=8, pre=2, post=2, coarse=40)
        r = O.poisson_residual_rel(N, N, p, d) if it > 0 else 1.0
        if solver == "multigrid" and it >= 5:
            # the reference's FP32 plateau (~9e-8); the f64 oracle keeps converging
            assert r < ref
            continue
        tol = 5e-3 if (solver == "multigrid" and it == 4) else 1e-4
        assert abs(r - ref) / ref < tol, (solver, it, r, ref)


@pytest.mark.parametrize("scheme", ["semi-lagrangian", "maccormack"])
def test_scheme_energy_after_200_tg_steps(O, golden, scheme):
    # gputest block 7: TG 64^2, nu=0, dt=0.01, multigrid 3 V-cycles, KE of raw faces
    s = O.Solver(n=N, nu=0.0, advection=scheme, projection="multigrid", iterations=3)
    u, v = O.taylor_green(N, N)
    s.u[:] = f32(u)
    s.v[:] = f32(v)
    for _ in range(200):
        s.step(0.01)
    ke = O.kinetic_energy_faces(s.u, s.v)
    ref = golden["schemeEnergy"][scheme]
    assert abs(ke - ref) / ref < 1e-5
    assert golden["schemeEnergy"]["maccormack"] > golden["schemeEnergy"]["semi-lagrangian"]


# ------------------------------------------------ the reference's known-answer tests

def test_wrap_known_answers(O):
    # test/core.test.js:35-42
    L = O.lib()
    assert [L.or_wrap(i, 8) for i in (0, 7, 8, -1, -9, 17)] == [0, 7, 0, 7, 7, 1]


def test_bilerp_known_answers_and_wrap(O):
    # test/core.test.js:44-64
    a = np.tile(np.arange(4, dtype=np.float64), 4)
    assert O.bilerp(a, 4, 4, 1, 1) == 1
    assert O.bilerp(a, 4, 4, 2, 0) == 2
    assert abs(O.bilerp(a, 4, 4, 1.5, 0) - 1.5) < 1e-15
    assert abs(O.bilerp(a, 4, 4, 1.25, 0) - 1.25) < 1e-15
    assert O.bilerp(a, 4, 4, 4, 0) == O.bilerp(a, 4, 4, 0, 0)
    assert O.bilerp(a, 4, 4, -1, 0) == O.bilerp(a, 4, 4, 3, 0)


def test_samplers_round_trip_their_own_positions(O):
    # test/core.test.js:66-83: catches a half-cell MAC offset
    n = 16
    dx = TWO_PI / n
    i = np.arange(n)[None, :]
    j = np.arange(n)[:, None]
    u = (np.sin(i * dx) * np.cos((j + 0.5) * dx)).reshape(-1)
    v = (np.cos((i + 0.5) * dx) * np.sin(j * dx)).reshape(-1)
    p = np.sin((i + 0.5) * dx + (j + 0.5) * dx).reshape(-1)
    for (a, b) in [(0, 0), (3, 5), (15, 15), (8, 2)]:
        assert abs(O.sample("u", n, n, u, a * dx, (b + 0.5) * dx) - u[b * n + a]) < 1e-12
        assert abs(O.sample("v", n, n, v, (a + 0.5) * dx, b * dx) - v[b * n + a]) < 1e-12
        assert abs(O.sample("p", n, n, p, (a + 0.5) * dx, (b + 0.5) * dx) - p[b * n + a]) < 1e-12


def test_divergence_gradient_exact_negative_adjoints(O):
    # test/core.test.js:85-120 (seed 12345, 16^2)
    n = 16
    vals, _ = O.js_lcg(12345, 3 * n * n)
    wu, wv, q = vals[: n * n], vals[n * n: 2 * n * n], vals[2 * n * n:]
    lhs = float(np.dot(O.divergence(n, n, wu, wv), q))
    gu, gv = O.subtract_gradient(n, n, np.zeros(n * n), np.zeros(n * n), q, scale=-1.0)
    rhs = float(np.dot(wu, gu) + np.dot(wv, gv))
    assert abs(lhs + rhs) / max(abs(lhs), abs(rhs)) < 1e-12


def test_taylor_green_discretely_divergence_free(O):
    # test/core.test.js:138-149
    for n in (16, 32, 64):
        u, v = O.taylor_green(n, n, 0.0, 0.01)
        assert O.max_divergence(n, n, u, v) < 1e-14
    u = np.full(32 * 32, 1.3)
    v = np.full(32 * 32, -0.7)
    assert O.max_divergence(32, 32, u, v) < 1e-14


def test_laplacian_symmetric_and_annihilates_constants(O):
    # test/solver.test.js:31-58
    n = 16
    vals, _ = O.js_lcg(4242, 2 * n * n)
    x, y = vals[: n * n], vals[n * n:]
    lhs = np.dot(O.laplacian(n, n, x), y)
    rhs = np.dot(x, O.laplacian(n, n, y))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
    assert np.max(np.abs(O.laplacian(n, n, np.full(n * n, 3.7)))) < 1e-12


def test_kinetic_energy_of_taylor_green(O):
    # test/core.test.js:160-173
    u, v = O.taylor_green(128, 128, 0.7, 0.03)
    exact = 0.25 * math.exp(-4 * 0.03 * 0.7)
    assert abs(O.kinetic_energy(128, 128, u, v) - exact) / exact < 2e-3


# ------------------------------------------------------------------ FFT / Poisson

def test_fft_matches_naive_dft(O):
    n = 32
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    re, im = x.copy(), y.copy()
    assert O.lib().or_fft(n, re, im, -1) == 0
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    ref = W @ (x + 1j * y)
    assert np.max(np.abs(re + 1j * im - ref)) < 1e-12


def test_fft2_single_cosine_mode(O):
    # test/core.test.js:298-316
    nx = ny = 32
    k0 = 5
    f = np.cos(2 * np.pi * k0 * np.arange(nx) / nx)[None, :].repeat(ny, 0).reshape(-1)
    re, im = O.fft2(f, np.zeros(nx * ny), nx, ny)
    mag = np.hypot(re, im).reshape(ny, nx) / (nx * ny)
    expect = np.zeros((ny, nx))
    expect[0, k0] = expect[0, nx - k0] = 0.5
    assert np.max(np.abs(mag - expect)) < 1e-12


def test_energy_spectrum_parseval_for_taylor_green(O):
    # test/core.test.js:318-327 (no corner modes for TG, so equality is exact)
    u, v = O.taylor_green(64, 64, 0.0, 0.01)
    cu, cv = O.velocity_at_centers(64, 64, u, v)
    E, _ = O.energy_spectrum(64, 64, cu, cv)
    phys = 0.5 * np.mean(cu * cu + cv * cv)
    assert abs(E.sum() - phys) / phys < 1e-12


def test_fft_poisson_single_mode_closed_form(O):
    # rhs = sin(kx x) sin(ky y) at centres solves exactly to rhs / lambda(k)
    n = 32
    h = TWO_PI / n
    c = (np.arange(n) + 0.5) * h
    kx, ky = 3, 5
    rhs = (np.sin(kx * c)[None, :] * np.sin(ky * c)[:, None]).reshape(-1)
    lam = -4 * math.sin(math.pi * kx / n) ** 2 / h ** 2 - 4 * math.sin(math.pi * ky / n) ** 2 / h ** 2
    p = O.poisson_fft(n, n, rhs)
    assert np.max(np.abs(p - rhs / lam)) < 1e-13
    assert np.max(np.abs(O.laplacian(n, n, p) - rhs)) < 1e-11


def test_cg_equals_fft_projection(O):
    # test/validation.test.js:201-221
    n = 32
    u, v = O.js_random_velocity(n, 777)
    d = O.divergence(n, n, u, v)
    pf = O.poisson_fft(n, n, d)
    pc, it, res = O.poisson_cg(n, n, d, tol=1e-12, maxit=2000)
    pf -= pf.mean()
    assert res < 1e-12
    assert np.max(np.abs(pf - pc)) < 1e-9


def test_projection_divergence_free_and_idempotent(O):
    # test/solver.test.js:133-181
    n = 32
    u, v = O.js_random_velocity(n, 777)
    p = O.poisson_fft(n, n, O.divergence(n, n, u, v))
    u1, v1 = O.subtract_gradient(n, n, u, v, p)
    assert O.max_divergence(n, n, u1, v1) < 1e-10
    p2 = O.poisson_fft(n, n, O.divergence(n, n, u1, v1))
    u2, v2 = O.subtract_gradient(n, n, u1, v1, p2)
    assert np.max(np.abs(u2 - u1)) < 1e-12 and np.max(np.abs(v2 - v1)) < 1e-12


# ------------------------------------------------------------------- diffusion

def test_explicit_diffusion_checkerboard_and_sine_factors(O):
    # checkerboard amplification |1 - 8d| (test/validation.test.js:122-146)
    n = 16
    h = TWO_PI / n
    for d in (0.1, 0.25, 0.26):
        nu, dt = 1.0, d * h * h
        cb = np.fromfunction(lambda j, i: (-1.0) ** (i + j), (n, n)).reshape(-1)
        out = O.diffuse_explicit(n, n, cb, nu, dt)
        assert np.max(np.abs(out - (1 - 8 * d) * cb)) < 1e-12
    # a sine mode decays by 1 - 2a(1 - cos k h) - 2b(1 - cos l h), and the mean is kept
    k, l = 2, 3
    c = np.arange(n) * h
    f = (np.sin(k * c)[None, :] * np.cos(l * c)[:, None]).reshape(-1) + 0.3
    nu, dt = 0.02, 0.05
    a = nu * dt / h ** 2
    fac = 1 - 2 * a * (1 - math.cos(k * h)) - 2 * a * (1 - math.cos(l * h))
    out = O.diffuse_explicit(n, n, f, nu, dt)
    assert np.max(np.abs((out - 0.3) - fac * (f - 0.3))) < 1e-13
    assert abs(out.mean() - f.mean()) < 1e-14


def test_implicit_diffusion_inverts_each_fourier_mode(O):
    # (I - nu dt L) out = f: every Fourier mode of f divided by 1 + 4a sin^2(kh/2) +
    # 4b sin^2(lh/2) (the 5-point Laplacian's eigenvalue); checked on a seeded field
    # through numpy's FFT, and a single sine mode by its closed factor
    nx, ny = 32, 16
    hx, hy = TWO_PI / nx, TWO_PI / ny
    nu, dt = 0.7, 0.3
    rng = np.random.default_rng(5)
    f = rng.standard_normal((ny, nx))
    out, it = O.diffuse_implicit(nx, ny, f.reshape(-1), nu, dt)
    assert 0 < it < 400
    kx = np.fft.fftfreq(nx) * nx
    ky = np.fft.fftfreq(ny) * ny
    den = (1 + 4 * nu * dt / hx ** 2 * np.sin(kx[None, :] * hx / 2) ** 2
           + 4 * nu * dt / hy ** 2 * np.sin(ky[:, None] * hy / 2) ** 2)
    want = np.real(np.fft.ifft2(np.fft.fft2(f) / den))
    assert np.max(np.abs(out.reshape(ny, nx) - want)) < 1e-10
    # a dropped factor 4, a wrong sign or an explicit step instead all miss by far more
    for bad in (np.real(np.fft.ifft2(np.fft.fft2(f) / (2 * den - 1))),
                np.real(np.fft.ifft2(np.fft.fft2(f) * (2 - den)))):
        assert np.max(np.abs(bad - want)) > 1e-2
    c = np.arange(nx) * hx
    g = np.tile(np.sin(3 * c), ny)
    fac = 1 + 4 * nu * dt / hx ** 2 * math.sin(3 * hx / 2) ** 2
    out, _ = O.diffuse_implicit(nx, ny, g, nu, dt)
    assert np.max(np.abs(out - g / fac)) < 1e-12
    same, it0 = O.diffuse_implicit(nx, ny, g, 0.0, dt)
    assert it0 == 0 and np.array_equal(same, g)


# ------------------------------------------------------------------- advection

@pytest.mark.parametrize("scheme", ["semi-lagrangian", "maccormack"])
def test_uniform_integer_shift_is_exact(O, scheme):
    # a uniform flow moving a whole number of cells per step transports exactly
    n = 32
    h = TWO_PI / n
    dt = 0.1
    U, V = 2 * h / dt, -1 * h / dt
    q = O.js_lcg(31337, n * n)[0]
    u = np.full(n * n, U)
    v = np.full(n * n, V)
    out = O.advect_scalar(scheme, n, n, u, v, q, dt)
    expect = np.roll(np.roll(q.reshape(n, n), 2, axis=1), -1, axis=0).reshape(-1)
    assert np.max(np.abs(out - expect)) < 1e-12
    # the velocity itself is a constant and must stay exactly constant
    ou, ov = O.advect_velocity(scheme, n, n, u, v, dt)
    assert np.max(np.abs(ou - U)) < 1e-12 and np.max(np.abs(ov - V)) < 1e-12


def _gaussian_error(O, scheme, n):
    # translating Gaussian (src/core/analytic.js:100-140) after T=1 with CFL ~0.5
    h = TWO_PI / n
    ax, ay, sig = 1.0, 0.6, 0.5
    c = (np.arange(n) + 0.5) * h

    def blob(t):
        out = np.zeros((n, n))
        for mx in (-1, 0, 1):
            for my in (-1, 0, 1):
                X = c[None, :] - (math.pi + ax * t) + mx * TWO_PI
                Y = c[:, None] - (math.pi + ay * t) + my * TWO_PI
                out += np.exp(-(X * X + Y * Y) / (2 * sig * sig))
        return out.reshape(-1)

    dt = 0.5 * h / (ax + ay)
    steps = int(round(1.0 / dt))
    dt = 1.0 / steps
    u = np.full(n * n, ax)
    v = np.full(n * n, ay)
    q = blob(0.0)
    for _ in range(steps):
        q = O.advect_scalar(scheme, n, n, u, v, q, dt)
    e = q - blob(1.0)
    return math.sqrt(np.mean(e * e))


def test_maccormack_second_order_and_less_diffusive(O):
    e_sl = [_gaussian_error(O, "semi-lagrangian", n) for n in (32, 64)]
    e_mc = [_gaussian_error(O, "maccormack", n) for n in (32, 64)]
    order_sl = math.log2(e_sl[0] / e_sl[1])
    order_mc = math.log2(e_mc[0] / e_mc[1])
    assert 0.7 < order_sl < 1.4
    assert order_mc > 1.6
    assert e_mc[1] < e_sl[1] / 4


def test_maccormack_limiter_keeps_bounds(O):
    # test/solver.test.js:275-295: a [0,1] step function stays in [0,1]
    n = 32
    u, v = O.taylor_green(n, n)
    q = np.zeros((n, n))
    q[8:20, 10:24] = 1.0
    q = q.reshape(-1)
    for _ in range(20):
        q = O.advect_scalar("maccormack", n, n, u, v, q, 0.05)
    assert q.min() >= -1e-15 and q.max() <= 1 + 1e-15


# ------------------------------------------------------------------- multigrid

def test_multigrid_levels_follow_reference_rule(O):
    # src/gpu/multigrid.js:182-197
    assert O.mg_levels(64, 8) == 4
    assert O.mg_levels(1024, 4) == 9
    assert O.mg_levels(1024, 8) == 8
    assert O.mg_levels(96, 4) == 5     # 96, 48, 24, 12, 6: 3 would be odd
    assert O.mg_levels(100, 4) == 2    # 100, 50: 25 would be odd


# ------------------------------------------------------------------- full step

def test_taylor_green_decay_f64(O):
    # Taylor-Green 64^2, nu=0.01, MacCormack, explicit diffusion, exact projection,
    # dt at CFL 0.5 (src/core/grid.js:322-327), 200 steps
    n, nu = 64, 0.01
    s = O.Solver(n=n, nu=nu, advection="maccormack", explicit_diffusion=True, projection="fft")
    s.u[:], s.v[:] = O.taylor_green(n, n)
    h = TWO_PI / n
    dt = 0.5 / (np.max(np.abs(s.u)) / h + np.max(np.abs(s.v)) / h)
    ke0 = O.kinetic_energy(n, n, s.u, s.v)
    for _ in range(200):
        s.step(dt)
    T = 200 * dt
    err = O.velocity_error_tg(n, n, s.u, s.v, T, nu)
    ke = O.kinetic_energy(n, n, s.u, s.v)
    rate = -math.log(ke / ke0) / T
    # first order in time (Lie splitting) at d = nu dt / h^2 = 0.0255: the scheme's own
    # dissipation adds to 4 nu, it never removes from it
    assert err < 0.05
    assert 4 * nu < rate < 1.35 * 4 * nu
    assert O.max_divergence(n, n, s.u, s.v) < 1e-12


def test_taylor_green_reference_gate(O):
    # test/solver.test.js:339-361 with the explicit diffusion the GPU path uses
    n, nu, dt = 64, 0.05, 0.01
    s = O.Solver(n=n, nu=nu, advection="maccormack", explicit_diffusion=True, projection="fft")
    s.u[:], s.v[:] = O.taylor_green(n, n)
    for _ in range(50):
        s.step(dt)
    assert O.velocity_error_tg(n, n, s.u, s.v, 0.5, nu) < 0.05
    ke = O.kinetic_energy(n, n, s.u, s.v)
    assert ke < 0.25 * math.exp(-4 * nu * 0.5) * 1.02


@pytest.mark.parametrize("n", [16, 32, 64])
def test_enstrophy_of_taylor_green_is_its_discrete_closed_form(O, n):
    # src/core/grid.js:297-303. For u = sin x cos y, v = -cos x sin y at the faces the
    # corner vorticity is 2 s sin x sin y (s = sin(h/2)/(h/2)), its centre average
    # 2 s cos^2(h/2) sin x_c sin y_c, and 1/2 <omega_c^2> = s^2 cos^4(h/2) / 2, which
    # tends to the continuum 1/2 <(2 sin x sin y)^2> = 1/2
    h = TWO_PI / n
    u, v = O.taylor_green(n, n)
    s = math.sin(h / 2) / (h / 2)
    assert abs(O.enstrophy(n, n, u, v) - 0.5 * s ** 2 * math.cos(h / 2) ** 4) < 1e-14

def test_js_lcg_emulation_equals_the_js_engine(O, tmp_path):
    # the oracle's IEEE-double emulation of gputest's seeded generator against Node
    # evaluating the reference's own expression (tools/gputest.html:113-114)
    import os
    import shutil
    import subprocess
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed on this machine")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    f = str(tmp_path / "lcg.f64")
    seeds = [991, 4242, 777, 1]
    out = subprocess.run([node, os.path.join(root, "tests", "node", "lcg_fields.mjs"), f, str(N),
                          *map(str, seeds)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    js = np.fromfile(f, np.float64).reshape(len(seeds), 2, N * N)
    for k, seed in enumerate(seeds):
        u, v = O.js_random_velocity(N, seed)
        assert np.array_equal(js[k, 0], u) and np.array_equal(js[k, 1], v), seed

def test_band_forcing_modes_deterministic_and_in_shell(O):
    kx, ky, ph = O.band_modes(1000, 32.0, 2.0)
    kx2, ky2, ph2 = O.band_modes(1000, 32.0, 2.0)
    assert (kx == kx2).all() and (ky == ky2).all() and (ph == ph2).all()
    k = np.hypot(kx, ky)
    assert (k >= 30 - 1).all() and (k <= 34 + 1).all()
    assert (ph >= 0).all() and (ph < TWO_PI).all()
    kx3, _, _ = O.band_modes(1001, 32.0, 2.0)
    assert not (kx3 == kx).all()



def _face_coefficients(F, n, comp):
    """Complex Fourier coefficient C(k) of a real face field sampled at its MAC positions,
    normalised so that c cos(k.x + phi) has C(k) = c e^{i phi}: u faces sit at
    (i h, (j + 1/2) h), v faces at ((i + 1/2) h, j h) on [0, 2pi]^2."""
    h = TWO_PI / n
    Z = np.fft.fft2(F.reshape(n, n)) * (2.0 / (n * n))   # Z[ky, kx] = 2/N sum F e^{-i k.(i h, j h)}
    k = np.fft.fftfreq(n, 1.0 / n)
    KX, KY = np.meshgrid(k, k)
    shift = np.exp(-0.5j * h * (KY if comp == "u" else KX))  # the staggered half cell
    return Z * shift, KX, KY


def _forcing_pin_error(O, Fu, Fv, n, seed, kf, width, amp, dt):
    """Largest deviation of the faces' Fourier coefficients from the band forcing's
    closed form: c_u = amp dt / sqrt(24) ky/|k|, c_v = -amp dt / sqrt(24) kx/|k| at each
    mode's wavevector with its phase (modes sharing a wavevector, or opposite ones, add),
    and nothing at any other wavevector. Relative to amp dt."""
    kx, ky, ph = O.band_modes(seed, kf, width)
    scale = amp * dt / math.sqrt(24)
    err = 0.0
    for comp, F in (("u", Fu), ("v", Fv)):
        C, KX, KY = _face_coefficients(F, n, comp)
        want = np.zeros_like(C)
        for a, b, p in zip(kx, ky, ph):
            if a == 0 and b == 0:
                continue
            c = scale * (b if comp == "u" else -a) / math.hypot(a, b)
            want[(KX == a) & (KY == b)] += c * np.exp(1j * p)
            want[(KX == -a) & (KY == -b)] += c * np.exp(-1j * p)
        err = max(err, float(np.max(np.abs(C - want))) / (amp * dt))
    return err


def test_band_forcing_field_is_the_closed_form_at_every_mode(O):
    # src/cpu/solver.js:200-222: each mode adds the curl of a streamfunction, sampled at
    # the faces; its Fourier coefficients on the grid are fixed by the mode list alone
    n, kf, width, amp, dt = 64, 8.0, 2.0, 0.4, 0.01
    z = np.zeros(n * n)
    for seed in (1, 1000, 4242):
        Fu, Fv = O.band_forcing(n, n, z, z, kf, width, amp, seed, dt)
        assert _forcing_pin_error(O, Fu, Fv, n, seed, kf, width, amp, dt) < 1e-12


def test_band_forcing_pin_rejects_planted_mistakes(O):
    # the pin above fails for a flipped v sign, a half-cell face offset and 1/M scaling
    n, kf, width, amp, dt, seed = 64, 8.0, 2.0, 0.4, 0.01, 1000
    z = np.zeros(n * n)
    Fu, Fv = O.band_forcing(n, n, z, z, kf, width, amp, seed, dt)
    assert _forcing_pin_error(O, Fu, -Fv, n, seed, kf, width, amp, dt) > 1e-2
    assert _forcing_pin_error(O, Fu * math.sqrt(24) / 24, Fv * math.sqrt(24) / 24, n, seed, kf, width, amp,
                              dt) > 1e-2
    # u sampled at the cell centre row y = j h instead of the face row (j + 1/2) h
    h = TWO_PI / n
    kx, ky, ph = O.band_modes(seed, kf, width)
    I, J = np.meshgrid(np.arange(n), np.arange(n))
    Gu = np.zeros((n, n))
    for a, b, p in zip(kx, ky, ph):
        if a or b:
            Gu += amp * dt / math.sqrt(24) * b / math.hypot(a, b) * np.cos(a * I * h + b * J * h + p)
    assert _forcing_pin_error(O, Gu.reshape(-1), Fv, n, seed, kf, width, amp, dt) > 1e-2


def test_mode_list_forcing_is_its_closed_form(O):
    # FORCING (src/gpu/shaders.js:532-566): drag on the old velocity, then
    # amp dt / sqrt(count) (ky/|k|, -kx/|k|) cos(k.x + phase + drift t) at the faces; a
    # zero wavevector counts in sqrt(count) but adds nothing
    n, amp, drag, drift, t, dt = 64, 1.9, 0.42, 0.7, 1.3, 0.01
    modes = [dict(kx=3, ky=-2, phase=0.4), dict(kx=0, ky=5, phase=2.0), dict(kx=0, ky=0, phase=1.0),
             dict(kx=-4, ky=1, phase=5.5)]
    z = np.zeros(n * n)
    Fu, Fv = O.mode_forcing(n, n, z, z, modes, amp, drag, drift, t, dt)
    scale = amp * dt / math.sqrt(len(modes))
    for comp, F in (("u", Fu), ("v", Fv)):
        C, KX, KY = _face_coefficients(F, n, comp)
        want = np.zeros_like(C)
        for m in modes:
            a, b = m["kx"], m["ky"]
            if a == 0 and b == 0:
                continue
            c = scale * (b if comp == "u" else -a) / math.hypot(a, b)
            ph = m["phase"] + drift * t
            want[(KX == a) & (KY == b)] += c * np.exp(1j * ph)
            want[(KX == -a) & (KY == -b)] += c * np.exp(-1j * ph)
        assert np.max(np.abs(C - want)) < 1e-12 * amp * dt
    # linear in the velocity with the drag factor on the old velocity alone: the CPU
    # order (u + f) exp(-drag dt) would scale the force as well
    rng = np.random.default_rng(1)
    u0, v0 = rng.standard_normal((2, n * n))
    Gu, Gv = O.mode_forcing(n, n, u0, v0, modes, amp, drag, drift, t, dt)
    damp = math.exp(-drag * dt)
    assert np.max(np.abs(Gu - (damp * u0 + Fu))) < 1e-14 and np.max(np.abs(Gv - (damp * v0 + Fv))) < 1e-14
    assert np.max(np.abs(Gu - damp * (u0 + Fu))) > 1e-6


def test_dye_step_fades_clamps_and_shifts_exactly(O):
    # ADVECT_SCALAR (src/gpu/shaders.js:234-266): at rest the dye is only faded and
    # clamped to [0, 1]; a uniform flow that moves every trace by whole cells shifts each
    # channel exactly (brute force)
    n, dt = 32, 0.1
    h = TWO_PI / n
    rng = np.random.default_rng(2)
    dye = rng.uniform(-0.5, 1.5, 4 * n * n)
    z = np.zeros(n * n)
    out = O.advect_dye(n, n, z, z, dye, dt, fade=0.1)
    assert np.max(np.abs(out - np.clip(dye * (1.0 - 0.1), 0.0, 1.0))) < 1e-14
    d01 = rng.uniform(0, 1, 4 * n * n)
    u = np.full(n * n, 2 * h / dt)      # two cells to the right
    v = np.full(n * n, -1 * h / dt)     # one cell down
    out = O.advect_dye(n, n, u, v, d01, dt)
    want = np.roll(np.roll(d01.reshape(n, n, 4), 2, axis=1), -1, axis=0).reshape(-1)
    assert np.max(np.abs(out - want)) < 1e-12


def test_splat_is_a_periodic_gaussian_about_the_cell_centres(O):
    # SPLAT (src/gpu/shaders.js:474-511): weight 1 at a cell centre the splat sits on,
    # exp(-h^2/r^2) one cell away, the same across the periodic seam, dye splats add
    # their colour and leave the velocity alone
    n = 64
    h = TWO_PI / n
    z = np.zeros(n * n)
    dz = np.zeros(4 * n * n)
    r = 3 * h
    s = dict(x=5.5 * h, y=0.5 * h, dx=2.0, dy=-1.0, radius=r)
    u, v, dye = O.splat(n, n, z, z, dz, s)
    U, V = u.reshape(n, n), v.reshape(n, n)
    assert U[0, 5] == 2.0 and V[0, 5] == -1.0
    assert abs(U[0, 6] - 2.0 * math.exp(-1 / 9)) < 1e-14 and abs(U[0, 4] - U[0, 6]) < 1e-14
    assert abs(U[n - 1, 5] - U[1, 5]) < 1e-14          # across the seam in y
    assert not dye.any()
    # the weights integrate to pi r^2 over the cells (r >> h, r << L)
    assert abs(U.sum() / 2.0 * h * h - math.pi * r * r) < 1e-9
    s2 = dict(x=0.0, y=0.0, r=0.5, g=0.25, b=1.0, radius=r, isDye=True)
    u, v, dye = O.splat(n, n, z, z, dz, s2)
    D = dye.reshape(n, n, 4)
    assert not u.any() and not v.any() and not D[:, :, 3].any()
    # the splat on the corner of four cells: they share the largest weight
    w = math.exp(-0.5 * h * h / (r * r))
    for (j, i) in ((0, 0), (0, n - 1), (n - 1, 0), (n - 1, n - 1)):
        assert abs(D[j, i, 0] - 0.5 * w) < 1e-14 and abs(D[j, i, 2] - w) < 1e-14

# ------------------------------------------------- immersed disk (src/cpu/solid.js)
# restating test/solid.test.js:29-160 on the oracle's coupling

def _disk(**kw):
    from fluidsim.solid import RigidDisk
    return RigidDisk(**kw)


def test_disk_mask_known_answers_and_wrap(O):
    d = dict(x=math.pi, y=math.pi, r=0.8)
    w = 1.5 * (TWO_PI / 128)
    assert O.disk_mask(d, math.pi, math.pi, TWO_PI, TWO_PI, w) > 0.999
    assert O.disk_mask(d, math.pi + 2.0, math.pi, TWO_PI, TWO_PI, w) < 1e-6
    assert abs(O.disk_mask(d, math.pi + 0.8, math.pi, TWO_PI, TWO_PI, w) - 0.5) < 1e-9
    seam = dict(x=0.1, y=math.pi, r=0.5)
    assert O.disk_mask(seam, TWO_PI - 0.1, math.pi, TWO_PI, TWO_PI, 0.05) > 0.9
    # Math.round semantics of the periodic separation: halves round up
    assert O.wrap_delta(math.pi, TWO_PI) == -math.pi
    assert O.wrap_delta(-math.pi, TWO_PI) == -math.pi


def test_disk_velocity_includes_rotation(O):
    d = dict(x=1, y=1, r=0.4, vx=2, vy=-1, omega=3)
    assert O.disk_velocity(d, 1, 1, TWO_PI, TWO_PI) == (2.0, -1.0)
    vx, vy = O.disk_velocity(d, 1.2, 1, TWO_PI, TWO_PI)
    assert abs(vx - 2) < 1e-12 and abs(vy - (-1 + 0.6)) < 1e-12


def _fluid_momentum(n, u, v):
    dA = (TWO_PI / n) ** 2
    return u.sum() * dA, v.sum() * dA


def test_coupling_conserves_total_momentum_with_a_free_disk(O):
    n = 64
    vals, _ = O.js_lcg(4242, 2 * n * n)
    # test/solid.test.js:64-69 uses its own LCG; any seeded field serves the invariant
    u = 1 + (vals[0::2] - 0.5)
    v = vals[1::2] - 0.5
    disk = _disk(x=math.pi, y=math.pi, r=0.7, density=2.5, vx=-0.3, vy=0.2)
    dt = 0.01
    f0, s0 = _fluid_momentum(n, u, v), disk.momentum()
    imp = O.apply_coupling(n, n, u, v, disk.as_dict())
    disk.apply_reaction(imp, dt)
    disk.integrate(dt)
    f1, s1 = _fluid_momentum(n, u, v), disk.momentum()
    before = math.hypot(f0[0] + s0[0], f0[1] + s0[1])
    drift = math.hypot(f1[0] + s1[0] - f0[0] - s0[0], f1[1] + s1[1] - f0[1] - s0[1]) / before
    assert drift < 1e-12


def test_reported_impulse_is_the_fluid_momentum_change(O):
    n = 48
    u = np.full(n * n, 1.4)
    v = np.full(n * n, -0.6)
    dt = 0.02
    f0 = _fluid_momentum(n, u, v)
    imp = O.apply_coupling(n, n, u, v, dict(x=3, y=3, r=0.6))
    f1 = _fluid_momentum(n, u, v)
    assert abs((f1[0] - f0[0]) - imp[0]) < 1e-12
    assert abs((f1[1] - f0[1]) - imp[1]) < 1e-12
    disk = _disk(x=3, y=3, r=0.6, density=1.7)
    disk.apply_reaction(imp, dt)
    assert abs(disk.force[0] * dt + imp[0]) < 1e-12 and abs(disk.force[1] * dt + imp[1]) < 1e-12


def test_pinned_disk_absorbs_momentum(O):
    n = 48
    u = np.ones(n * n)
    v = np.zeros(n * n)
    f0 = _fluid_momentum(n, u, v)[0]
    O.apply_coupling(n, n, u, v, dict(x=3, y=3, r=0.7))
    assert _fluid_momentum(n, u, v)[0] < f0 - 1e-6


@pytest.mark.parametrize("passes", [1, 2, 3])
def test_slip_falls_geometrically_with_passes(O, passes):
    # uniform flow past a disk at rest: after p passes u = (1 - chi)^p exactly, with
    # chi evaluated independently from the tanh mask at each u-face
    n = 64
    h = TWO_PI / n
    u = np.ones(n * n)
    v = np.zeros(n * n)
    d = dict(x=3.0, y=2.5, r=0.9)
    O.apply_coupling(n, n, u, v, d, passes=passes)
    i = np.arange(n)
    X, Y = np.meshgrid(i * h, (i + 0.5) * h)
    dx = (X - 3.0) - TWO_PI * np.floor((X - 3.0) / TWO_PI + 0.5)
    dy = (Y - 2.5) - TWO_PI * np.floor((Y - 2.5) / TWO_PI + 0.5)
    chi = 0.5 * (1 - np.tanh((np.hypot(dx, dy) - 0.9) / (1.5 * h)))
    expect = np.where(chi < 1e-6, 1.0, (1 - chi) ** passes).ravel()
    assert np.max(np.abs(u - expect)) < 1e-14
    assert np.all(v == 0)


def test_torque_impulse_is_the_angular_momentum_change(O):
    # solid-body rotation about the disk centre, disk at rest: the fluid loses
    # angular momentum and the reported torque impulse is exactly that change
    n = 64
    h = TWO_PI / n
    cx, cy, om = 3.1, 2.9, 0.7
    i = np.arange(n)
    Xu, Yu = np.meshgrid(i * h, (i + 0.5) * h)
    Xv, Yv = np.meshgrid((i + 0.5) * h, i * h)
    wd = lambda d: d - TWO_PI * np.floor(d / TWO_PI + 0.5)
    u = (-om * wd(Yu - cy)).ravel()
    v = (om * wd(Xv - cx)).ravel()
    L = lambda u, v: (np.sum(-wd(Yu - cy).ravel() * u) + np.sum(wd(Xv - cx).ravel() * v)) * h * h
    L0 = L(u, v)
    imp = O.apply_coupling(n, n, u, v, dict(x=cx, y=cy, r=0.8), passes=2)
    assert imp[2] < 0
    assert abs((L(u, v) - L0) - imp[2]) < 1e-12 * abs(L0)


# ---------------------------------------- refinement projection, storage rounding

def test_bf16_rounding_known_answers(O):
    # 7 explicit mantissa bits: spacing 2^-7 at 1.0; ties go to the even mantissa
    assert O.round_storage(1.0, "bf16") == 1.0
    assert O.round_storage(1.0 + 2 ** -8, "bf16") == 1.0
    assert O.round_storage(1.0 + 3 * 2 ** -8, "bf16") == 1.0 + 2 * 2 ** -7
    assert O.round_storage(1.0 + 2 ** -8 + 2 ** -20, "bf16") == 1.0 + 2 ** -7
    assert O.round_storage(-3.0, "bf16") == -3.0
    assert O.round_storage(1e-30, "bf16") == pytest.approx(1e-30, rel=2 ** -8)
    assert O.round_storage(3e38, "bf16") == pytest.approx(3e38, rel=2 ** -8)
    # FP32 storage rounds only to FP32
    assert O.round_storage(1.0 + 2 ** -30, "fp32") == 1.0
    assert O.round_storage(0.1, "fp32") == float(np.float32(0.1))


def _zero_mean_div(O, n, seed):
    vals, _ = O.js_lcg(seed, 2 * n * n)
    u = vals[0::2] - 0.5
    v = vals[1::2] - 0.5
    return O.divergence(n, n, u, v)


@pytest.mark.parametrize("n,coarse_n", [(64, 64), (128, 64), (256, 32)])
def test_refinement_without_rounding_equals_plain_vcycles(O, n, coarse_n):
    # a V-cycle is affine in its rhs, so k refinement steps from p = 0 reproduce k
    # plain V-cycles (src/gpu/solver.js:686-703), whatever the on-chip split
    div = _zero_mean_div(O, n, 5)
    p_ref, _ = O.mg_vcycles(n, div, 4, min_size=4)
    p, rmax = O.mg_refine(n, div, 4, min_size=4, coarse_n=coarse_n, storage=None or "fp32")
    # fp32 storage rounding is tiny but not zero: compare against an exact-storage run
    # through the same cycles by checking the residual histories instead
    lap = O.laplacian(n, n, p_ref)
    r_ref = np.max(np.abs(div - lap))
    lap2 = O.laplacian(n, n, p)
    r = np.max(np.abs(div - lap2))
    assert abs(r - r_ref) <= 1e-3 * r_ref + 1e-7 * np.max(np.abs(div))
    assert rmax[0] == pytest.approx(np.max(np.abs(div)), rel=1e-15)
    assert np.all(np.diff(np.log(rmax)) < -1.0)   # each step gains more than a factor e


@pytest.mark.parametrize("n", [64, 48, 50, 16, 6])
def test_poisson_dft_is_the_exact_zero_mean_inverse(O, n):
    # L x = r - mean(r) to rounding, x of zero mean, for any n (the exact coarsest-level
    # solve); equal to the FFT projection where that applies; a dropped factor of the
    # eigenvalue, a wrong sign or a kept mean each fail it
    rng = np.random.default_rng(n)
    r = rng.standard_normal(n * n)
    x = O.poisson_dft(n, r)
    lap = O.laplacian(n, n, x)
    assert np.max(np.abs(lap - (r - r.mean()))) < 1e-11 * np.max(np.abs(r))
    assert abs(x.mean()) < 1e-13 * np.max(np.abs(x))
    if n & (n - 1) == 0:
        assert np.max(np.abs(x - O.poisson_fft(n, n, r))) < 1e-12 * np.max(np.abs(x))
    for bad in (2 * x, -x, x + 0.1):
        assert np.max(np.abs(O.laplacian(n, n, bad) - (r - r.mean()))) > 1e-3 or abs(bad.mean()) > 1e-3


@pytest.mark.parametrize("n", [256, 96, 64])
def test_exact_coarse_refinement_converges_faster_than_sweeps(O, n):
    # the hierarchy stopped at its first level <= 64 and solved exactly: each V-cycle
    # contracts at least as fast as the reference's cycle down to 4^2; a grid of at most
    # 64^2 is solved in one refinement step
    div = _zero_mean_div(O, n, 5)
    _, r_ex = O.mg_refine(n, div, 4, min_size=4, coarse_exact=True)
    _, r_sw = O.mg_refine(n, div, 4, min_size=4)
    assert r_ex[0] == r_sw[0]
    if n <= 64:
        # exact but for the FP32 storage of the level-0 residual (2^-24 relative)
        assert r_ex[1] < 2e-7 * r_ex[0]
        assert r_sw[1] > 1e-3 * r_sw[0]
    else:
        assert np.all(r_ex[1:] <= r_sw[1:] * 1.02)
        assert np.all(np.diff(np.log(r_ex)) < -1.0)
    with pytest.raises(ValueError):
        O.mg_refine(260, _zero_mean_div(O, 260, 1), 1, min_size=130, coarse_exact=True)


def test_bf16_refinement_still_converges_to_the_gate(O):
    # BF16-stored corrections: 8-bit mantissa per step, FP64 residual; the iteration
    # keeps contracting and reaches a 1e-6 gate on a unit-scale rhs
    n = 128
    div = _zero_mean_div(O, n, 9)
    div = div / np.max(np.abs(div))
    _, r32 = O.mg_refine(n, div, 8, min_size=4, storage="fp32")
    _, rbf = O.mg_refine(n, div, 10, min_size=4, storage="bf16")
    k32 = int(np.argmax(r32 <= 1e-6))
    k = int(np.argmax(rbf <= 1e-6))
    assert r32[k32] <= 1e-6 and rbf[k] <= 1e-6 and k <= k32 + 1
    # each BF16 step down to the gate still contracts by about 10x or more (FP32: 13-100x)
    assert np.all(rbf[1:k + 1] / rbf[:k] < 0.12)


# --------------------------------------------------- learned-guess weights format

def test_guess_blob_round_trip_and_reference_forward():
    import torch
    from fluidsim.guess_net import LAYERS, load_blob
    from guess_torch import build_model, export_blob, model_from_blob
    torch.manual_seed(0)
    m = build_model()
    blob = export_blob(m)
    layers = load_blob(blob)
    assert len(layers) == 15 and layers[-1][0].shape == (1, 16, 1, 1)
    for (ci, co), (w, b) in zip(LAYERS, layers):
        assert w.shape == (co, ci, 3, 3) and b.shape == (co,)
    m2 = model_from_blob(blob)
    x = torch.randn(1, 2, 64, 64)
    with torch.no_grad():
        assert torch.allclose(m(x), m2(x), atol=1e-6)
    # the network is periodic: a circular shift of the input shifts the output
    with torch.no_grad():
        y = m(x)
        ys = m(torch.roll(x, (8, 16), (2, 3)))
    assert torch.allclose(torch.roll(y, (8, 16), (2, 3)), ys, atol=1e-5)


def test_bf16_round_matches_the_oracle_rounding(O):
    from guess_torch import bf16_round
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(500) * 10.0 ** rng.integers(-5, 5, 500)).astype(np.float32)
    ours = bf16_round(x).astype(np.float64)
    ref = np.array([O.round_storage(float(v), "bf16") for v in x])
    assert np.array_equal(ours, ref)


# ------------------------------------------------ lid-driven cavity (walled box)

def test_cavity_neumann_poisson_single_mode_closed_form(O):
    n = 32
    h = 1.0 / n
    i = np.arange(n)
    for k, l in ((1, 0), (2, 3), (5, 7)):
        rhs = np.outer(np.cos(np.pi * l * (i + 0.5) / n), np.cos(np.pi * k * (i + 0.5) / n)).ravel()
        lam = -4 / h ** 2 * (np.sin(np.pi * k / (2 * n)) ** 2 + np.sin(np.pi * l / (2 * n)) ** 2)
        p = O.cav_poisson_exact(n, rhs)
        assert np.max(np.abs(p - rhs / lam)) < 1e-12 * np.max(np.abs(rhs / lam))


def test_cavity_poisson_inverts_the_neumann_laplacian(O):
    n = 48
    rhs = np.random.default_rng(2).standard_normal(n * n)
    p = O.cav_poisson_exact(n, rhs)
    lap = O.cav_laplacian(n, p)
    assert np.max(np.abs(lap - (rhs - rhs.mean()))) < 1e-9 * np.max(np.abs(rhs))
    # the Neumann Laplacian is symmetric and annihilates constants
    a, b = np.random.default_rng(3).standard_normal((2, n * n))
    assert abs(a @ O.cav_laplacian(n, b) - b @ O.cav_laplacian(n, a)) < 1e-9 * n * n * n * n
    assert np.max(np.abs(O.cav_laplacian(n, np.full(n * n, 2.5)))) == 0


def test_cavity_projection_is_divergence_free_and_keeps_walls(O):
    n = 40
    rng = np.random.default_rng(4)
    u, v = rng.standard_normal((2, n * n))
    u.reshape(n, n)[:, 0] = 0
    v.reshape(n, n)[0, :] = 0
    d = O.cav_divergence(n, u, v)
    p = O.cav_poisson_exact(n, d)
    u2, v2 = O.cav_subtract_gradient(n, u, v, p)
    assert np.max(np.abs(O.cav_divergence(n, u2, v2))) < 1e-9 * np.max(np.abs(d))
    assert np.all(u2.reshape(n, n)[:, 0] == 0) and np.all(v2.reshape(n, n)[0, :] == 0)
    # idempotent
    p2 = O.cav_poisson_exact(n, O.cav_divergence(n, u2, v2))
    u3, v3 = O.cav_subtract_gradient(n, u2, v2, p2)
    assert np.max(np.abs(u3 - u2)) < 1e-10 and np.max(np.abs(v3 - v2)) < 1e-10


def test_cavity_samplers_put_the_wall_values_on_the_walls(O):
    n, U = 16, 1.7
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal((2, n * n))
    u.reshape(n, n)[:, 0] = 0
    v.reshape(n, n)[0, :] = 0
    for x in (0.13, 0.5, 0.77):
        assert abs(O.cav_sample_u(n, U, u, x, 0.0)) < 1e-15           # no slip at the bottom
        assert abs(O.cav_sample_u(n, U, u, x, 1.0) - U) < 1e-14       # the lid
        assert abs(O.cav_sample_u(n, U, u, x, -0.3)) < 1e-15          # clamped into the box
    for y in (0.2, 0.6):
        assert abs(O.cav_sample_v(n, v, 0.0, y)) < 1e-15 and abs(O.cav_sample_v(n, v, 1.0, y)) < 1e-15
        assert abs(O.cav_sample_u(n, U, u, 0.0, y)) < 1e-15 and abs(O.cav_sample_u(n, U, u, 1.0, y)) < 1e-15
    # at a stored face the sample is the stored value
    assert O.cav_sample_u(n, U, u, 5 / n, 7.5 / n) == pytest.approx(u[7 * n + 5], abs=1e-14)


def test_cavity_couette_profile_is_steady_under_diffusion(O):
    # u = U y between the resting bottom wall and the lid is a steady solution away
    # from the side walls
    n, U = 32, 1.3
    y = (np.arange(n) + 0.5) / n
    u = np.tile((U * y)[:, None], (1, n)).ravel()
    u.reshape(n, n)[:, 0] = 0
    v = np.zeros(n * n)
    ou, ov = O.cav_diffuse(n, U, u, v, 0.01, 1e-3)
    inner = np.s_[:, 2:n - 1]
    assert np.max(np.abs(ou.reshape(n, n)[inner] - u.reshape(n, n)[inner])) < 1e-13
    assert np.all(ov == 0)


@pytest.mark.parametrize("scheme", ["semi-lagrangian", "maccormack"])
def test_cavity_advection_at_rest_and_bounded(O, scheme):
    n = 32
    z = np.zeros(n * n)
    ou, ov = O.cav_advect(scheme, n, 1.0, z, z, 0.01)
    assert np.all(ou == 0) and np.all(ov == 0)
    rng = np.random.default_rng(6)
    u, v = rng.uniform(-1, 1, (2, n * n))
    u.reshape(n, n)[:, 0] = 0
    v.reshape(n, n)[0, :] = 0
    ou, ov = O.cav_advect(scheme, n, 1.0, u, v, 0.02)
    # values come from the field, the lid ghosts (<= 2U + 1) or walls: bounded, walls kept
    assert np.all(ou.reshape(n, n)[:, 0] == 0) and np.all(ov.reshape(n, n)[0, :] == 0)
    assert np.max(np.abs(ou)) <= 3.0 + 1e-12 and np.max(np.abs(ov)) <= 1.0 + 1e-12



@pytest.mark.parametrize("scheme", ["semi-lagrangian", "maccormack"])
def test_cavity_advection_away_from_walls_equals_the_periodic_operator(O, scheme):
    # far from every wall a walled trace, sample and limiter read only interior faces,
    # so the cavity operator must give the periodic one's numbers on the unit box
    n, dt = 64, 0.01
    h = 1.0 / n
    i = np.arange(n)
    X, Y = np.meshgrid(i * h, (i + 0.5) * h)       # u faces
    Xv, Yv = np.meshgrid((i + 0.5) * h, i * h)     # v faces
    blob = lambda x, y: np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    u = (0.6 + 0.3 * blob(X, Y) + 0.05 * np.sin(9 * X) * np.cos(7 * Y)).reshape(-1)
    v = (-0.4 + 0.2 * blob(Xv, Yv) + 0.05 * np.cos(5 * Xv) * np.sin(11 * Yv)).reshape(-1)
    u.reshape(n, n)[:, 0] = 0
    v.reshape(n, n)[0, :] = 0
    cu, cv = O.cav_advect(scheme, n, 1.0, u, v, dt)
    pu, pv = O.advect_velocity(scheme, n, n, u, v, dt, lx=1.0, ly=1.0)
    inner = np.s_[8:n - 8, 8:n - 8]   # traces move < 0.02 cells; stencils reach 2 cells
    assert np.max(np.abs(cu.reshape(n, n)[inner] - pu.reshape(n, n)[inner])) < 1e-13
    assert np.max(np.abs(cv.reshape(n, n)[inner] - pv.reshape(n, n)[inner])) < 1e-13


def test_cavity_samples_take_the_wall_values_on_the_walls(O):
    # no slip: a sample on the lid is the lid speed, on the bottom wall 0, and v on a
    # side wall 0, whatever the interior holds (ghost 2U - u above the lid, -u below
    # the bottom, -v beyond a side wall)
    n, U = 16, 1.7
    rng = np.random.default_rng(3)
    u, v = rng.uniform(-1, 1, (2, n * n))
    u.reshape(n, n)[:, 0] = 0
    v.reshape(n, n)[0, :] = 0
    for i in range(1, n):
        x = i / n                                  # on a u face column
        assert abs(O.cav_sample_u(n, U, u, x, 1.0) - U) < 1e-15
        assert abs(O.cav_sample_u(n, U, u, x, 0.0)) < 1e-15
        y = i / n                                  # on a v face row
        assert abs(O.cav_sample_v(n, v, 0.0, y)) < 1e-15
        assert abs(O.cav_sample_v(n, v, 1.0, y)) < 1e-15
    # and inside, a sample on a face node is that face's value
    assert O.cav_sample_u(n, U, u, 5 / n, 7.5 / n) == u[7 * n + 5]
    assert O.cav_sample_v(n, v, 5.5 / n, 7 / n) == v[7 * n + 5]

def test_cavity_flow_spins_up_divergence_free(O):
    s = O.Cavity(32, U=1.0, nu=1e-2, disk=dict(x=0.5, y=0.4, r=0.12), passes=2)
    for _ in range(20):
        s.step(s.dt_cfl(0.5))
    d = O.cav_divergence(32, s.u, s.v)
    assert np.max(np.abs(d)) < 1e-9
    uu = s.u.reshape(32, 32)
    assert uu[-1, 1:-1].mean() > 0.2      # the lid drags the top row
    assert np.all(uu[:, 0] == 0)
