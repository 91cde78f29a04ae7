"""Pins of the vorticity-confinement oracle (oracle/confine_oracle.py,
VORTICITY_CONFINEMENT of src/gpu/shaders.js:437-467) by closed forms, on the CPU."""
import math

import numpy as np

import confine_oracle as C

N = 48
TH = 2 * math.pi / N               # phase step of the test profiles
HX, HY = 2 * math.pi / N, 3 * math.pi / N  # a 2 pi x 3 pi domain: dx != dy
H = HX
EPS, DT = 0.8, 0.01


def shear_u(A):
    y = (np.arange(N) + 0.5) * TH
    return np.broadcast_to((A * np.sin(y))[:, None], (N, N)).astype(np.float32)


def test_uniform_flow_is_unchanged():
    u = np.full((N, N), 1.25, np.float32)
    v = np.full((N, N), -0.5, np.float32)
    cu, cv = C.confine(u, v, H, H, EPS, DT)
    assert np.array_equal(cu, u.astype(np.float64)) and np.array_equal(cv, v.astype(np.float64))


def test_shear_in_u_pushes_u_along_the_gradient_of_its_vorticity_magnitude():
    # u_j = A sin((j + 1/2) t): the corner vorticity of row j is c_j = -(u_j - u_{j-1}) / dy
    # = -(2A/dy) cos(j t) sin(t/2), the same along the row, so the unit gradient of |w| is
    # (0, sign(|c_{j+1}| - |c_{j-1}|)) and the force eps dx dt (n_y w, 0)
    A = 1.7
    u = shear_u(A)
    z = np.zeros((N, N), np.float32)
    cu, cv = C.confine(u, z, HX, HY, EPS, DT)
    c = -(2 * A / np.float32(HY)) * np.cos(np.arange(N) * TH) * math.sin(TH / 2)
    s = np.float32(EPS) * np.float32(HX) * np.float32(DT)
    a = np.abs(c)
    sgn = np.sign(np.roll(a, -1) - np.roll(a, 1))
    want = (s * sgn * c)[:, None]
    ok = np.abs(np.roll(a, -1) - np.roll(a, 1)) > 1e-6 * a.max()  # rows where |w| is not at an extremum
    assert np.count_nonzero(ok) > N // 2
    assert np.allclose((cu - u.astype(np.float64))[ok], np.broadcast_to(want, (N, N))[ok], rtol=1e-5,
                       atol=1e-9 * np.max(np.abs(want)))
    assert np.array_equal(cv, z.astype(np.float64))


def test_shear_in_v_pushes_v_against_the_gradient_of_its_vorticity_magnitude():
    # v_i = B sin((i + 1/2) t): c_i = (v_i - v_{i-1}) / dx = (2B/dx) cos(i t) sin(t/2) along each
    # column; n = (sign(|c_{i+1}| - |c_{i-1}|), 0) and the force eps dx dt (0, -n_x w)
    B = -0.9
    x = (np.arange(N) + 0.5) * TH
    v = np.broadcast_to((B * np.sin(x))[None, :], (N, N)).astype(np.float32)
    z = np.zeros((N, N), np.float32)
    cu, cv = C.confine(z, v, HX, HY, EPS, DT)
    c = (2 * B / np.float32(HX)) * np.cos(np.arange(N) * TH) * math.sin(TH / 2)
    s = np.float32(EPS) * np.float32(HX) * np.float32(DT)
    a = np.abs(c)
    sgn = np.sign(np.roll(a, -1) - np.roll(a, 1))
    want = (-s * sgn * c)[None, :]
    ok = np.abs(np.roll(a, -1) - np.roll(a, 1)) > 1e-6 * a.max()
    assert np.count_nonzero(ok) > N // 2
    assert np.allclose((cv - v.astype(np.float64))[:, ok], np.broadcast_to(want, (N, N))[:, ok], rtol=1e-5,
                       atol=1e-9 * np.max(np.abs(want)))
    assert np.array_equal(cu, z.astype(np.float64))


def test_force_is_odd_and_of_degree_one_in_the_velocity():
    rng = np.random.default_rng(5)
    u = rng.standard_normal((N, N)).astype(np.float32)
    v = rng.standard_normal((N, N)).astype(np.float32)
    cu, cv = C.confine(u, v, H, H, EPS, DT)
    fu, fv = cu - u, cv - v
    cu2, cv2 = C.confine(-2 * u, -2 * v, H, H, EPS, DT)
    assert np.allclose(cu2 + 2 * u, -2 * fu, rtol=1e-9, atol=1e-12)
    assert np.allclose(cv2 + 2 * v, -2 * fv, rtol=1e-9, atol=1e-12)
    # and of size eps h dt |w| at most
    w = C.corner_vorticity(u, v, H, H)
    bound = np.float32(EPS) * np.float32(H) * np.float32(DT) * np.abs(w) * (1 + 1e-12)
    assert np.all(np.hypot(fu, fv) <= bound + 1e-15)
