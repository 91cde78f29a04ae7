"""render() and its bloom chain (RENDER, BLOOM_BRIGHT, BLOOM_BLUR,
src/gpu/shaders.js:677-843, chained as src/gpu/solver.js:771-823) on the GPU against
the f64 display oracle (oracle/render_oracle.py) on the GPU's own fields.

The kernels compute in FP32 and store rgba16float like the reference's textures; the
oracle computes in f64 and rounds to float16 between the passes. So a pixel may differ
by one float16 spacing where the two values straddle a rounding boundary: the bound
is 1 spacing for the image and 4 for the bloom (fed the GPU's own image, four blur
passes each rounding once)."""
import math

import numpy as np
import pytest

import render_oracle as R

pytestmark = pytest.mark.gpu
TWO_PI = 2 * math.pi


def spacings(got, want):
    """max |got - want| in float16 spacings of want (floored at the spacing of 2^-14)."""
    w = want.astype(np.float64)
    sp = np.maximum(np.spacing(np.abs(want).astype(np.float16)).astype(np.float64), 2.0 ** -24)
    return float(np.max(np.abs(got.astype(np.float64) - w) / sp))


@pytest.fixture(scope="module")
def fs():
    import fluidsim
    fluidsim.createDevice()
    return fluidsim


def turbulent_solver(fs, n, nx=None, ny=None):
    from fluidsim import initial
    if nx is None:
        g = fs.GPUFluidSolver(0, n=n, nu=1e-4, pressureSolver="multigrid", pressureIterations=8)
    else:
        g = fs.GPUFluidSolver(0, nx=nx, ny=ny, nu=1e-4, pressureSolver="jacobi", pressureIterations=20)
    if nx is None:
        u, v = initial.decaying_turbulence(n, seed=3)
    else:
        rng = np.random.default_rng(4)
        u, v = rng.standard_normal((2, nx * ny))
    g.uploadVelocity(np.asarray(u, np.float64), np.asarray(v, np.float64))
    return g


@pytest.mark.parametrize("mode,scale", [(0, 1.5), (0, 0.05), (1, 1.4)])
def test_render_and_bloom_match_oracle(fs, mode, scale):
    n = 128
    with turbulent_solver(fs, n) as g:
        for _ in range(3):
            g.stepCFL(0.5)
        g.step(0.01, splats=[dict(x=2.0, y=3.0, r=0.9, g=0.3, b=0.6, radius=0.6, isDye=True)])
        g.render(mode=mode, scale=scale, bloomThreshold=0.75, bloomPasses=2)
        img = g.readField("output")
        bloom = g.readField("bloom")
        vel = g.readVelocity()
        dye = g.readField("dye").reshape(n, n, 4)
    u = vel["u"].reshape(n, n)
    v = vel["v"].reshape(n, n)
    h = TWO_PI / n
    want = R.render(u, v, dye, h, h, mode, scale)
    assert img.shape == (n, n, 4) and bloom.shape == (n // 2, n // 2, 4)
    assert np.all(img[..., 3] == 1) and np.all(bloom[..., 3] == 1)
    assert spacings(img, want) <= 1.0
    # the picture is not trivial: both hues (mode 0) and the bloom's knee are exercised
    assert np.ptp(img[..., :3].astype(np.float64)) > 0.5
    wb = R.bloom_bright(img, 0.75)
    for _ in range(2):
        wb = R.bloom_blur(R.bloom_blur(wb, 0), 1)
    assert spacings(bloom, wb) <= 4.0
    assert np.max(bloom[..., :3].astype(np.float64)) > 0.01


def test_render_reads_the_state_and_changes_none_of_it(fs):
    n = 64
    with turbulent_solver(fs, n) as g:
        g.stepCFL(0.5)
        before = g.readVelocity()
        g.render()
        g.render(mode=1, bloomPasses=0)
        after = g.readVelocity()
        assert np.array_equal(before["u"], after["u"]) and np.array_equal(before["v"], after["v"])
        with pytest.raises(Exception):
            g.readField("dye")  # rendering allocates no dye
        g.stepCFL(0.5)


def test_no_bloom_pass_keeps_the_last_bloom(fs):
    n = 64
    with turbulent_solver(fs, n) as g:
        g.render(scale=1.5, bloomPasses=1)
        b1 = g.readField("bloom")
        g.stepCFL(0.5)
        g.render(mode=1, bloomPasses=0)
        assert np.array_equal(g.readField("bloom"), b1)
        img = g.readField("output")
        vel = g.readVelocity()
    h = TWO_PI / n
    want = R.render(vel["u"].reshape(n, n), vel["v"].reshape(n, n), None, h, h, 1, 1.5)
    assert spacings(img, want) <= 1.0


def test_fractional_pass_count_runs_as_the_reference_loop(fs):
    # for (k = 0; k < 1.5; k++) runs two passes
    n = 64
    with turbulent_solver(fs, n) as g:
        g.render(bloomPasses=1.5)
        b15 = g.readField("bloom")
        g.render(bloomPasses=2)
        assert np.array_equal(g.readField("bloom"), b15)


def test_render_of_a_rectangular_odd_grid(fs):
    # nx odd: the bright pass reads 2x2 blocks of the first 2 (nx // 2) columns
    nx, ny = 81, 50
    with turbulent_solver(fs, 0, nx=nx, ny=ny) as g:
        g.render(scale=0.8, bloomThreshold=0.3, bloomPasses=3)
        img = g.readField("output")
        bloom = g.readField("bloom")
        vel = g.readVelocity()
    want = R.render(vel["u"].reshape(ny, nx), vel["v"].reshape(ny, nx), None, TWO_PI / nx, TWO_PI / ny, 0, 0.8)
    assert spacings(img, want) <= 1.0
    wb = R.bloom_bright(img, 0.3)
    for _ in range(3):
        wb = R.bloom_blur(R.bloom_blur(wb, 0), 1)
    assert bloom.shape == (ny // 2, nx // 2, 4) and spacings(bloom, wb) <= 4.0


def test_render_before_any_render_read_is_an_error(fs):
    with turbulent_solver(fs, 64) as g:
        with pytest.raises(Exception):
            g.readField("output")
