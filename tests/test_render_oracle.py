"""Pins of the display oracle (oracle/render_oracle.py, RENDER / BLOOM_BRIGHT /
BLOOM_BLUR of src/gpu/shaders.js:677-843) by closed forms, on the CPU.

Each check is a case whose answer follows from the shader's definition without
running its full formula: a fluid at rest, a uniform flow, shear flows whose discrete
vorticity is known in closed form, a uniform dye, a NaN; box averages, the soft knee,
the blur's normalisation, its taps on a point source and at the clamped edge.
"""
import math

import numpy as np
import pytest

import render_oracle as R

N = 32
H = 2 * math.pi / N


def close16(got, want, ulps=1):
    """got (float16) within `ulps` float16 spacings of the float64 value want."""
    want16 = np.asarray(want, np.float64)
    tol = ulps * np.spacing(np.abs(want16).astype(np.float16)).astype(np.float64) + 1e-12
    return np.all(np.abs(got.astype(np.float64) - want16) <= tol)


def f32(x):
    return np.float64(np.float32(x))


VOID = np.array([f32(0.004), f32(0.003), f32(0.016)])
WHITE = np.array([f32(0.55), f32(0.78), f32(1.0)])
MAGENTA = np.array([f32(1.0), f32(0.10), f32(0.62)])
CYAN = np.array([f32(0.10), f32(0.92), f32(1.0)])


@pytest.mark.parametrize("mode,lift", [(0, 0.30), (1, 0.35)])
def test_fluid_at_rest_renders_the_void_colour(mode, lift):
    z = np.zeros((N, N), np.float32)
    img = R.render(z, z, None, H, H, mode=mode, scale=1.5)
    assert img.dtype == np.float16 and img.shape == (N, N, 4)
    assert np.all(img[..., 3] == 1)
    assert close16(img[..., :3], np.broadcast_to(VOID * f32(lift), (N, N, 3)))


def test_uniform_flow_adds_only_the_white_hot_speed_term():
    U, scale = 2.0, 1.5
    u = np.full((N, N), U, np.float32)
    z = np.zeros((N, N), np.float32)
    sp = math.tanh(U * f32(scale) * f32(0.45))
    want = VOID * f32(0.30) + WHITE * sp ** 5 * f32(0.85)
    assert close16(R.render(u, z, None, H, H, 0, scale)[..., :3], np.broadcast_to(want, (N, N, 3)))


def test_discrete_vorticity_of_shear_flows_is_the_closed_form():
    A, B = 1.3, -0.7
    y = (np.arange(N) + 0.5) * H
    u = np.broadcast_to((A * np.sin(y))[:, None], (N, N)).astype(np.float32)
    v = np.broadcast_to((B * np.sin(y))[None, :], (N, N)).astype(np.float32)  # v depends on x
    z = np.zeros((N, N), np.float32)
    wu = R.vorticity_at_centres(u, z, H, H)
    wv = R.vorticity_at_centres(z, v, H, H)
    assert np.allclose(wu, (-A * np.cos(y) * math.sin(H) / H)[:, None], rtol=0, atol=1e-5)
    assert np.allclose(wv, (B * np.cos(y) * math.sin(H) / H)[None, :], rtol=0, atol=1e-5)


def test_saturated_vorticity_takes_its_sign_hue_lifted_to_4_9():
    # u = A sin y: w = -A cos(y) sin(h)/h, so magenta (w > 0) where cos y < 0, cyan where
    # cos y > 0; with |w| scale and |u| scale beyond the tanh clamp both terms saturate
    A, scale = 100.0, 10.0
    y = (np.arange(N) + 0.5) * H
    u = np.broadcast_to((A * np.sin(y))[:, None], (N, N)).astype(np.float32)
    z = np.zeros((N, N), np.float32)
    img = R.render(u, z, None, H, H, 0, scale)
    rows = np.nonzero((np.abs(np.sin(y)) > 0.5) & (np.abs(np.cos(y)) > 0.5))[0]
    assert rows.size >= 8
    for j in rows:
        hue = MAGENTA if math.cos(y[j]) < 0 else CYAN
        want = hue * (f32(0.30) + f32(4.6)) + WHITE * f32(0.85)
        assert close16(img[j, :, :3], np.broadcast_to(want, (N, 3)), ulps=2), j


def test_dye_lifts_by_its_luminance_and_adds_a_fifth_of_its_colour():
    z = np.zeros((N, N), np.float32)
    d = np.array([0.6, 0.2, 0.9, 0.5], np.float32)
    dye = np.broadcast_to(d, (N, N, 4)).astype(np.float32)
    c0 = VOID * f32(0.30)
    lum = f32(d[0]) * f32(0.2126) + f32(d[1]) * f32(0.7152) + f32(d[2]) * f32(0.0722)
    want = c0 + c0 * lum * f32(1.5) + d[:3].astype(np.float64) * f32(0.20)
    assert close16(R.render(z, z, dye, H, H, 0, 1.5)[..., :3], np.broadcast_to(want, (N, N, 3)))


def test_non_finite_input_paints_black():
    u = np.zeros((N, N), np.float32)
    u[5, 7] = np.nan
    img = R.render(u, np.zeros_like(u), None, H, H, 0, 1.5)
    assert np.all(img[5, 7] == np.array([0, 0, 0, 1], np.float16))
    assert np.all(np.isfinite(img.astype(np.float64)))


def test_bright_pass_box_average_and_soft_knee():
    thr = 0.75
    for a, b in [(3.0, 1.0), (0.2, 0.1), (0.9, 0.7)]:
        img = np.ones((8, 8, 4))
        img[..., :3] = np.where((np.add.outer(np.arange(8), np.arange(8)) % 2 == 0)[..., None], a, b)
        img = img.astype(np.float16)
        out = R.bloom_bright(img, thr)
        assert out.shape == (4, 4, 4)
        m = (np.float64(np.float16(a)) + np.float64(np.float16(b))) / 2
        lum = m * (f32(0.2126) + f32(0.7152) + f32(0.0722))
        t = min(max((lum - f32(thr)) / (f32(np.float32(thr) + np.float32(0.6)) - f32(thr)), 0.0), 1.0)
        want = m * t * t * (3 - 2 * t)
        assert close16(out[..., :3], np.full((4, 4, 3), want)), (a, b)


def test_blur_is_normalised_and_its_taps_are_the_weights():
    W = [f32(w) for w in (0.1964, 0.1747, 0.1210, 0.0656, 0.0278, 0.0092, 0.0024)]
    norm = W[0] + 2 * sum(W[1:])
    const = np.ones((10, 40, 4))
    const[..., :3] = 0.37
    const = const.astype(np.float16)
    for axis in (0, 1):
        assert close16(R.bloom_blur(const, axis)[..., :3], np.full((10, 40, 3), np.float64(np.float16(0.37))))
    pt = np.zeros((9, 40, 4), np.float16)
    pt[..., 3] = 1
    pt[4, 20, :3] = 1
    row = R.bloom_blur(pt, 0)[4, :, 0]
    want = np.zeros(40)
    for k in range(-6, 7):
        want[20 + k] = W[abs(k)] / norm
    assert close16(row, want)
    assert np.all(R.bloom_blur(pt, 0)[3, :, :3] == 0)  # x blur stays in its row
    col = R.bloom_blur(pt, 1)[:, 20, 1]
    assert close16(col, np.array([W[abs(j - 4)] / norm if abs(j - 4) <= 6 else 0 for j in range(9)]))


def test_blur_clamps_at_the_image_edge():
    W = [f32(w) for w in (0.1964, 0.1747, 0.1210, 0.0656, 0.0278, 0.0092, 0.0024)]
    norm = W[0] + 2 * sum(W[1:])
    img = np.zeros((3, 20, 4), np.float16)
    img[:, 0, :3] = 1  # a source on the left edge: taps left of it read it again
    row = R.bloom_blur(img, 0)[1, :, 2]
    want = np.zeros(20)
    want[0] = (W[0] + sum(W[1:])) / norm
    for m in range(1, 7):
        want[m] = sum(W[m:]) / norm
    assert close16(row, want)


def test_no_bloom_pass_keeps_the_previous_bloom():
    z = np.zeros((8, 8), np.float32)
    prev = np.full((4, 4, 4), 0.5, np.float16)
    img, bl = R.render_and_bloom(z, z, None, H, H, bloom_passes=0, bloom=prev)
    assert bl is prev and img.shape == (8, 8, 4)
