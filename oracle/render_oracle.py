"""f64 restatement of the reference's display passes: RENDER, BLOOM_BRIGHT and
BLOOM_BLUR (src/gpu/shaders.js:677-843) as src/gpu/solver.js:771-823 chains them.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's render
kernels. The product library in fluid-sim_amd/ never imports this module.

Arithmetic is float64 on the float32 inputs, with the shader's parameters and literals
rounded to float32 first (they are f32 uniforms and f32 literals in WGSL). Each pass
stores rgba16float, so its output is rounded to float16 (round to nearest even) before
the next pass reads it, as the textures between the passes hold it.

Pinned in tests/test_render_oracle.py by closed forms: the colour of a fluid at rest,
of a uniform flow (the speed term alone), of a shear flow whose discrete vorticity is
known in closed form (hue by sign, the lift at saturation), of a uniform dye; the NaN
guard; the bright pass's box average and knee; the blur's normalisation, its tap
weights on a point source and its clamped edge.
"""
from __future__ import annotations

import numpy as np


def _f(x):
    """A WGSL f32 literal or uniform: the value rounded to float32, used in float64."""
    return np.float64(np.float32(x))


MAGENTA = np.array([_f(1.00), _f(0.10), _f(0.62)])
CYAN = np.array([_f(0.10), _f(0.92), _f(1.00)])
VIOLET = np.array([_f(0.16), _f(0.03), _f(0.42)])
VOID = np.array([_f(0.004), _f(0.003), _f(0.016)])
WHITE_HOT = np.array([_f(0.55), _f(0.78), _f(1.0)])
LUMA = np.array([_f(0.2126), _f(0.7152), _f(0.0722)])
BLUR_W = np.array([_f(w) for w in (0.1964, 0.1747, 0.1210, 0.0656, 0.0278, 0.0092, 0.0024)])


def smoothstep(e0, e1, x):
    """WGSL smoothstep: t^2 (3 - 2t), t = clamp((x - e0) / (e1 - e0), 0, 1)."""
    t = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def mix(a, b, t):
    """WGSL mix(a, b, t) = a (1 - t) + b t, t broadcast over the colour axis."""
    t = np.asarray(t)[..., None]
    return a * (1.0 - t) + b * t


def _lum(c):
    return c[..., 0] * LUMA[0] + c[..., 1] * LUMA[1] + c[..., 2] * LUMA[2]


def vorticity_at_centres(u, v, dx, dy):
    """RENDER's w (src/gpu/shaders.js:682-686, :694): v_x - u_y at the four corners of
    cell (i, j), (i..i+1, j..j+1), averaged; periodic indices. u, v: (ny, nx)."""
    dx, dy = _f(dx), _f(dy)
    U = u.astype(np.float64)
    V = v.astype(np.float64)

    def curl(si, sj):  # curlAt(i + si, j + sj) for every (i, j)
        Vc = np.roll(V, (-sj, -si), (0, 1))
        Vl = np.roll(V, (-sj, -si + 1), (0, 1))
        Uc = np.roll(U, (-sj, -si), (0, 1))
        Ub = np.roll(U, (-sj + 1, -si), (0, 1))
        return (Vc - Vl) / dx - (Uc - Ub) / dy

    return 0.25 * (curl(0, 0) + curl(1, 0) + curl(0, 1) + curl(1, 1))


def render(u, v, dye, dx, dy, mode=0.0, scale=1.5):
    """RENDER (src/gpu/shaders.js:677-780) of the (ny, nx) face fields u, v and the
    (ny, nx, 4) dye (None: no dye). Returns (ny, nx, 4) float16, alpha 1."""
    a = _f(scale)
    w = vorticity_at_centres(u, v, dx, dy)
    cu = u.astype(np.float64)
    cv = v.astype(np.float64)
    speed = np.sqrt(cu * cu + cv * cv)
    if np.float32(mode) < np.float32(0.5):
        sig = np.tanh(np.clip(w * a, -12.0, 12.0))
        mag = np.abs(sig)
        hue = np.where((sig >= 0.0)[..., None], MAGENTA, CYAN)
        ramp = mix(mix(VOID, VIOLET, smoothstep(0.0, _f(0.22), mag)), hue, smoothstep(_f(0.12), _f(0.62), mag))
        col = ramp * (_f(0.30) + _f(4.6) * mag * mag)[..., None]
    else:
        t = np.tanh(np.clip(speed * a, -12.0, 12.0))
        hue = mix(VIOLET, CYAN, smoothstep(_f(0.15), _f(0.95), t))
        col = mix(VOID, hue, smoothstep(0.0, _f(0.6), t)) * (_f(0.35) + _f(2.9) * t * t)[..., None]
    sp = np.tanh(np.clip(speed * a * _f(0.45), 0.0, 12.0))
    sp2 = sp * sp
    col = col + WHITE_HOT * ((sp2 * sp2 * sp) * _f(0.85))[..., None]
    d = np.zeros(u.shape + (3,)) if dye is None else dye[..., :3].astype(np.float64)
    col = col + (col * _lum(d)[..., None]) * _f(1.5) + d * _f(0.20)
    col = np.where(np.isnan(col), 0.0, col)
    out = np.ones(u.shape + (4,))
    out[..., :3] = col
    return out.astype(np.float16)


def bloom_bright(img, threshold):
    """BLOOM_BRIGHT (src/gpu/shaders.js:783-812): 2x2 box average of the (ny, nx, 4)
    float16 render at (ny//2, nx//2), scaled by smoothstep(thr, thr + 0.6, luminance)."""
    h, w = img.shape[0] >> 1, img.shape[1] >> 1
    c = img[: 2 * h, : 2 * w, :3].astype(np.float64)
    acc = ((c[0::2, 0::2] + c[0::2, 1::2]) + c[1::2, 0::2]) + c[1::2, 1::2]
    acc = acc * 0.25
    thr = np.float32(threshold)
    knee = smoothstep(np.float64(thr), np.float64(thr + np.float32(0.6)), _lum(acc))
    out = np.ones((h, w, 4))
    out[..., :3] = acc * knee[..., None]
    return out.astype(np.float16)


def bloom_blur(img, axis):
    """BLOOM_BLUR (src/gpu/shaders.js:816-843) along x (axis 0) or y (axis 1): the
    13-tap Gaussian normalised by its weights, taps clamped to the image."""
    c = img[..., :3].astype(np.float64)
    h, w = c.shape[:2]
    ii = np.arange(w)
    jj = np.arange(h)

    def tap(k):
        if axis == 0:
            return c[:, np.clip(ii + k, 0, w - 1)]
        return c[np.clip(jj + k, 0, h - 1), :]

    acc = c * BLUR_W[0]
    norm = BLUR_W[0]
    for k in range(1, 7):
        acc = acc + (tap(k) + tap(-k)) * BLUR_W[k]
        norm = norm + 2.0 * BLUR_W[k]
    out = np.ones((h, w, 4))
    out[..., :3] = acc / norm
    return out.astype(np.float16)


def render_and_bloom(u, v, dye, dx, dy, mode=0.0, scale=1.5, bloom_threshold=0.75, bloom_passes=2, bloom=None):
    """render() (src/gpu/solver.js:779-823): the image, then with passes > 0 the bright
    pass and `passes` rounds of the x and y blur. `bloom` is the bloom before the call
    (returned unchanged when there are no passes). Returns (image, bloom)."""
    img = render(u, v, dye, dx, dy, mode, scale)
    if bloom_passes > 0:
        bloom = bloom_bright(img, bloom_threshold)
        k = 0
        while k < bloom_passes:
            for axis in (0, 1):
                bloom = bloom_blur(bloom, axis)
            k += 1
    return img, bloom
