"""Initial conditions of the benchmark workloads (host-side set-up, not timed).

taylor_green      src/core/analytic.js:38-72 at the MAC faces.
decaying_turbulence  seeded random-phase streamfunction on cell corners with
                  E(k) ~ k^4 exp(-2 (k/k0)^2); u = d(psi)/dy, v = -d(psi)/dx as
                  one-sided corner differences, so the field is discretely
                  divergence-free; scaled to unit RMS face speed.
"""
from __future__ import annotations

import math

import numpy as np

TWO_PI = 2.0 * math.pi
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """`count` outputs of splitmix64 (Steele, Lea, Flood 2014) as uint64."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def taylor_green(nx, ny, t=0.0, nu=0.0, lx=TWO_PI, ly=TWO_PI):
    dx, dy = lx / nx, ly / ny
    i = np.arange(nx)[None, :]
    j = np.arange(ny)[:, None]
    a = math.exp(-2 * nu * t)
    u = np.sin(i * dx) * np.cos((j + 0.5) * dy) * a
    v = -np.cos((i + 0.5) * dx) * np.sin(j * dy) * a
    return u.reshape(-1).copy(), v.reshape(-1).copy()


def decaying_turbulence(n, seed=1, k0=8.0, lx=TWO_PI):
    """Decaying-turbulence initial state: returns (u, v) float64 of n*n at the MAC faces."""
    h = lx / n
    kx = np.fft.fftfreq(n, d=1.0 / n)[None, :]
    ky = np.fft.fftfreq(n, d=1.0 / n)[:, None]
    k = np.sqrt(kx * kx + ky * ky)
    theta = TWO_PI * uniform01(seed, n * n).reshape(n, n)
    with np.errstate(divide="ignore", invalid="ignore"):
        E = k ** 4 * np.exp(-2.0 * (k / k0) ** 2)
        amp = np.sqrt(E / (math.pi * k ** 3))
    amp[0, 0] = 0.0
    spec = amp * np.exp(1j * theta)
    psi = np.real(np.fft.ifft2(spec)) * (n * n)   # psi at corners (i*h, j*h)
    u = (np.roll(psi, -1, axis=0) - psi) / h      # (psi(i, j+1) - psi(i, j)) / h
    v = -(np.roll(psi, -1, axis=1) - psi) / h     # -(psi(i+1, j) - psi(i, j)) / h
    rms = math.sqrt(np.mean(u * u) + np.mean(v * v))
    return (u / rms).reshape(-1).copy(), (v / rms).reshape(-1).copy()
