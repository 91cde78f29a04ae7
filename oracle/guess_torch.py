"""The learned guess restated in torch (float64 on the CPU), as a test oracle: the
network of fluid-sim_amd/fluidsim/guess_net.py built from torch.nn layers, the
weights-blob export of such a model, and the library's BF16 inference (weights of
the 3x3 layers after the first and every activation rounded to BF16) evaluated in
float64. Test infrastructure only: tests import it, the package never does.
"""
from __future__ import annotations

import struct

import numpy as np

from fluidsim.guess_net import LAYERS, MAGIC, load_blob


def bf16_round(a):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def build_model():
    import torch.nn as nn
    import torch.nn.functional as F

    class GuessNet(nn.Module):
        def __init__(self):
            super().__init__()
            self.convs = nn.ModuleList([nn.Conv2d(ci, co, 3) for ci, co in LAYERS])
            self.out = nn.Conv2d(16, 1, 1)

        def forward(self, x, act=None):
            import torch

            def c(i, t):
                y = F.relu(self.convs[i](F.pad(t, (1, 1, 1, 1), mode="circular")))
                return act(y) if act else y

            def pool(t):
                y = F.avg_pool2d(t, 2)
                return act(y) if act else y

            def up(t):
                return F.interpolate(t, scale_factor=2, mode="nearest")

            s0 = c(1, c(0, x))
            s1 = c(3, c(2, pool(s0)))
            s2 = c(5, c(4, pool(s1)))
            b3 = c(7, c(6, pool(s2)))
            d2 = c(9, c(8, torch.cat([up(b3), s2], 1)))
            d1 = c(11, c(10, torch.cat([up(d2), s1], 1)))
            d0 = c(13, c(12, torch.cat([up(d1), s0], 1)))
            return self.out(d0)

    return GuessNet()


def export_blob(model) -> bytes:
    """Weights blob: "FSGN", version 1, 15 layers of (cin, cout, k, w[cout][k*k][cin], b[cout])."""
    out = [MAGIC, struct.pack("<II", 1, 15)]
    mods = list(model.convs) + [model.out]
    for m in mods:
        w = m.weight.detach().double().cpu().numpy()  # [cout, cin, ky, kx]
        b = m.bias.detach().double().cpu().numpy()
        cout, cin, k, _ = w.shape
        out.append(struct.pack("<III", cin, cout, k))
        out.append(np.ascontiguousarray(w.transpose(0, 2, 3, 1), np.float32).tobytes())
        out.append(np.ascontiguousarray(b, np.float32).tobytes())
    return b"".join(out)


def model_from_blob(blob: bytes, dtype=None):
    import torch
    model = build_model()
    mods = list(model.convs) + [model.out]
    for m, (w, b) in zip(mods, load_blob(blob)):
        m.weight.data = torch.from_numpy(w.copy())
        m.bias.data = torch.from_numpy(b.copy())
    if dtype is not None:
        model = model.to(dtype)
    return model


def reference_guess(blob: bytes, r, p_prev, n, h):
    """The library's guess restated on the CPU in float64: weights of the 3x3 layers
    after the first rounded to BF16, every activation rounded to BF16 (as the HIP
    kernels store them), the first and output layers in full precision.
    Returns p_prev + max|r| h^2 o."""
    import torch
    model = model_from_blob(blob, torch.float64)
    with torch.no_grad():
        for i, m in enumerate(model.convs):
            if i > 0:
                m.weight.data = torch.from_numpy(bf16_round(m.weight.data.numpy()).astype(np.float64))
        sr = float(np.max(np.abs(r)))
        sp = float(np.max(np.abs(p_prev)))
        x0 = np.asarray(r, np.float32) * np.float32(1.0 / sr) if sr > 0 else np.zeros(n * n, np.float32)
        x1 = (np.asarray(p_prev, np.float64).astype(np.float32) * np.float32(1.0 / sp)) if sp > 0 \
            else np.zeros(n * n, np.float32)
        x = torch.from_numpy(np.stack([x0, x1]).reshape(1, 2, n, n).astype(np.float64))

        def act(t):
            return torch.from_numpy(bf16_round(t.numpy().astype(np.float32)).astype(np.float64))

        o = model(x, act).numpy().reshape(-1)
    return np.asarray(p_prev, np.float64) + sr * h * h * o
