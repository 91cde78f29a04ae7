"""Forward pass, loss and gradient of the learned pressure guess, in float64 numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
training kernels (fs_trainer_*). The product library never imports it.

The network (channels 16/32/64/64, periodic 3x3 convolutions + ReLU, 2x2 average
pooling down, nearest-neighbour upsampling and [upsampled, skip] concatenation up,
1x1 output) and its input/target are those of fluid-sim_amd/fluidsim/guess_net.py
and train_guess.py; the reference has no learned component (its projection starts
from p = 0 every step, src/gpu/solver.js:669-672), so the network is pinned by the
mathematics instead: central finite differences of the loss (brute force) and an
independent implementation of the same network (torch autograd), both in the tests.

Tensors are NHWC: [batch][row][column][channel]. Parameters are one flat vector in
the weights-blob order: per layer w[cout][k*k][cin] (tap = 3 (dy + 1) + dx + 1),
then b[cout].

With bf16=True the library's storage points are reproduced: weights of the 3x3
layers after the first, every activation, the pooled inputs and every
pre-activation gradient are rounded to BF16 (nearest, ties to even, through
float32 as the library rounds its FP32 accumulators); the skip gradients stay
unrounded (the library keeps them in FP32).
"""
from __future__ import annotations

import numpy as np

# (cin, cout) of the 3x3 layers, then the 1x1 output 16 -> 1
LAYERS = [(2, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 64), (64, 64),
          (128, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16), (16, 1)]
KSIZE = [3] * 14 + [1]
# input of each 3x3 layer: ("x" | activation index, mode, skip activation index)
#   mode "direct": the tensor; "pool": its 2x2 average; "upcat": [up(tensor), skip]
INPUTS = [("x", "direct", None), (0, "direct", None), (1, "pool", None), (2, "direct", None),
          (3, "pool", None), (4, "direct", None), (5, "pool", None), (6, "direct", None),
          (7, "upcat", 5), (8, "direct", None), (9, "upcat", 3), (10, "direct", None),
          (11, "upcat", 1), (12, "direct", None)]


def layout():
    """[(w_offset, b_offset, cin, cout, k)] per layer, and the parameter count."""
    out, off = [], 0
    for (ci, co), k in zip(LAYERS, KSIZE):
        out.append((off, off + co * k * k * ci, ci, co, k))
        off += co * k * k * ci + co
    return out, off


def unpack(theta):
    """Flat parameters -> [(w[cout][k*k][cin], b[cout])] (views)."""
    lay, total = layout()
    theta = np.asarray(theta)
    assert theta.shape == (total,)
    return [(theta[w:b].reshape(co, k * k, ci), theta[b:b + co]) for w, b, ci, co, k in lay]


def init_params(seed):
    """Uniform +-1/sqrt(fan_in) weights and biases (fan_in = cin k^2), seeded."""
    rng = np.random.default_rng(seed)
    parts = []
    for (ci, co), k in zip(LAYERS, KSIZE):
        bound = 1.0 / np.sqrt(ci * k * k)
        parts.append(rng.uniform(-bound, bound, co * k * k * ci))
        parts.append(rng.uniform(-bound, bound, co))
    return np.concatenate(parts).astype(np.float32).astype(np.float64)


def bf16(a):
    """Round to the nearest BF16 (ties to even) through float32, returned as float64."""
    f = np.ascontiguousarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(a):
    return a


# ------------------------------------------------------------ inputs and targets

def laplacian(p, h):
    """5-point periodic Laplacian of p [.., n, n] (rows, columns), in the order the
    library evaluates it: ((east + west) + (north + south)) - 4 p, over h^2."""
    e = np.roll(p, -1, -1)
    w = np.roll(p, 1, -1)
    nn = np.roll(p, -1, -2)
    s = np.roll(p, 1, -2)
    return (((e + w) + (nn + s)) - 4.0 * p) * (1.0 / (h * h))


def make_inputs(p, pp, h):
    """Network input x [B][n][n][2] (float32 values, as the library forms them), the
    target t [B][n][n] and the scales (max|r|, max|p_prev|) per image, from the pairs
    (p_prev, p) = (pp[b], p[b]) of FP64 pressures [B][n][n]."""
    r = laplacian(p, h) - laplacian(pp, h)
    B = p.shape[0]
    x = np.zeros(p.shape + (2,), np.float32)
    t = np.zeros(p.shape, np.float32)
    scales = np.zeros((B, 2))
    for b in range(B):
        sr = float(np.max(np.abs(r[b])))
        sp = float(np.max(np.abs(pp[b])))
        scales[b] = sr, sp
        ir = np.float32(1.0 / sr) if sr > 0 else np.float32(0)
        ip = np.float32(1.0 / sp) if sp > 0 else np.float32(0)
        x[b, ..., 0] = r[b].astype(np.float32) * ir
        x[b, ..., 1] = pp[b].astype(np.float32) * ip
        if sr > 0:
            t[b] = ((p[b] - pp[b]) * (1.0 / (sr * h * h))).astype(np.float32)
    return x.astype(np.float64), t.astype(np.float64), scales


# ------------------------------------------------------------ layer operators

_TAPS = [(dy, dx) for dy in range(3) for dx in range(3)]


def im2col(x):
    """x [B][m][m][C] -> cols [B][m][m][9][C], cols[p][tap] = x[p + (dy - 1, dx - 1)]."""
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)), mode="wrap")
    m = x.shape[1]
    return np.stack([xp[:, dy:dy + m, dx:dx + m, :] for dy, dx in _TAPS], axis=3)


def conv3x3(x, w, b):
    """Periodic 3x3 correlation: y[p][co] = b[co] + sum_tap,ci w[co][tap][ci] x[p + off(tap)][ci]."""
    B, m, _, c = x.shape
    cols = im2col(x).reshape(B * m * m, 9 * c)
    return (cols @ w.reshape(w.shape[0], -1).T + b).reshape(B, m, m, -1)


def conv3x3_backward(gz, x, w):
    """Gradients of conv3x3 for the pre-activation gradient gz: (dx, dw, db)."""
    B, m, _, c = x.shape
    co = w.shape[0]
    g = gz.reshape(B * m * m, co)
    cols = im2col(x).reshape(B * m * m, 9 * c)
    dw = (g.T @ cols).reshape(co, 9, c)
    db = g.sum(0)
    dcols = (g @ w.reshape(co, -1)).reshape(B, m, m, 9, c)
    dx = np.zeros_like(x)
    for tap, (dy, dxx) in enumerate(_TAPS):
        dx += np.roll(dcols[:, :, :, tap, :], shift=(dy - 1, dxx - 1), axis=(1, 2))
    return dx, dw, db


def pool(x):
    return 0.25 * (((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + x[:, 1::2, 0::2]) + x[:, 1::2, 1::2])


def unpool_grad(g):
    """Gradient of pool: each fine pixel gets a quarter of its coarse pixel's gradient."""
    return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_grad(g):
    """Gradient of up: the sum over each 2x2 block."""
    return (g[:, 0::2, 0::2] + g[:, 0::2, 1::2]) + (g[:, 1::2, 0::2] + g[:, 1::2, 1::2])


# ------------------------------------------------------------ network

def forward(theta, x, bf16_storage=False):
    """Activations (after ReLU) of the 14 3x3 layers, the formed input of each layer
    and the output o [B][n][n]."""
    R = bf16 if bf16_storage else _ident
    params = unpack(theta)
    acts, ins = [], []
    for l in range(14):
        src, mode, skip = INPUTS[l]
        t = x if src == "x" else acts[src]
        if mode == "pool":
            t = R(pool(t))
        elif mode == "upcat":
            t = np.concatenate([up(t), acts[skip]], axis=3)
        ins.append(t)
        w, b = params[l]
        if l > 0:
            w = R(w)
        acts.append(R(np.maximum(conv3x3(t, w, b), 0.0)))
    wo, bo = params[14]
    o = acts[13] @ wo.reshape(-1) + bo[0]
    return acts, ins, o


def loss_and_gradient(theta, x, t, bf16_storage=False):
    """Loss mean((o - t)^2) over the batch and the pixels and its gradient (flat,
    blob order), with the library's storage points when bf16_storage."""
    R = bf16 if bf16_storage else _ident
    params = unpack(theta)
    lay, total = layout()
    grad = np.zeros(total)
    acts, ins, o = forward(theta, x, bf16_storage)
    e = o - t
    count = e.size
    loss = float(np.mean(e * e))
    dO = e * (2.0 / count)
    wo, _ = params[14]
    w_off, b_off = lay[14][0], lay[14][1]
    grad[w_off:b_off] = np.einsum("bijc,bij->c", acts[13], dO)
    grad[b_off] = dO.sum()
    gz = [None] * 14
    gz[13] = np.where(acts[13] > 0, R(dO[..., None] * wo.reshape(-1)), 0.0)
    skip_grad = {}
    for l in range(13, -1, -1):
        w, _ = params[l]
        if l > 0:
            w = R(w)
        gx, dw, db = conv3x3_backward(gz[l], ins[l], w)
        w_off, b_off = lay[l][0], lay[l][1]
        grad[w_off:b_off] = dw.reshape(-1)
        grad[b_off:b_off + db.size] = db
        if l == 0:
            break
        src, mode, skip = INPUTS[l]
        if mode == "direct":
            gz[src] = np.where(acts[src] > 0, R(gx), 0.0)
        elif mode == "upcat":
            c2 = acts[src].shape[3]
            gz[src] = np.where(acts[src] > 0, R(up_grad(gx[..., :c2])), 0.0)
            skip_grad[skip] = gx[..., c2:]
        else:  # pool: the skip gradient of the same tensor was stored by the decoder
            g = unpool_grad(gx) + skip_grad.get(src, 0.0)
            gz[src] = np.where(acts[src] > 0, R(g), 0.0)
    return loss, grad


def loss_only(theta, x, t):
    _, _, o = forward(theta, x, False)
    e = o - t
    return float(np.mean(e * e))


def adam(theta, grad, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float32 (PyTorch's formulation); returns (theta, m1, m2)."""
    f = np.float32
    g = grad.astype(f)
    m1 = f(b1) * m1.astype(f) + f(1 - b1) * g
    m2 = f(b2) * m2.astype(f) + f(1 - b2) * (g * g)
    bc1 = 1.0 - b1 ** step
    bc2 = 1.0 - b2 ** step
    theta = theta.astype(f) - f(lr / bc1) * (m1 / (np.sqrt(m2) * f(1.0 / np.sqrt(bc2)) + f(eps)))
    return theta, m1, m2


def to_blob(theta):
    """The weights blob of flat parameters (float32)."""
    import struct
    out = [b"FSGN", struct.pack("<II", 1, 15)]
    lay, _ = layout()
    th = np.asarray(theta, np.float32)
    for w, b, ci, co, k in lay:
        out.append(struct.pack("<III", ci, co, k))
        out.append(th[w:b + co].tobytes())
    return b"".join(out)
