"""The learned pressure guess: its layer table and the weights-blob format the HIP
library loads (fs_load_guess), and seeded initial weights for training.

The library evaluates and trains the network itself (fs_guess.hpp: BF16
matrix-core convolutions; fs_train.hpp: loss, backward pass, Adam).

Network (channels 16/32/64/64, periodic 3x3 convolutions + ReLU, 2x2 average
pooling down, nearest-neighbour upsampling and [upsampled, skip] concatenation
up, 1x1 output). Inputs: the residual the previous pressure leaves, r = div u* -
L p_prev, and p_prev, each divided by its maximum magnitude. Output o:
p_guess = p_prev + max|r| h^2 o.
"""
from __future__ import annotations

import struct

import numpy as np

# (cin, cout) of the 3x3 layers in evaluation order; then the 1x1 output 16 -> 1
LAYERS = [(2, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 64), (64, 64),
          (128, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]
MAGIC = b"FSGN"


def init_blob(seed: int) -> bytes:
    """Initial weights blob: every weight and bias uniform in +-1/sqrt(cin k^2), seeded."""
    rng = np.random.default_rng(seed)
    out = [MAGIC, struct.pack("<II", 1, 15)]
    for (ci, co), k in list(zip(LAYERS, [3] * len(LAYERS))) + [((16, 1), 1)]:
        bound = 1.0 / np.sqrt(ci * k * k)
        out.append(struct.pack("<III", ci, co, k))
        out.append(rng.uniform(-bound, bound, co * k * k * ci).astype(np.float32).tobytes())
        out.append(rng.uniform(-bound, bound, co).astype(np.float32).tobytes())
    return b"".join(out)


def load_blob(blob: bytes):
    """Parse a blob into [(w[cout][cin][k][k] float32, b[cout])] in torch's layout."""
    if blob[:4] != MAGIC or struct.unpack("<II", blob[4:12]) != (1, 15):
        raise ValueError("not a version-1 guess network blob")
    off = 12
    layers = []
    for _ in range(15):
        cin, cout, k = struct.unpack("<III", blob[off:off + 12])
        off += 12
        nw = cout * k * k * cin
        w = np.frombuffer(blob, np.float32, nw, off).reshape(cout, k, k, cin).transpose(0, 3, 1, 2)
        off += 4 * nw
        b = np.frombuffer(blob, np.float32, cout, off)
        off += 4 * cout
        layers.append((np.ascontiguousarray(w), b.copy()))
    if off != len(blob):
        raise ValueError("trailing bytes in the blob")
    return layers
