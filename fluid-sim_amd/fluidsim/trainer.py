"""Host side of the learned guess's trainer (fs_trainer_* in include/fluidsim.h).

The forward pass, the loss, the backward pass and Adam all run in libfluidsim.so;
the pressure snapshots the pairs come from live in the trainer's own device
store, filled from a solver's pressure (device to device) or from host arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, load


class GuessTrainer:
    """Trains the guess network on pairs of consecutive pressure snapshots."""

    def __init__(self, device=0, n=256, batch=16, h=None):
        import math
        self._L = load()
        self.n = n
        self.batch = batch
        self.h = h if h is not None else 2 * math.pi / n
        handle = C.c_void_p()
        check(self._L.fs_trainer_create(device, n, batch, self.h, C.byref(handle)))
        self._h = handle
        cnt = C.c_size_t()
        check(self._L.fs_trainer_param_count(self._h, C.byref(cnt)))
        self.param_count = cnt.value

    def close(self):
        if getattr(self, "_h", None):
            self._L.fs_trainer_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, blob: bytes):
        """Parameters from a weights blob; clears the Adam moments."""
        check(self._L.fs_trainer_load(self._h, blob, len(blob)))

    def export(self) -> bytes:
        need = C.c_size_t()
        check(self._L.fs_trainer_export(self._h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        check(self._L.fs_trainer_export(self._h, buf, need.value, C.byref(need)))
        return buf.raw

    def reserve(self, count):
        """Room for `count` snapshots (zeroed); replaces any held."""
        check(self._L.fs_trainer_reserve(self._h, int(count)))
        self.snapshots = int(count)

    def store(self, slot, p):
        """Snapshot `slot` from a host array of n x n pressures."""
        p = np.ascontiguousarray(p, np.float64).reshape(-1)
        check(self._L.fs_trainer_store(self._h, int(slot), p.ctypes.data, p.size))

    def store_pressure(self, slot, solver):
        """Snapshot `slot` from a GPUFluidSolver's current pressure, device to device."""
        check(self._L.fs_trainer_store_pressure(self._h, int(slot), solver.handle))

    def batch_step(self, idx):
        """Forward and backward over the pairs (snapshot idx[b], snapshot idx[b] + 1)."""
        idx = np.ascontiguousarray(idx, np.int32)
        if idx.shape != (self.batch,):
            raise ValueError(f"idx must hold {self.batch} snapshot indices")
        check(self._L.fs_trainer_batch(self._h, idx.ctypes.data_as(C.POINTER(C.c_int)), self.batch))

    def loss(self) -> float:
        out = C.c_double()
        check(self._L.fs_trainer_loss(self._h, C.byref(out)))
        return out.value

    def gradients(self) -> np.ndarray:
        g = np.zeros(self.param_count, np.float32)
        check(self._L.fs_trainer_read_gradients(self._h, g.ctypes.data, g.size))
        return g

    def error(self) -> np.ndarray:
        """The last batch's output minus target, [batch][n][n]."""
        e = np.zeros((self.batch, self.n, self.n), np.float32)
        check(self._L.fs_trainer_read_error(self._h, e.ctypes.data, e.size))
        return e

    def params(self) -> np.ndarray:
        p = np.zeros(self.param_count, np.float32)
        check(self._L.fs_trainer_read_params(self._h, p.ctypes.data, p.size))
        return p

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(self._L.fs_trainer_adam(self._h, lr, beta1, beta2, eps))

    def sync(self):
        check(self._L.fs_trainer_sync(self._h))
