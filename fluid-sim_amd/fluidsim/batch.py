"""Batched lid-driven cavities (walls, lid, pinned disk) over the C ABI.

Every stage of a step is one launch over all cavities of the batch, each cavity
with its own size, lid velocity, viscosity, disk and dt. The reference's domain
is periodic; the walled box is this build's own counterpart, checked against the
f64 oracle's cavity step (oracle/fsoracle.c, or_cav_*).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FsBatchConfig, FsBatchStats, FsCavity, check, load
from .initial import uniform01


def random_cavities(count=256, seed=1, sizes=(64, 128, 256, 512)):
    """The batch workload: sizes uniform over `sizes`, lid U ~ U[0.5, 2], Re ~ U[100, 5000]
    (nu = U / Re on the unit box), a pinned disk of radius r ~ U[0.05, 0.2] centred
    uniformly in [r + 2h, 1 - r - 2h]^2; one splitmix64 stream, seven draws per cavity."""
    r = uniform01(seed, 7 * count).reshape(count, 7)
    out = []
    for k in range(count):
        n = int(sizes[min(int(r[k, 0] * len(sizes)), len(sizes) - 1)])
        h = 1.0 / n
        lid = 0.5 + 1.5 * r[k, 1]
        re = 100.0 + 4900.0 * r[k, 2]
        rad = 0.05 + 0.15 * r[k, 3]
        lo, hi = rad + 2 * h, 1 - rad - 2 * h
        out.append(dict(n=n, lid=lid, nu=lid / re, disk=dict(x=lo + (hi - lo) * r[k, 4],
                                                               y=lo + (hi - lo) * r[k, 5], r=rad)))
    return out


class CavityBatch:
    """A batch of cavities: list of dict(n, lid, nu, disk=None | dict(x, y, r))."""

    def __init__(self, cavities, device=0, advection="maccormack", tolerance=1e-6, maxVcycles=60, mgPre=2,
                 mgPost=2, couplingPasses=2, maskWidth=1.5, guess="zero", mgFused=True, graph=True,
                 autotune=1, tile=None, coarseSolve="sweeps"):
        if advection not in _lib.ADVECTION:
            raise ValueError(f"unknown advection scheme {advection}")
        if guess not in ("zero", "previous"):
            raise ValueError("a batch takes the zero or the previous pressure guess")
        self._L = load()
        cfg = FsBatchConfig()
        self._L.fs_batch_config_default(C.byref(cfg))
        cfg.advection = _lib.ADVECTION[advection]
        cfg.tolerance, cfg.max_vcycles = float(tolerance), int(maxVcycles)
        cfg.mg_pre, cfg.mg_post = int(mgPre), int(mgPost)
        cfg.coupling_passes, cfg.mask_width = int(couplingPasses), float(maskWidth)
        cfg.guess = _lib.GUESS[guess]
        cfg.device = int(device)
        cfg.mg_fused = 1 if mgFused else 0
        cfg.graph = 1 if graph else 0
        cfg.autotune = int(autotune)
        cfg.tile = 0 if tile is None else {"64x32": 1, "32x32": 2, "16x16": 3, "8x8": 4, "64x64": 5,
                                           "32x32/1024": 6, "16x16/256": 7}[tile]
        if coarseSolve not in ("sweeps", "exact"):
            raise ValueError(f"unknown coarse solve {coarseSolve}; have sweeps, exact")
        cfg.mg_coarse_exact = 1 if coarseSolve == "exact" else 0
        arr = (FsCavity * len(cavities))()
        for a, c in zip(arr, cavities):
            a.n, a.lid, a.nu = int(c["n"]), float(c["lid"]), float(c["nu"])
            d = c.get("disk")
            a.has_disk = 1 if d else 0
            if d:
                a.disk_x, a.disk_y, a.disk_r = float(d["x"]), float(d["y"]), float(d["r"])
        h = C.c_void_p()
        check(self._L.fs_batch_create(C.byref(cfg), arr, len(cavities), C.byref(h)))
        self._h = h
        self.sizes = [int(c["n"]) for c in cavities]

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("batch has been destroyed")
        return self._h

    def stepCFL(self, cfl=0.5):
        check(self._L.fs_batch_step_cfl(self.handle, float(cfl)))

    def sync(self):
        check(self._L.fs_batch_sync(self.handle))

    def readVelocity(self, k):
        n = self.sizes[k]
        u = np.empty(n * n, np.float32)
        v = np.empty(n * n, np.float32)
        check(self._L.fs_batch_read_velocity(self.handle, k, u.ctypes.data_as(C.c_void_p),
                                             v.ctypes.data_as(C.c_void_p), n * n))
        return {"u": u, "v": v}

    def lastDt(self):
        dt = (C.c_double * len(self.sizes))()
        check(self._L.fs_batch_read_dt(self.handle, dt, len(self.sizes)))
        return np.array(dt[:])

    def tuneReport(self) -> str:
        buf = C.create_string_buffer(8192)
        check(self._L.fs_batch_tune_report(self.handle, buf, len(buf)))
        return buf.value.decode()

    def stats(self):
        st = FsBatchStats()
        check(self._L.fs_batch_get_stats(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in FsBatchStats._fields_}

    def destroy(self):
        if getattr(self, "_h", None) is not None:
            self._L.fs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
