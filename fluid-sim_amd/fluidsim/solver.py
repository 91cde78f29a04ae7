"""Python mirror of the reference's GPU solver API (src/gpu/solver.js) over libfluidsim.so.

`createDevice()` and `GPUFluidSolver` keep the reference's names, options,
defaults and error behaviour, so callers and parity tests read like the
reference's own harness (tools/gputest.html, tools/bench.html). The differences
are deliberate and listed on each method: unknown selectors raise instead of
silently running Jacobi, readbacks return numpy arrays instead of Promises, and
the multigrid projection can iterate to an FP64 residual tolerance.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import FsConfig, FsDiag, FsForcing, FsStats, check, load

TWO_PI = 2.0 * math.pi


def createDevice(requireTimestamps: bool = False, device: int = 0):
    """createDevice() (src/gpu/solver.js:43-62): the HIP device and its name.

    Timestamps are always available (HIP events), so `requireTimestamps` never fails.
    """
    L = load()
    n = C.c_int()
    check(L.fs_device_count(C.byref(n)))
    if n.value == 0:
        raise RuntimeError("no HIP device available")
    buf = C.create_string_buffer(256)
    check(L.fs_device_name(device, buf, 256))
    return {"device": device, "hasTimestamps": True,
            "info": {"vendor": "amd", "description": buf.value.decode()}}


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class GPUFluidSolver:
    """new GPUFluidSolver(device, opts) (src/gpu/solver.js:64-98).

    Options and defaults follow the reference: n=256, lx=ly=2pi, nu=0,
    advection="maccormack", pressureSolver="jacobi", pressureIterations=40,
    multigrid V(2,2) with 40 coarse sweeps and minSize 8. Extra options:
    tolerance (FP64 residual gate for multigrid; 0 = fixed V-cycle count),
    mgMinSize, mgPrecision, guess ("zero" | "previous"), gated (the tolerance
    checked on the device, no host wait inside a step), graph (gated steps
    replayed as captured hipGraphs) and mgCoarseSolve ("sweeps": the reference's
    V-cycle down to mgMinSize; "exact": the hierarchy stops at its first level of at
    most 64 x 64 cells and solves it exactly).
    """

    def __init__(self, device=0, *, n=256, nx=None, ny=None, lx=TWO_PI, ly=TWO_PI, nu=0.0,
                 confinement=0, dyeFade=0, advection="maccormack", solidMaskWidth=1.5,
                 pressureSolver="jacobi", pressureIterations=40, hasTimestamps=False,
                 tolerance=0.0, mgMinSize=8, mgPre=2, mgPost=2, mgCoarse=40,
                 mgPrecision="fp32", guess="zero", mgFused=True, autotune=True,
                 advFused=True, gated=True, graph=False, mgCoarseSolve="sweeps", tile=None):
        if isinstance(device, dict):
            device = device.get("device", 0)
        nx = n if nx is None else nx
        ny = n if ny is None else ny
        if advection not in _lib.ADVECTION:
            raise ValueError(f"unknown advection scheme {advection}; have {', '.join(_lib.ADVECTION)}")
        if pressureSolver not in _lib.PRESSURE:
            raise ValueError(f"unknown pressure solver {pressureSolver}; have {', '.join(_lib.PRESSURE)}")
        if mgPrecision not in _lib.PRECISION:
            raise ValueError(f"unknown multigrid precision {mgPrecision}")
        if guess not in _lib.GUESS:
            raise ValueError(f"unknown guess mode {guess}")
        if not (0.0 <= dyeFade <= 1.0):
            raise ValueError("dyeFade must be in [0, 1]")
        if mgCoarseSolve not in ("sweeps", "exact"):
            raise ValueError(f"unknown coarse solve {mgCoarseSolve}; have sweeps, exact")
        cfg = _lib.default_config()
        cfg.nx, cfg.ny, cfg.lx, cfg.ly, cfg.nu = nx, ny, lx, ly, nu
        cfg.advection = _lib.ADVECTION[advection]
        cfg.pressure_solver = _lib.PRESSURE[pressureSolver]
        cfg.pressure_iterations = int(pressureIterations)
        cfg.tolerance = float(tolerance)
        cfg.mg_min_size, cfg.mg_pre, cfg.mg_post, cfg.mg_coarse = mgMinSize, mgPre, mgPost, mgCoarse
        cfg.mg_precision = _lib.PRECISION[mgPrecision]
        cfg.guess = _lib.GUESS[guess]
        cfg.solid_mask_width = solidMaskWidth
        cfg.device = int(device)
        cfg.mg_fused = 1 if mgFused else 0
        cfg.autotune = 1 if autotune else 0
        cfg.adv_fused = 1 if advFused else 0
        cfg.gated = 1 if gated else 0
        cfg.graph = 1 if graph else 0
        cfg.dye_fade = float(dyeFade)
        cfg.confinement = float(confinement)
        cfg.mg_coarse_exact = 1 if mgCoarseSolve == "exact" else 0
        cfg.tile = 0 if tile is None else {"64x32": 1, "32x32": 2, "16x16": 3, "8x8": 4, "64x64": 5,
                                           "32x32/1024": 6, "16x16/256": 7}[tile]
        self._L = load()
        h = C.c_void_p()
        check(self._L.fs_solver_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.nx, self.ny = nx, ny
        self.dx, self.dy = lx / nx, ly / ny
        self.lx, self.ly = lx, ly
        self._nu = nu
        self._confinement = float(confinement)
        self.advection = advection
        self.pressureSolver = pressureSolver
        self._pressureIterations = int(pressureIterations)
        self.hasTimestamps = True
        self.solid = None

    # ---------------------------------------------------------------- state
    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("solver has been destroyed")
        return self._h

    @property
    def nu(self):
        return self._nu

    @nu.setter
    def nu(self, value):
        check(self._L.fs_set_nu(self.handle, float(value)))
        self._nu = float(value)

    @property
    def confinement(self):
        return self._confinement

    @confinement.setter
    def confinement(self, value):  # mutable, as index.html:471 uses it; <= 0 is off
        check(self._L.fs_set_confinement(self.handle, float(value)))
        self._confinement = float(value)

    @property
    def pressureIterations(self):
        return self._pressureIterations

    @pressureIterations.setter
    def pressureIterations(self, value):  # mutable, as index.html:472 uses it
        check(self._L.fs_set_pressure_iterations(self.handle, int(value)))
        self._pressureIterations = int(value)

    def setTolerance(self, tol: float):
        check(self._L.fs_set_tolerance(self.handle, float(tol)))

    def loadGuess(self, blob_or_path):
        """Weights of the learned pressure guess (bytes, or a path to a .fsgn file)."""
        if isinstance(blob_or_path, (str, bytes)) and not isinstance(blob_or_path, bytes):
            with open(blob_or_path, "rb") as f:
                blob = f.read()
        else:
            blob = bytes(blob_or_path)
        check(self._L.fs_load_guess(self.handle, blob, len(blob)))

    def setGuess(self, guess: str):
        """"zero", "previous", "learned" (the network at the grid's resolution),
        "learned-coarse" or "learned-quarter" (the network at half or quarter
        resolution, its increment interpolated back)."""
        if guess not in _lib.GUESS:
            raise ValueError(f"unknown guess mode {guess}")
        check(self._L.fs_set_guess(self.handle, _lib.GUESS[guess]))

    def stats(self) -> dict:
        st = FsStats()
        check(self._L.fs_get_stats(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in FsStats._fields_}

    @property
    def t(self):
        return self.stats()["t"]

    @property
    def steps(self):
        return self.stats()["steps"]

    # ------------------------------------------------------------ data in/out
    def uploadVelocity(self, u, v):
        """uploadVelocity(u, v) (src/gpu/solver.js:301-315): float64 or float32, nx*ny MAC order."""
        u = np.ascontiguousarray(u).reshape(-1)
        v = np.ascontiguousarray(v).reshape(-1)
        n = self.nx * self.ny
        if u.size != n or v.size != n:
            raise ValueError(f"expected {n} values per component, got {u.size}, {v.size}")
        if u.dtype == np.float32 and v.dtype == np.float32:
            check(self._L.fs_upload_velocity_f32(self.handle, _ptr(u), _ptr(v), n))
        else:
            u = np.ascontiguousarray(u, np.float64)
            v = np.ascontiguousarray(v, np.float64)
            check(self._L.fs_upload_velocity_f64(self.handle, _ptr(u), _ptr(v), n))

    def readVelocity(self):
        """readVelocity() (src/gpu/solver.js:356-362) -> {"u": float32[nx*ny], "v": ...}."""
        n = self.nx * self.ny
        u = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        check(self._L.fs_read_velocity(self.handle, _ptr(u), _ptr(v), n))
        return {"u": u, "v": v}

    def readField(self, name: str):
        """readTexture(tex) (src/gpu/solver.js:336-353) by field name. "output" and "bloom"
        (the last render()) come back as float16 rgba, (ny, nx, 4) and (ny/2, nx/2, 4)."""
        if name not in _lib.FIELD:
            raise ValueError(f"unknown field {name}; have {', '.join(_lib.FIELD)}")
        if name in ("output", "bloom"):
            h, w = (self.ny, self.nx) if name == "output" else (self.ny >> 1, self.nx >> 1)
            out = np.empty((h, w, 4), np.float16)
        else:
            dtype = np.float64 if name == "pressure" else np.float32
            out = np.empty(self.nx * self.ny * (4 if name == "dye" else 1), dtype)
        check(self._L.fs_read_field(self.handle, _lib.FIELD[name], _ptr(out), out.nbytes))
        return out

    def render(self, mode=0, scale=1.5, bloomThreshold=0.75, bloomPasses=2):
        """render({mode, scale, bloomThreshold, bloomPasses}) (src/gpu/solver.js:771-823):
        the display image of the current velocity and dye and its bloom, enqueued; read
        them with readField("output") and readField("bloom"). Nothing it computes feeds
        back into the solve. Passes run while k < bloomPasses, as the reference's loop."""
        passes = math.ceil(bloomPasses) if bloomPasses > 0 else 0  # NaN or <= 0: no pass
        check(self._L.fs_render(self.handle, float(mode), float(scale), float(bloomThreshold), int(passes)))

    # ------------------------------------------------------------------ step
    def step(self, dt, splats=()):
        """step(dt, {splats}) (src/gpu/solver.js:518-591): enqueued, returns before the device
        finishes, as the reference's does. With a tolerance the projection is checked
        on the device (stats()["gate_misses"] counts steps that ended above it); a
        step checked on the host (the first after an upload) raises NotConverged if
        the tolerance was not met within pressureIterations V-cycles.

        splats: dicts {x, y, dx, dy, r, g, b, radius, isDye} as the reference's
        (_writeSplat, src/gpu/solver.js:745-757); each keeps its own parameters."""
        self._push_solid()
        if splats:
            arr = (_lib.FsSplat * len(splats))()
            for a, sp in zip(arr, splats):
                a.x, a.y = float(sp["x"]), float(sp["y"])
                a.dx, a.dy = float(sp.get("dx", 0.0) or 0.0), float(sp.get("dy", 0.0) or 0.0)
                a.r, a.g, a.b = (float(sp.get(c, 0.0) or 0.0) for c in ("r", "g", "b"))
                a.radius = float(sp.get("radius", 0.2) or 0.2)
                a.is_dye = 1 if sp.get("isDye") else 0
            check(self._L.fs_step_splats(self.handle, float(dt), arr, len(splats)))
        else:
            check(self._L.fs_step(self.handle, float(dt)))

    def uploadDye(self, rgba):
        """uploadDye(rgba) (src/gpu/solver.js:317-323): nx*ny*4 values, rgba per cell centre."""
        a = np.ascontiguousarray(rgba, np.float32).reshape(-1)
        if a.size != 4 * self.nx * self.ny:
            raise ValueError(f"expected {4 * self.nx * self.ny} values, got {a.size}")
        check(self._L.fs_upload_dye(self.handle, _ptr(a), a.size))

    def stepCFL(self, cfl=0.5):
        """One step with dt = cfl / (max|u|/dx + max|v|/dy), computed on the device.
        Enqueued like step(); the dt it took is stats()["last_dt"] afterwards."""
        self._push_solid()
        check(self._L.fs_step_cfl(self.handle, float(cfl)))

    # ----------------------------------------------------------- immersed disk
    def setSolid(self, disk, passes=1):
        """setSolid(disk | None) (src/gpu/solver.js:425-440). `disk` is a
        fluidsim.solid.RigidDisk (or anything with x, y, r, vx, vy, omega); its
        current state is sent before every step, as _encodeSolid does
        (:442-452). `passes` forcing passes per step (reference GPU: 1, CPU: 2)."""
        self.solid = disk
        self._solid_passes = int(passes)
        if disk is None:
            check(self._L.fs_set_solid(self.handle, None))
        else:
            self._push_solid()

    def _push_solid(self):
        d = getattr(self, "solid", None)
        if d is None:
            return
        fd = _lib.FsDisk(float(d.x), float(d.y), float(d.r), float(d.vx), float(d.vy), float(d.omega),
                         self._solid_passes)
        check(self._L.fs_set_solid(self.handle, C.byref(fd)))

    def solidImpulse(self):
        """FP64 device sum of the last coupling: (fluid impulse x, y, torque impulse), dA included."""
        out = (C.c_double * 3)()
        check(self._L.fs_solid_impulse(self.handle, out))
        return [out[0], out[1], out[2]]

    def integrateSolid(self, dt, gravity=0.0):
        """integrateSolid(dt) (src/gpu/solver.js:454-510) without the one-step lag: the
        impulse of the step just taken, Newton's third law, explicit Euler, and the
        centre wrapped into the periodic box."""
        d = getattr(self, "solid", None)
        if d is None:
            return
        d.apply_reaction(self.solidImpulse(), dt)
        d.integrate(dt, gravity=gravity)
        d.x %= self.lx
        d.y %= self.ly

    def sync(self):
        """device.queue.onSubmittedWorkDone()."""
        check(self._L.fs_sync(self.handle))

    def setForcing(self, forcing, amplitude=1.0, drag=0.0, phaseDrift=0.3):
        """setForcing (src/gpu/solver.js:370-414).

        A list of modes [{kx, ky, phase}] is the reference's own call: its GPU kernel's
        semantics (u exp(-drag dt) + amplitude dt / sqrt(count) * curl modes, phases
        drifting by phaseDrift * t). A dict(kf, width, amp, drag, seed) is the seeded band
        forcing of the CPU solver (src/cpu/solver.js:179-237; step k uses seed + k).
        None or an empty list disables."""
        if forcing is None or (isinstance(forcing, (list, tuple)) and len(forcing) == 0):
            check(self._L.fs_set_forcing(self.handle, None))
            return
        if isinstance(forcing, (list, tuple)):
            kx = np.ascontiguousarray([float(m["kx"]) for m in forcing], np.float64)
            ky = np.ascontiguousarray([float(m["ky"]) for m in forcing], np.float64)
            ph = np.ascontiguousarray([float(m.get("phase", 0.0)) for m in forcing], np.float64)
            check(self._L.fs_set_forcing_modes(self.handle, _ptr(kx), _ptr(ky), _ptr(ph), len(forcing),
                                               float(amplitude), float(drag), float(phaseDrift)))
            return
        f = FsForcing(float(forcing.get("kf", 28)), float(forcing.get("width", 2)),
                      float(forcing.get("amp", 0.4)), float(forcing.get("drag", 0.0)),
                      int(forcing.get("seed", 1)) & 0xFFFFFFFF)
        check(self._L.fs_set_forcing(self.handle, C.byref(f)))

    # ------------------------------------------------------------ diagnostics
    def maxDivergence(self):
        """maxDivergence() (src/gpu/solver.js:835-850), reduced on the device in FP64."""
        out = C.c_double()
        check(self._L.fs_max_divergence(self.handle, C.byref(out)))
        return out.value

    def poissonResidualAfter(self, iterations, dt=0.01):
        """poissonResidualAfter(it, dt) (src/gpu/solver.js:857-896): ||L p - div|| / ||div||."""
        out = C.c_double()
        check(self._L.fs_poisson_residual_after(self.handle, int(iterations), C.byref(out)))
        return out.value

    def diagnostics(self) -> dict:
        d = FsDiag()
        check(self._L.fs_diagnostics(self.handle, C.byref(d)))
        return {k: getattr(d, k) for k, _ in FsDiag._fields_}

    def runKernel(self, name: str, dt: float = 0.0):
        if name not in _lib.KERNEL:
            raise ValueError(f"unknown kernel {name}")
        check(self._L.fs_run_kernel(self.handle, _lib.KERNEL[name], float(dt)))

    def profileEnable(self, kinds=None):
        """Time every launch of the named kernel classes (all when None, none when [])."""
        names = _lib.PROFILE_KINDS if kinds is None else kinds
        mask = 0
        for k in names:
            mask |= 1 << _lib.PROFILE_KINDS[k]
        check(self._L.fs_profile_enable(self.handle, mask))

    def energySpectrum(self):
        """energySpectrum (src/core/fft.js:196-234) of the current velocity, on the device in FP64:
        returns (k, E, counts) for shells k = 0..n/2."""
        kmax = self.nx // 2
        E = (C.c_double * (kmax + 1))()
        cnt = (C.c_int64 * (kmax + 1))()
        check(self._L.fs_energy_spectrum(self.handle, E, cnt, kmax + 1))
        return np.arange(kmax + 1), np.array(E[:]), np.array(cnt[:])

    def tileReport(self) -> str:
        """Tile shape per grid-wide level and the autotuner's timings."""
        buf = C.create_string_buffer(4096)
        check(self._L.fs_tile_report(self.handle, buf, 4096))
        return buf.value.decode()

    def profileRead(self) -> dict:
        """{kind: (total_ms, launches, algorithmic_bytes)} since profileEnable."""
        out = {}
        for name, k in _lib.PROFILE_KINDS.items():
            ms, n, b = C.c_double(), C.c_int64(), C.c_double()
            check(self._L.fs_profile_read(self.handle, k, C.byref(ms), C.byref(n), C.byref(b)))
            out[name] = (ms.value, n.value, b.value)
        return out

    # -------------------------------------------------------------- lifetime
    def destroy(self):
        """destroy() (src/gpu/solver.js:898-911); idempotent."""
        if getattr(self, "_h", None) is not None:
            self._L.fs_solver_destroy(self._h)
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
