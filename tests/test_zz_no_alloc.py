"""fs_diagnostics_noalloc (the variant of fs_diagnostics that allocates nothing) against
fs_diagnostics on the same fields, through the C ABI: two calls in a row, the second on
a smaller grid, so the shared scratch of the first call cannot leak into the second."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ["kinetic_energy", "max_divergence", "l2_divergence", "momentum_x", "momentum_y", "max_abs_u",
          "max_abs_v", "nonfinite", "enstrophy"]
EXACT = {"max_divergence", "max_abs_u", "max_abs_v", "nonfinite"}   # maxima and counts: no rounding order


class Hip:
    """The few HIP runtime calls the test needs to own device buffers, from the HIP
    runtime libfluidsim.so is bound to (a process can map another copy, e.g. torch's,
    whose device pointers that runtime does not know)."""

    def __init__(self):
        import fluidsim
        fluidsim.load()
        with open("/proc/self/maps") as f:
            paths = sorted({ln.split()[-1] for ln in f if "libamdhip64.so" in ln})
        # one copy: the one libfluidsim.so resolved; two (torch's own besides the system's,
        # under different sonames): libfluidsim.so links the system's
        if len(paths) > 1:
            paths = [q for q in paths if "site-packages" not in q and "dist-packages" not in q] or paths
        assert paths, "no HIP runtime mapped"
        self.h = C.CDLL(paths[0])

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        return p

    def h2d(self, dst, a):
        assert self.h.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def d2h(self, a, src):
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, C.c_size_t(a.nbytes), 2) == 0

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self, p):
        self.h.hipFree(p)


def test_noalloc_diagnostics_match_the_solver_diagnostics_on_two_calls():
    import fluidsim
    from fluidsim import initial
    fluidsim.createDevice()
    L = fluidsim.load()
    fn = L.fs_diagnostics_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    hip = Hip()
    for n, seed in ((256, 1), (64, 2)):
        print(f"noalloc diagnostics n={n} seed={seed}", flush=True)
        u0, v0 = initial.decaying_turbulence(n, seed=seed)
        u32, v32 = np.asarray(u0, np.float32), np.asarray(v0, np.float32)
        with fluidsim.GPUFluidSolver(0, n=n) as g:
            g.uploadVelocity(u32, v32)
            ref = g.diagnostics()
        du, dv, dout = hip.malloc(u32.nbytes), hip.malloc(v32.nbytes), hip.malloc(9 * 8)
        try:
            hip.h2d(du, u32)
            hip.h2d(dv, v32)
            rc = fn(du, dv, n, n, 2 * math.pi, 2 * math.pi, dout, None)
            assert rc == 0, fluidsim._lib.last_error()
            hip.sync()
            out = np.empty(9)
            hip.d2h(out, dout)
        finally:
            for p in (du, dv, dout):
                hip.free(p)
        for k, name in enumerate(FIELDS):
            want = ref[name]
            if name in EXACT:
                assert out[k] == want, (name, out[k], want)
            else:
                assert abs(out[k] - want) <= 1e-12 * max(abs(want), 1e-300) + 1e-15, (name, out[k], want)


def test_noalloc_diagnostics_reject_bad_sizes():
    import fluidsim
    L = fluidsim.load()
    fn = L.fs_diagnostics_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    dummy = C.c_void_p(16)  # never dereferenced: the sizes are rejected first
    assert fn(dummy, dummy, 1 << 15, 1 << 14, 1.0, 1.0, dummy, None) == fluidsim._lib.FS_E_INVALID_ARG
    assert "2^28" in fluidsim._lib.last_error()
    assert fn(dummy, dummy, 1, 64, 1.0, 1.0, dummy, None) == fluidsim._lib.FS_E_INVALID_ARG
    assert fn(None, dummy, 64, 64, 1.0, 1.0, dummy, None) == fluidsim._lib.FS_E_INVALID_ARG
