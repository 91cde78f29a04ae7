"""The C-ABI library loads and exports every entry point include/fluidsim.h
declares; host-side validation rejects bad configurations before any device
work. CPU only: no compute call is made."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "fluidsim.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(fs_\w+)\s*\(", text, re.M)))


def test_library_builds_and_exports_every_declared_symbol():
    import fluidsim
    L = fluidsim.load()
    names = declared_entry_points()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), f"{n} declared in fluidsim.h but not exported"
    assert L.fs_abi_version() == 3


def test_python_binding_covers_every_declared_symbol():
    from fluidsim import _lib
    assert set(declared_entry_points()) <= set(_lib._SIGNATURES)


def test_defaults_follow_the_reference_constructor():
    # src/gpu/solver.js:64-98, src/gpu/multigrid.js:159-161
    from fluidsim import _lib
    c = _lib.default_config()
    assert (c.nx, c.ny) == (256, 256)
    assert abs(c.lx - 2 * 3.141592653589793) < 1e-15
    assert c.nu == 0.0
    assert c.advection == _lib.ADVECTION["maccormack"]
    assert c.pressure_solver == _lib.PRESSURE["jacobi"]
    assert c.pressure_iterations == 40
    assert (c.mg_min_size, c.mg_pre, c.mg_post, c.mg_coarse) == (8, 2, 2, 40)
    assert c.solid_mask_width == 1.5


@pytest.mark.parametrize("field,value,msg", [
    ("advection", 7, "unknown advection"),
    ("pressure_solver", 9, "unknown pressure solver"),
    ("nx", 1, "grid must be"),
    ("pressure_iterations", -1, "pressure_iterations"),
])
def test_invalid_config_is_rejected_with_a_message(field, value, msg):
    from fluidsim import _lib
    L = _lib.load()
    c = _lib.default_config()
    setattr(c, field, value)
    h = C.c_void_p()
    rc = L.fs_solver_create(C.byref(c), C.byref(h))
    assert rc == _lib.FS_E_INVALID_ARG
    assert msg in _lib.last_error()
    assert h.value is None


def test_odd_grid_rejected_for_red_black_orderings():
    from fluidsim import _lib
    L = _lib.load()
    for solver in ("red-black", "multigrid"):
        c = _lib.default_config()
        c.pressure_solver = _lib.PRESSURE[solver]
        c.nx = c.ny = 63
        h = C.c_void_p()
        assert L.fs_solver_create(C.byref(c), C.byref(h)) == _lib.FS_E_INVALID_ARG
        assert "even grid" in _lib.last_error()


@pytest.mark.parametrize("prec,fused,solver,msg", [
    ("fp16", 1, "multigrid", "FP16 multigrid storage is not built"),
    ("bf16", 0, "multigrid", "needs the fused multigrid"),
    ("bf16", 1, "jacobi", "needs the fused multigrid"),
])
def test_storage_precision_limits_are_reported(prec, fused, solver, msg):
    from fluidsim import _lib
    L = _lib.load()
    c = _lib.default_config()
    c.mg_precision = _lib.PRECISION[prec]
    c.mg_fused = fused
    c.pressure_solver = _lib.PRESSURE[solver]
    h = C.c_void_p()
    assert L.fs_solver_create(C.byref(c), C.byref(h)) == _lib.FS_E_UNSUPPORTED
    assert msg in _lib.last_error()


def test_tolerance_needs_multigrid():
    from fluidsim import _lib
    L = _lib.load()
    c = _lib.default_config()
    c.tolerance = 1e-6
    h = C.c_void_p()
    assert L.fs_solver_create(C.byref(c), C.byref(h)) == _lib.FS_E_INVALID_ARG
    assert "multigrid" in _lib.last_error()


def test_python_mirror_rejects_unknown_selectors():
    # the reference silently runs Jacobi / SL (src/gpu/solver.js:606, :725);
    # this build raises, like the CPU solver (src/cpu/solver.js:265-290)
    import fluidsim
    with pytest.raises(ValueError, match="unknown pressure solver"):
        fluidsim.GPUFluidSolver(n=64, pressureSolver="sor")
    with pytest.raises(ValueError, match="unknown advection"):
        fluidsim.GPUFluidSolver(n=64, advection="weno")


def test_null_handles_are_errors_not_crashes():
    from fluidsim import _lib
    L = _lib.load()
    assert L.fs_step(None, 0.01) == _lib.FS_E_INVALID_ARG
    assert L.fs_solver_destroy(None) == _lib.FS_OK


def test_initial_turbulence_field_is_discretely_divergence_free(O):
    from fluidsim import initial
    import numpy as np
    u, v = initial.decaying_turbulence(128, seed=1)
    assert O.max_divergence(128, 128, u, v) < 1e-11
    assert abs(np.sqrt(np.mean(u * u) + np.mean(v * v)) - 1.0) < 1e-12
    u2, _ = initial.decaying_turbulence(128, seed=1)
    assert (u == u2).all()


def test_node_addon_loads_and_exports_the_shim_surface():
    # the Node-API addon over the C ABI (drop-in for src/gpu/solver.js); no device call
    import shutil
    import subprocess
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed")
    addon = os.path.join(ROOT, "fluid-sim_amd", "node", "fluidsim.node")
    out = subprocess.run([node, "-e", f"console.log(Object.keys(require({addon!r})).join(','))"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    names = set(out.stdout.strip().split(","))
    assert {"create", "destroy", "uploadVelocity", "readVelocity", "step", "stepCFL", "sync", "maxDivergence",
            "poissonResidualAfter", "setParam", "setForcing", "stats", "setSolid", "solidImpulse", "readField",
            "energySpectrum", "syncAsync", "runKernel", "setForcingModes"} <= names
    shim = open(os.path.join(ROOT, "fluid-sim_amd", "node", "solver.mjs")).read()
    assert "export async function createDevice" in shim and "export class GPUFluidSolver" in shim
