"""ctypes binding of libfluidsim.so (the C ABI declared in include/fluidsim.h).

No fallback exists: if the HIP library is missing, importing the solver raises.
"""
from __future__ import annotations

import ctypes as C
import os

from . import build as _build

ABI_VERSION = 3  # FS_ABI_VERSION of include/fluidsim.h these structures mirror

FS_OK = 0
FS_E_INVALID_ARG = -1
FS_E_DEVICE = -2
FS_E_OOM = -3
FS_E_NONFINITE = -4
FS_E_NOT_CONVERGED = -5
FS_E_UNSUPPORTED = -6

ADVECTION = {"semi-lagrangian": 0, "maccormack": 1}
PRESSURE = {"jacobi": 0, "red-black": 1, "multigrid": 2}
PRECISION = {"fp32": 0, "fp16": 1, "bf16": 2}
GUESS = {"zero": 0, "previous": 1, "learned": 2, "learned-coarse": 3, "learned-quarter": 4,
         "extrapolated": 5}
FIELD = {"u": 0, "v": 1, "div": 2, "pressure": 3, "residual": 4, "dye": 5, "output": 6, "bloom": 7}
PROFILE_KINDS = {"advect_sl": 0, "advect_mc": 1, "diffuse": 2, "down0": 3, "down": 4, "coarse": 5, "up": 6,
                 "gradient": 7, "residual": 8, "rb": 9, "guess": 10,
                 "advect_fused": 11, "forcing": 12, "up_coarse": 13}
KERNEL = {"advect-sl": 0, "advect-mc": 1, "divergence": 2, "diffuse": 3, "project": 4, "gradient": 5, "solid": 6,
          "forcing": 7, "advect-dye": 8, "confine": 9}


class FsConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double),
        ("nu", C.c_double),
        ("advection", C.c_int),
        ("pressure_solver", C.c_int),
        ("pressure_iterations", C.c_int),
        ("tolerance", C.c_double),
        ("mg_min_size", C.c_int),
        ("mg_pre", C.c_int), ("mg_post", C.c_int),
        ("mg_coarse", C.c_int),
        ("mg_precision", C.c_int),
        ("guess", C.c_int),
        ("solid_mask_width", C.c_double),
        ("device", C.c_int),
        ("mg_fused", C.c_int),
        ("autotune", C.c_int),
        ("adv_fused", C.c_int),
        ("gated", C.c_int),
        ("graph", C.c_int),
        ("dye_fade", C.c_double),
        ("mg_coarse_exact", C.c_int),
        ("tile", C.c_int),
        ("confinement", C.c_double),
    ]


class FsSplat(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("r", C.c_double), ("g", C.c_double), ("b", C.c_double), ("radius", C.c_double),
                ("is_dye", C.c_int)]


class FsStats(C.Structure):
    _fields_ = [
        ("t", C.c_double),
        ("steps", C.c_int64),
        ("last_vcycles", C.c_int),
        ("total_vcycles", C.c_int64),
        ("last_residual", C.c_double),
        ("last_dt", C.c_double),
        ("gated_steps", C.c_int64),
        ("gate_misses", C.c_int64),
        ("graph_captures", C.c_int64),
        ("graph_launches", C.c_int64),
        ("predicted_vcycles", C.c_int),
    ]


class FsDiag(C.Structure):
    _fields_ = [
        ("kinetic_energy", C.c_double),
        ("max_divergence", C.c_double),
        ("l2_divergence", C.c_double),
        ("momentum_x", C.c_double), ("momentum_y", C.c_double),
        ("max_abs_u", C.c_double), ("max_abs_v", C.c_double),
        ("nonfinite", C.c_int),
        ("enstrophy", C.c_double),
    ]


class FsForcing(C.Structure):
    _fields_ = [
        ("kf", C.c_double), ("width", C.c_double), ("amp", C.c_double),
        ("drag", C.c_double),
        ("seed", C.c_uint32),
    ]


class FsDisk(C.Structure):
    _fields_ = [
        ("x", C.c_double), ("y", C.c_double), ("r", C.c_double),
        ("vx", C.c_double), ("vy", C.c_double), ("omega", C.c_double),
        ("passes", C.c_int),
    ]


class FsCavity(C.Structure):
    _fields_ = [("n", C.c_int), ("lid", C.c_double), ("nu", C.c_double), ("has_disk", C.c_int),
                ("disk_x", C.c_double), ("disk_y", C.c_double), ("disk_r", C.c_double)]


class FsBatchConfig(C.Structure):
    _fields_ = [("advection", C.c_int), ("tolerance", C.c_double), ("max_vcycles", C.c_int),
                ("mg_pre", C.c_int), ("mg_post", C.c_int), ("coupling_passes", C.c_int),
                ("mask_width", C.c_double), ("guess", C.c_int), ("device", C.c_int), ("mg_fused", C.c_int),
                ("graph", C.c_int), ("autotune", C.c_int), ("tile", C.c_int), ("mg_coarse_exact", C.c_int)]


class FsBatchStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("cells", C.c_int64), ("last_vcycles", C.c_int),
                ("total_vcycles", C.c_int64), ("last_residual", C.c_double), ("graph_captures", C.c_int64),
                ("graph_replays", C.c_int64)]


class FluidSimError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (status {code})")
        self.code = code


class NotConverged(FluidSimError):
    pass


_VP = C.c_void_p
_SIGNATURES = {
    "fs_abi_version": (C.c_int, []),
    "fs_last_error": (C.c_char_p, []),
    "fs_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "fs_device_name": (C.c_int, [C.c_int, C.c_char_p, C.c_size_t]),
    "fs_config_default": (None, [C.POINTER(FsConfig)]),
    "fs_solver_create": (C.c_int, [C.POINTER(FsConfig), C.POINTER(_VP)]),
    "fs_solver_destroy": (C.c_int, [_VP]),
    "fs_upload_velocity_f64": (C.c_int, [_VP, _VP, _VP, C.c_size_t]),
    "fs_upload_velocity_f32": (C.c_int, [_VP, _VP, _VP, C.c_size_t]),
    "fs_read_velocity": (C.c_int, [_VP, _VP, _VP, C.c_size_t]),
    "fs_read_field": (C.c_int, [_VP, C.c_int, _VP, C.c_size_t]),
    "fs_step": (C.c_int, [_VP, C.c_double]),
    "fs_step_splats": (C.c_int, [_VP, C.c_double, _VP, C.c_int]),
    "fs_upload_dye": (C.c_int, [_VP, _VP, C.c_size_t]),
    "fs_render": (C.c_int, [_VP, C.c_double, C.c_double, C.c_double, C.c_int]),
    "fs_set_confinement": (C.c_int, [_VP, C.c_double]),
    "fs_step_cfl": (C.c_int, [_VP, C.c_double]),
    "fs_sync": (C.c_int, [_VP]),
    "fs_max_divergence": (C.c_int, [_VP, C.POINTER(C.c_double)]),
    "fs_poisson_residual_after": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_double)]),
    "fs_set_pressure_iterations": (C.c_int, [_VP, C.c_int]),
    "fs_set_tolerance": (C.c_int, [_VP, C.c_double]),
    "fs_set_nu": (C.c_int, [_VP, C.c_double]),
    "fs_set_guess": (C.c_int, [_VP, C.c_int]),
    "fs_set_forcing": (C.c_int, [_VP, C.POINTER(FsForcing)]),
    "fs_set_forcing_modes": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int, C.c_double, C.c_double, C.c_double]),
    "fs_set_solid": (C.c_int, [_VP, C.POINTER(FsDisk)]),
    "fs_solid_impulse": (C.c_int, [_VP, C.POINTER(C.c_double)]),
    "fs_get_stats": (C.c_int, [_VP, C.POINTER(FsStats)]),
    "fs_diagnostics": (C.c_int, [_VP, C.POINTER(FsDiag)]),
    "fs_diagnostics_noalloc": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_double, C.c_double, _VP, _VP]),
    "fs_run_kernel": (C.c_int, [_VP, C.c_int, C.c_double]),
    "fs_profile_enable": (C.c_int, [_VP, C.c_uint]),
    "fs_tile_report": (C.c_int, [_VP, C.c_char_p, C.c_size_t]),
    "fs_load_guess": (C.c_int, [_VP, C.c_char_p, C.c_size_t]),
    "fs_energy_spectrum": (C.c_int, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int]),
    "fs_batch_config_default": (None, [C.POINTER(FsBatchConfig)]),
    "fs_batch_create": (C.c_int, [C.POINTER(FsBatchConfig), C.POINTER(FsCavity), C.c_int, C.POINTER(_VP)]),
    "fs_batch_destroy": (C.c_int, [_VP]),
    "fs_batch_step_cfl": (C.c_int, [_VP, C.c_double]),
    "fs_batch_sync": (C.c_int, [_VP]),
    "fs_batch_read_velocity": (C.c_int, [_VP, C.c_int, _VP, _VP, C.c_size_t]),
    "fs_batch_read_dt": (C.c_int, [_VP, C.POINTER(C.c_double), C.c_int]),
    "fs_batch_get_stats": (C.c_int, [_VP, C.POINTER(FsBatchStats)]),
    "fs_batch_tune_report": (C.c_int, [_VP, C.c_char_p, C.c_size_t]),
    "fs_profile_read": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_double)]),
    "fs_trainer_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(_VP)]),
    "fs_trainer_destroy": (C.c_int, [_VP]),
    "fs_trainer_param_count": (C.c_int, [_VP, C.POINTER(C.c_size_t)]),
    "fs_trainer_load": (C.c_int, [_VP, C.c_char_p, C.c_size_t]),
    "fs_trainer_export": (C.c_int, [_VP, _VP, C.c_size_t, C.POINTER(C.c_size_t)]),
    "fs_trainer_reserve": (C.c_int, [_VP, C.c_size_t]),
    "fs_trainer_store": (C.c_int, [_VP, C.c_size_t, _VP, C.c_size_t]),
    "fs_trainer_store_pressure": (C.c_int, [_VP, C.c_size_t, _VP]),
    "fs_trainer_batch": (C.c_int, [_VP, C.POINTER(C.c_int), C.c_int]),
    "fs_trainer_loss": (C.c_int, [_VP, C.POINTER(C.c_double)]),
    "fs_trainer_read_gradients": (C.c_int, [_VP, _VP, C.c_size_t]),
    "fs_trainer_read_params": (C.c_int, [_VP, _VP, C.c_size_t]),
    "fs_trainer_read_error": (C.c_int, [_VP, _VP, C.c_size_t]),
    "fs_trainer_adam": (C.c_int, [_VP, C.c_double, C.c_double, C.c_double, C.c_double]),
    "fs_trainer_sync": (C.c_int, [_VP]),
}

_lib = None


def library_path() -> str:
    return _build.LIB


def load():
    """Load the in-tree libfluidsim.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise ImportError(
            f"libfluidsim.so not found at {path}; build it with "
            "`python fluid-sim_amd/fluidsim/build.py` (hipcc, gfx950)")
    L = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.fs_abi_version() != ABI_VERSION:
        raise ImportError(f"{path} has ABI version {L.fs_abi_version()}, this binding expects {ABI_VERSION}; "
                          "rebuild it")
    _lib = L
    return L


def last_error() -> str:
    msg = load().fs_last_error()
    return msg.decode() if msg else ""


def check(code: int):
    if code == FS_OK:
        return
    msg = last_error()
    if code == FS_E_NOT_CONVERGED:
        raise NotConverged(code, msg)
    if code == FS_E_INVALID_ARG:
        raise ValueError(f"{msg} (status {code})")
    raise FluidSimError(code, msg)


def default_config() -> FsConfig:
    cfg = FsConfig()
    load().fs_config_default(C.byref(cfg))
    return cfg
