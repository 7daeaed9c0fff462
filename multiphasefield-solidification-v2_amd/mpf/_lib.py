"""ctypes binding of libmpf.so (include/mpf.h).

The library is the only compute path: if it is missing or cannot be loaded
this module raises, it never falls back to a host implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmpf.so")

FP32, FP64 = 0, 1
MPF_EINVAL, MPF_EHIP, MPF_ENOMEM, MPF_ECAPACITY = -1, -2, -3, -4

SCALAR_NAMES = ("dx", "dt", "T_melt", "Sf", "eps0_sl", "w0_sl", "a0", "delta_a", "mu_a",
                "p_round", "g2_floor", "ksi", "theta_c_rad")

# Every entry point that include/mpf.h declares (checked by the CPU tests).
EXPORTS = (
    "mpf_create", "mpf_destroy", "mpf_set_tables", "mpf_upload_phi", "mpf_init_columns",
    "mpf_upload_temp", "mpf_download_temp", "mpf_step", "mpf_download_phi", "mpf_download_apt",
    "mpf_get_stats", "mpf_reset_stats", "mpf_set_kernel_timing", "mpf_observables", "mpf_device_bytes",
    "mpf_compare_phi",
    "mpf_dense_apt", "mpf_dense_update", "mpf_dense_update_noalloc", "mpf_dense_temp",
    "mpf_malloc", "mpf_free", "mpf_memcpy_htod", "mpf_memcpy_dtoh", "mpf_memcpy_dtod", "mpf_memset",
    "mpf_device_synchronize", "mpf_last_error", "mpf_abi_version",
    "mpf_batch_create", "mpf_batch_size", "mpf_batch_set_tables", "mpf_batch_upload_phi",
    "mpf_batch_init_columns", "mpf_batch_upload_temp", "mpf_batch_download_temp", "mpf_batch_download_phi",
    "mpf_batch_download_apt", "mpf_batch_observables",
    "mpf_set_launch", "mpf_get_launch", "mpf_set_graphs", "mpf_autotune", "mpf_set_tiling", "mpf_get_tiling",
    "mpf_batch_observables_all", "mpf_probe_rays", "mpf_phase_map", "mpf_batch_phase_map",
    "mpf_upload_phi_device", "mpf_download_phi_device", "mpf_batch_download_phi_device",
)


class MPFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mpf error {code}: {msg}")
        self.code = code


class CapacityError(MPFError):
    pass


class Scalars(C.Structure):
    _fields_ = [(n, C.c_double) for n in SCALAR_NAMES]


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dim", C.c_int),
        ("n_phase", C.c_int), ("kmax", C.c_int), ("precision", C.c_int),
        ("temp_dense", C.c_int), ("temp_f32", C.c_int),
        ("enable_anisotropy", C.c_int), ("enable_torque", C.c_int), ("halt_on_kmax", C.c_int),
        ("slot_capacity", C.c_int),
        ("s", Scalars), ("cooling_rate", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("max_active", C.c_int), ("kmax_overflow", C.c_int),
        ("overflow_cells", C.c_ulonglong), ("general_cells", C.c_ulonglong),
        ("halt_step", C.c_int), ("steps_done", C.c_int), ("last_step_ms", C.c_double),
        ("merge_cells", C.c_ulonglong), ("front_cells", C.c_ulonglong),
        ("class_tables", C.c_int), ("kernel_steps", C.c_int), ("kernel_ms", C.c_double * 4),
        ("light_cells", C.c_ulonglong), ("n111_pairs", C.c_int)
    ]


_lib = None


def loaded() -> bool:
    """Whether libmpf.so (and with it a HIP runtime) is loaded in this process."""
    return _lib is not None


def lib():
    """Load libmpf.so (raises when it has not been built).  The library
    binds to whichever libamdhip64.so.7 the process loaded first: the
    system's, or PyTorch-ROCm's when torch was imported before."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)
    H = C.c_void_p
    sig = {
        "mpf_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
        "mpf_destroy": (None, [H]),
        "mpf_set_tables": (C.c_int, [H, dp, dp, dp, dp]),
        "mpf_upload_phi": (C.c_int, [H, vp, C.c_int]),
        "mpf_init_columns": (C.c_int, [H, ip, ip, C.c_int, dp, dp]),
        "mpf_upload_temp": (C.c_int, [H, dp]),
        "mpf_download_temp": (C.c_int, [H, dp]),
        "mpf_step": (C.c_int, [H, C.c_int]),
        "mpf_download_phi": (C.c_int, [H, vp, C.c_int]),
        "mpf_download_apt": (C.c_int, [H, ip, C.c_int, ip, ip]),
        "mpf_get_stats": (C.c_int, [H, C.POINTER(Stats)]),
        "mpf_reset_stats": (C.c_int, [H]),
        "mpf_set_kernel_timing": (C.c_int, [H, C.c_int]),
        "mpf_observables": (C.c_int, [H, dp, ip, dp]),
        "mpf_compare_phi": (C.c_int, [H, H, C.c_double, dp]),
        "mpf_device_bytes": (C.c_size_t, [H]),
        "mpf_dense_apt": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, C.c_int, C.c_int, vp]),
        "mpf_dense_update": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, vp, vp, vp, vp, vp, vp,
                                       C.POINTER(Scalars), C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
        "mpf_dense_update_noalloc": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, vp, vp, vp, vp, vp, vp,
                                               C.POINTER(Scalars), C.c_int, C.c_int, C.c_int, C.c_int,
                                               C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
        "mpf_dense_temp": (C.c_int, [vp, C.c_int, C.c_double, C.c_size_t, vp]),
        "mpf_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
        "mpf_free": (C.c_int, [vp]),
        "mpf_memcpy_htod": (C.c_int, [vp, vp, C.c_size_t]),
        "mpf_memcpy_dtoh": (C.c_int, [vp, vp, C.c_size_t]),
        "mpf_memcpy_dtod": (C.c_int, [vp, vp, C.c_size_t]),
        "mpf_memset": (C.c_int, [vp, C.c_int, C.c_size_t]),
        "mpf_device_synchronize": (C.c_int, []),
        "mpf_last_error": (C.c_char_p, []),
        "mpf_abi_version": (C.c_int, []),
        "mpf_batch_create": (C.c_int, [C.POINTER(Config), C.c_int, ip, C.POINTER(C.c_void_p)]),
        "mpf_batch_size": (C.c_int, [H]),
        "mpf_batch_set_tables": (C.c_int, [H, C.c_int, dp, dp, dp, dp]),
        "mpf_batch_upload_phi": (C.c_int, [H, C.c_int, vp, C.c_int]),
        "mpf_batch_init_columns": (C.c_int, [H, C.c_int, ip, ip, C.c_int, dp, dp]),
        "mpf_batch_upload_temp": (C.c_int, [H, C.c_int, dp]),
        "mpf_batch_download_temp": (C.c_int, [H, C.c_int, dp]),
        "mpf_batch_download_phi": (C.c_int, [H, C.c_int, vp, C.c_int]),
        "mpf_batch_download_apt": (C.c_int, [H, C.c_int, ip, C.c_int, ip, ip]),
        "mpf_batch_observables": (C.c_int, [H, C.c_int, dp, ip, dp]),
        "mpf_set_launch": (C.c_int, [H, C.c_int, C.c_int, C.c_int, C.c_int]),
        "mpf_get_launch": (C.c_int, [H, ip]),
        "mpf_set_graphs": (C.c_int, [H, C.c_int]),
        "mpf_set_tiling": (C.c_int, [H, C.c_int, C.c_int]),
        "mpf_get_tiling": (C.c_int, [H, ip]),
        "mpf_autotune": (C.c_int, [H, C.c_int, ip]),
        "mpf_batch_observables_all": (C.c_int, [H, dp, ip, dp]),
        "mpf_probe_rays": (C.c_int, [H, C.c_int, ip, ip, dp, dp, C.c_double, C.c_double, dp]),
        "mpf_phase_map": (C.c_int, [H, C.c_int, C.c_int, ip]),
        "mpf_upload_phi_device": (C.c_int, [H, vp, C.c_int]),
        "mpf_download_phi_device": (C.c_int, [H, vp, C.c_int]),
        "mpf_batch_download_phi_device": (C.c_int, [H, C.c_int, vp, C.c_int]),
        "mpf_batch_phase_map": (C.c_int, [H, C.c_int, C.c_int, C.c_int, ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int) -> int:
    """Raise MPFError for a negative return code; pass counts through."""
    if rc < 0:
        msg = lib().mpf_last_error().decode(errors="replace")
        if rc == MPF_ECAPACITY:
            raise CapacityError(rc, msg)
        raise MPFError(rc, msg)
    return rc


def ptr(a: np.ndarray, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


def scalars_struct(s: dict) -> Scalars:
    out = Scalars()
    for n in SCALAR_NAMES:
        setattr(out, n, float(s[n]))
    return out
