"""ctypes binding of libswd.so (the C ABI declared in include/swd_abi.h).

The library is the only compute path: when it is missing, every entry point
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswd.so")

SWD_OK = 0
ERRORS = {-2: "SWD_ERR_ARG", -3: "SWD_ERR_CFL", -4: "SWD_ERR_NOMEM", -5: "SWD_ERR_HIP",
          -6: "SWD_ERR_PICK", -7: "SWD_ERR_NODEV", -8: "SWD_ERR_BUSY"}


class SwdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double

# name -> (restype, argtypes); kept in the order of include/swd_abi.h
SIGNATURES = {
    "swd_version": (ctypes.c_char_p, []),
    "swd_last_error": (ctypes.c_char_p, []),
    "swd_device_info": (_I, [_P, ctypes.c_char_p, _I]),
    "swd_staggerfd_f32": (_I, [_P] * 15),
    "swd_staggerfd_f64": (_I, [_P] * 15),
    "swd_e2drtm_f32": (_I, [_P] * 16),
    "swd_e2drtm_f64": (_I, [_P] * 16),
    "swd_set_fd_path": (_I, [_I]),
    "swd_set_fd_time_block": (_I, [_I, _I]),
    "swd_prepare_model": (_I, [_P, _P, _I, _I, _I, _D, _D, _I] + [_P] * 7),
    "swd_ricker": (_I, [_D, _D, _I, _P]),
    "swd_radon": (_I, [_I, _P, _I, _I, _I, _D, _D, _I, _D, _D, _D, _D, _D, _I, _I, _I, _P, _P, _P, _P, _P]),
    "swd_radon_f32": (_I, [_I, _P, _I, _I, _I, _D, _D, _I, _D, _D, _D, _D, _D, _I, _I, _I, _P, _P, _P, _P, _P]),
    "swd_lhdisppick": (_I, [_P, _I, _I, _D, _P, _I, _I, _P, _P]),
    "swd_lhdisppick_f64": (_I, [_P, _I, _I, _D, _P, _I, _I, _P, _P]),
    "swd_lhdisppick_noalloc": (_I, [_P, _I, _I, _D, _P, _I, _I, _P, _P]),
    "swd_adwdgrad1": (_I, [_I, _I, _I, _I, _I, _D, _I, _I, _D, _P, _P, _P, _P, _I, _I, _P, _P, _I, _I, _I, _D,
                           _D, _D, _I, _I, _P]),
    "swd_engine_create": (_I, [_P, _P]),
    "swd_engine_destroy": (_I, [_P]),
    "swd_engine_set_model": (_I, [_P, _P, _P]),
    "swd_engine_set_wavelet": (_I, [_P, _P]),
    "swd_engine_set_observed_picks": (_I, [_P, _P, _P]),
    "swd_engine_observe": (_I, [_P, _P, _P]),
    "swd_engine_gradient": (_I, [_P, _P, _P, _P, _P, _P]),
    "swd_engine_forward_picks": (_I, [_P, _P, _P, _P, _P]),
    "swd_engine_set_fd_path": (_I, [_P, _I]),
    "swd_engine_set_graph": (_I, [_P, _I]),
    "swd_engine_set_variant": (_I, [_P, _P]),
    "swd_engine_run": (_I, [_P, _I]),
    "swd_engine_set_picks": (_I, [_P, _P, _P]),
    "swd_engine_set_tile_rows": (_I, [_P, _I]),
    "swd_engine_set_time_block": (_I, [_P, _I, _I]),
    "swd_engine_set_phase_time_block": (_I, [_P, _I, _I, _I]),
    "swd_engine_autotune": (_I, [_P, _I, _P, _P, _I, _P]),
    "swd_batch_autotune": (_I, [_P, _I, _I, _P, _P, _I, _P]),
    "swd_engine_info": (_I, [_P, _P, _I]),
    "swd_engine_timings": (_I, [_P, _P, _I]),
    "swd_engine_read": (_I, [_P, _I, _P, ctypes.c_longlong]),
    "swd_batch_gradient": (_I, [_P, _I, _P, _P]),
    "swd_problem_npair": (_I, [_P, _P]),
    "swd_set_radon_precision": (_I, [_I]),
    "swd_engine_iterate": (_I, [_P, _P]),
    "swd_engine_read_model": (_I, [_P, _P]),
    "swd_smooth2a": (_I, [_P, _I, _I, _I, _I, _P]),
    "swd_dispersion_create": (_I, [_P, _P]),
    "swd_dispersion_destroy": (_I, [_P]),
    "swd_dispersion_info": (_I, [_P, _P, _I]),
    "swd_dispersion_set_gathers": (_I, [_P, _P]),
    "swd_dispersion_set_observed_picks": (_I, [_P, _P, _P]),
    "swd_dispersion_set_variant": (_I, [_P, _P]),
    "swd_dispersion_pick": (_I, [_P, _P, _P, _P, _P, _P]),
    "swd_dispersion_adjoint": (_I, [_P, ctypes.c_double, _P, _P]),
    "swd_calib_copy": (_I, [ctypes.c_longlong, _I, _P]),
    "swd_copy_peak": (_I, [ctypes.c_longlong, _I, _P]),
    "swd_server_create": (_I, [ctypes.c_longlong, _I, _I, _I, _P]),
    "swd_server_destroy": (_I, [_P]),
    "swd_server_submit": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "swd_server_wait": (_I, [_P, ctypes.c_longlong, ctypes.c_double, _P, _P, _P, _P]),
    "swd_server_stats": (_I, [_P, _P, _I]),
    "swd_server_set_wave_slots": (_I, [_P, ctypes.c_longlong]),
    "swd_server_set_deadline_slack": (_I, [_P, _D]),
    "swd_server_cancel": (_I, [_P, ctypes.c_longlong]),
}

_lock = threading.Lock()
_lib = None


def load() -> ctypes.CDLL:
    """Load libswd.so (raises if it was not built).  SWD_LIBRARY names another
    build of the same library instead (the sanitized one, libswd_san.so)."""
    global _lib
    with _lock:
        if _lib is None:
            path = os.environ.get("SWD_LIBRARY") or LIB_PATH
            if not os.path.exists(path):
                raise RuntimeError(f"{path} is missing: run `python -m swd_amd.build` first")
            L = ctypes.CDLL(path)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != SWD_OK:
        msg = load().swd_last_error()
        raise SwdError(rc, msg.decode() if msg else "")


def ptr(a) -> ctypes.c_void_p:
    """Data pointer of a numpy array; the returned object keeps the array alive
    for as long as it exists (so temporaries survive until the C call returns)."""
    if a is None:
        return ctypes.c_void_p(None)
    return a.ctypes.data_as(ctypes.c_void_p)


def fortran(a, dtype) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def gpu_available() -> bool:
    try:
        n = ctypes.c_int(0)
        arch = ctypes.create_string_buffer(64)
        rc = load().swd_device_info(ctypes.byref(n), arch, 64)
        return rc == SWD_OK and n.value > 0
    except Exception:
        return False
