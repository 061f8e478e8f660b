"""ctypes binding of ``libakitsushima_amd.so`` (the C ABI in
``include/akitsushima_amd.h``).

The library is the only compute path: there is no CPU fallback.  Loading
fails with a RuntimeError when the library has not been built, and every
entry point fails with AK_ERR_NO_DEVICE when no GPU is visible.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libakitsushima_amd.so")

AK_OK = 0
AK_ERR_ARG = 1
AK_ERR_MIN_IMAGE = 2
AK_ERR_HIP = 3
AK_ERR_CAPACITY = 4
AK_ERR_STATE = 5
AK_ERR_NO_DEVICE = 6

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_vp = ctypes.c_void_p
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pf64 = ctypes.POINTER(ctypes.c_double)
_pu8 = ctypes.POINTER(ctypes.c_uint8)
_pf32 = ctypes.POINTER(ctypes.c_float)
_handle = ctypes.c_void_p
_phandle = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "ak_abi_version": (ctypes.c_int, []),
    "ak_last_error": (ctypes.c_char_p, []),
    "ak_device_count": (ctypes.c_int, [_pi32]),
    "ak_tersoff_create": (ctypes.c_int, [_i32, _pf64, _pf64, _f64, _phandle]),
    "ak_sw_create": (ctypes.c_int, [_pf64, _phandle]),
    "ak_sw_multi_create": (ctypes.c_int, [_i32, _pf64, _pf64, _pf64, _f64, _phandle]),
    "ak_eam_create": (ctypes.c_int, [_i32, _i32, _pf64, _pi32, _f64, _phandle]),
    "ak_eam_tables_create": (ctypes.c_int, [_i32, _i32, _pf64, _pi32, _i64, _pf64, _pf64, _f64,
                                            _phandle]),
    "ak_eam_eval": (ctypes.c_int, [ctypes.c_void_p, _i32, _i64, _pf64, _pf64, _pf64]),
    "ak_potential_destroy": (ctypes.c_int, [_handle]),
    "ak_potential_set_lanes": (ctypes.c_int, [_handle, _i32]),
    "ak_potential_set_launch": (ctypes.c_int, [_handle, _i32, _i32, _i32]),
    "ak_autotune_batched": (ctypes.c_int, [_handle, _i32, _pi64, _pf64, _pf64, _pi32, _i32,
                                           _pi32, _pf64, _i32, _pi32, _pf64, _pi32]),
    "ak_nlist_create": (ctypes.c_int, [_f64, _f64, _phandle]),
    "ak_nlist_destroy": (ctypes.c_int, [_handle]),
    "ak_nlist_build": (ctypes.c_int, [_handle, _i32, _pf64, _pf64, _pf64]),
    "ak_nlist_check_and_rebuild": (ctypes.c_int, [_handle, _i32, _pf64, _pf64, _pf64, _pi32]),
    "ak_nlist_candidate_pairs": (ctypes.c_int, [_handle, _pi64, _pi32, _pi32, _i64]),
    "ak_nlist_counters": (ctypes.c_int, [_handle, _pi64, _pi64]),
    "ak_compute": (ctypes.c_int, [_handle, _handle, _i32, _pf64, _pf64, _pf64, _pi32, _i32,
                                  _pf64, _pf64, _pf64]),
    "ak_compute_device": (ctypes.c_int, [_handle, _handle, _i32, _pf64, _pf64, _vp, _vp, _vp,
                                         _vp, _vp]),
    "ak_compute_device_noalloc": (ctypes.c_int, [_handle, _handle, _i32, _pf64, _pf64, _vp,
                                                 _vp, _vp, _vp, _vp]),
    "ak_compute_batched": (ctypes.c_int, [_handle, _i32, _pi64, _pf64, _pf64, _pf64, _pi32,
                                          _pf64, _pf64, _pf64]),
    "ak_compute_batched_device": (ctypes.c_int, [_handle, _i32, _pi64, _pf64, _vp, _vp, _vp,
                                                 _vp, _i32, _vp]),
    "ak_md_nve": (ctypes.c_int, [_handle, _i32, _pf64, _pf64, _pf64, _pf64, _pf64, _pi32, _f64,
                                 _i32, _f64, _i32, _pf64, _pf64, _pi32, _pf64]),
    "ak_md_run": (ctypes.c_int, [_handle, _i32, _pf64, _pf64, _pf64, _pf64, _pf64, _pi32, _f64,
                                 _i32, _f64, _i32, _pf64, _pf64, _pu8, _i32, _i32, _pf64, _pf64,
                                 _pf64, _pf64, _pf64, _pf64, _pf64, _pf64, _pf64, _pi32, _pf64,
                                 _pf64, _pi32, _i32, _i32, _pf64, _i32, _pf64, _pf32, _pi32,
                                 _pf64]),
    "ak_nemd_run": (ctypes.c_int, [_handle, _i32, _pf64, _pf64, _pf64, _pf64, _pf64, _pi32, _f64,
                                   _i32, _f64, _i32, _pu8, _f64, _f64, _f64, _f64, _pf64, _i32,
                                   _i32, _i32, _i32, _pf64, _pf64, _pf64, _pf64, _pf64, _pi32,
                                   _pf64]),
    "ak_site_terms": (ctypes.c_int, [_handle, _handle, _i32, _pf64, _pf64, _pf64, _pi32, _pf64,
                                     _i64, _pi32, _pi32, _pf64, _pf64,
                                     ctypes.POINTER(ctypes.c_int64)]),
    "ak_heat_flux": (ctypes.c_int, [_handle, _handle, _i32, _pf64, _pf64, _pf64, _pf64, _pf64,
                                    _pi32, _pf64, _pf64]),
    "ak_profile_enable": (ctypes.c_int, [_handle, _i32]),
    "ak_profile_read": (ctypes.c_int, [_handle, _pf64, _pf64, _pi64]),
    "ak_tersoff_loss_grad": (ctypes.c_int, [_handle, _i32, _pi64, _pf64, _pf64, _pf64, _pf64,
                                            _pf64, _pf64, _pf64, _pf64, _pf64, _pf64, _pf64,
                                            _pf64]),
    "ak_tersoff_fit_load": (ctypes.c_int, [_handle, _i32, _pi64, _pf64, _pf64, _pf64, _pf64,
                                           _pf64]),
    "ak_tersoff_fit_terms": (ctypes.c_int, [_handle, _i32, _pf64, _i32, _pf64]),
    "ak_tersoff_fit_eval": (ctypes.c_int, [_handle, _pf64, _pf64, _pf64, _pf64, _pf64, _pf64,
                                           _pf64]),
}

_lib = None
_lock = threading.Lock()


class EngineError(RuntimeError):
    """A failure reported by the HIP engine."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """The loaded library (raises RuntimeError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    "akitsushima_amd: the HIP library is not built "
                    f"({LIB_PATH} missing); run `python -m akitsushima_amd.build`")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def check(status):
    if status == AK_OK:
        return
    msg = lib().ak_last_error().decode(errors="replace")
    if status in (AK_ERR_ARG, AK_ERR_MIN_IMAGE):
        raise ValueError(msg)
    raise EngineError(status, msg)


def device_count():
    n = _i32(0)
    lib().ak_device_count(ctypes.byref(n))
    return int(n.value)


def f64(a):
    """C-contiguous float64 copy/view and its pointer."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_pf64)


def i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_pi32)


def i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(_pi64)
