"""Loader for the HIP library libxrd_tomo.so (ctypes, C ABI of include/xrd_tomo.h).

The library must have been built (``python -m xrd_tomography_amd.build`` or
``__graft_entry__.build()``).  There is no fallback: a missing or unloadable
library raises immediately.
"""
from __future__ import annotations

import ctypes
import os

from .descriptors import Bf16TablesStruct, Discretization, Geometry

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libxrd_tomo.so")

XRD_OK = 0
ERRORS = {
    -1: "XRD_ERR_INVALID_ARG", -2: "XRD_ERR_NULL_POINTER", -3: "XRD_ERR_DIMS",
    -4: "XRD_ERR_ALIGN", -5: "XRD_ERR_HIP", -6: "XRD_ERR_ALLOC", -7: "XRD_ERR_UNSUPPORTED",
    -8: "XRD_ERR_NONFINITE",
}
DTYPE_F32, DTYPE_BF16, DTYPE_F16, DTYPE_BF24 = 0, 1, 2, 3
MLEM_FUSED, MLEM_UNFUSED = 0, 1

# Exported symbols (kept in sync with include/xrd_tomo.h; a test checks both).
EXPORTS = (
    "xrd_geometry_reference", "xrd_discretization_reference", "xrd_num_rows", "xrd_num_cols",
    "xrd_context_create", "xrd_context_destroy", "xrd_context_workspace_bytes",
    "xrd_last_error_string", "xrd_version", "xrd_forward", "xrd_trace_rays", "xrd_project",
    "xrd_backproject", "xrd_sensitivity", "xrd_convert_matrix", "xrd_mlem_plan_create",
    "xrd_mlem_plan_iterate", "xrd_mlem_plan_loglik", "xrd_mlem_plan_sensitivity",
    "xrd_mlem_plan_destroy", "xrd_forward_coeffs", "xrd_vjp_coeffs", "xrd_poisson_nll",
    "xrd_convert_matrix_scaled", "xrd_mlem_plan_autotune", "xrd_mlem_plan_set_launch",
    "xrd_mlem_batch_create", "xrd_mlem_batch_iterate", "xrd_mlem_batch_set_launch",
    "xrd_mlem_batch_destroy", "xrd_mlem_plan_time_pass", "xrd_mlem_batch_autotune",
    "xrd_mlem_batch_time_pass", "xrd_mlem_batch_groups", "xrd_forward_ct", "xrd_vjp_ct",
    "xrd_forward_slice", "xrd_mlem_iterate_static", "xrd_forward_bf16",
    "xrd_forward_set_launch", "xrd_forward_autotune",
)
MLEM_STATIC_SCRATCH_BYTES = 4259840  # XRD_MLEM_STATIC_SCRATCH_BYTES


class XrdError(RuntimeError):
    def __init__(self, code: int, fn: str, msg: str):
        super().__init__(f"{fn} failed with {ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None
_vp = ctypes.c_void_p
_i32, _i64, _dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double


def _declare(L):
    G, D = ctypes.POINTER(Geometry), ctypes.POINTER(Discretization)
    sig = {
        "xrd_geometry_reference": (ctypes.c_int, [_i32, _i32, _i32, _i32, _i32, G]),
        "xrd_discretization_reference": (ctypes.c_int, [_i32, _i32, _i32, _i32, D]),
        "xrd_num_rows": (_i64, [G, D]),
        "xrd_num_cols": (_i64, [D]),
        "xrd_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
        "xrd_context_destroy": (ctypes.c_int, [_vp]),
        "xrd_context_workspace_bytes": (_i64, [_vp]),
        "xrd_last_error_string": (ctypes.c_char_p, []),
        "xrd_version": (ctypes.c_int, []),
        "xrd_forward": (ctypes.c_int, [_vp, G, D, _vp, _vp, _vp, _vp, _vp, _vp,
                                       _dbl, _dbl, _dbl, _dbl, _vp, _vp, _vp, _vp]),
        "xrd_forward_slice": (ctypes.c_int, [_vp, G, D, _vp, _vp, _vp, _vp, _vp, _i32, _i32,
                                             _vp, _dbl, _dbl, _dbl, _dbl, _vp, _vp, _vp, _vp]),
        "xrd_forward_bf16": (ctypes.c_int, [_vp, G, D, _vp, ctypes.POINTER(Bf16TablesStruct),
                                            _vp, _i32, _i32, _vp, _dbl, _dbl, _dbl, _dbl, _vp,
                                            _vp, _vp, _vp]),
        "xrd_forward_set_launch": (ctypes.c_int, [_vp, _i32, _i32]),
        "xrd_forward_autotune": (ctypes.c_int, [_vp, G, D, _vp, _vp, _vp, _vp, _vp, _vp, _dbl,
                                                _dbl, _dbl, _dbl, _i32, _vp,
                                                ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                                ctypes.POINTER(ctypes.c_float)]),
        "xrd_trace_rays": (ctypes.c_int, [_vp, _i32, _i32, _dbl, _i64, _vp, _vp, _vp, _vp,
                                          _i32, _vp, _i32, _vp, _vp, _vp]),
        "xrd_project": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _i64, _i32, _vp, _vp, _vp]),
        "xrd_backproject": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _i64, _i32, _vp, _vp, _vp]),
        "xrd_sensitivity": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _i64, _i32, _vp, _vp]),
        "xrd_convert_matrix": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_int, _vp, _vp]),
        "xrd_mlem_plan_create": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _i64, _i32, _vp, _vp,
                                                _vp, _vp, ctypes.c_int, _vp,
                                                ctypes.POINTER(_vp)]),
        "xrd_mlem_plan_autotune": (ctypes.c_int, [_vp, _i32, _vp, ctypes.POINTER(_i32),
                                                  ctypes.POINTER(_i32),
                                                  ctypes.POINTER(ctypes.c_float)]),
        "xrd_mlem_plan_set_launch": (ctypes.c_int, [_vp, _i32, _i32]),
        "xrd_mlem_plan_time_pass": (ctypes.c_int, [_vp, _i32, _vp, ctypes.POINTER(ctypes.c_float)]),
        "xrd_mlem_batch_create": (ctypes.c_int, [_vp, _i32, ctypes.c_int, ctypes.POINTER(_vp),
                                                 ctypes.POINTER(_i64), ctypes.POINTER(_i32),
                                                 ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                 ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i32,
                                                 _vp, ctypes.POINTER(_vp)]),
        "xrd_mlem_batch_iterate": (ctypes.c_int, [_vp, _i32, ctypes.c_int, _vp]),
        "xrd_mlem_batch_set_launch": (ctypes.c_int, [_vp, _i32, _i32]),
        "xrd_mlem_batch_destroy": (ctypes.c_int, [_vp]),
        "xrd_mlem_batch_autotune": (ctypes.c_int, [_vp, _i32, _vp, ctypes.POINTER(_i32),
                                                   ctypes.POINTER(_i32),
                                                   ctypes.POINTER(ctypes.c_float)]),
        "xrd_mlem_batch_time_pass": (ctypes.c_int, [_vp, _i32, _vp,
                                                    ctypes.POINTER(ctypes.c_float)]),
        "xrd_mlem_batch_groups": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_i32),
                                                 ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                                 ctypes.POINTER(_i32)]),
        "xrd_convert_matrix_scaled": (ctypes.c_int, [_vp, _vp, _i64, _i32, ctypes.c_int, _vp,
                                                     _vp, _vp]),
        "xrd_mlem_plan_iterate": (ctypes.c_int, [_vp, _i32, ctypes.c_int, _vp]),
        "xrd_mlem_iterate_static": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _i64, _i32, _vp,
                                                   _vp, _vp, _vp, _i32, _vp]),
        "xrd_mlem_plan_loglik": (ctypes.c_int, [_vp, ctypes.POINTER(_dbl), _vp]),
        "xrd_mlem_plan_sensitivity": (_vp, [_vp]),
        "xrd_mlem_plan_destroy": (ctypes.c_int, [_vp]),
        "xrd_forward_coeffs": (ctypes.c_int, [_vp, G, D, _vp, _vp, _vp, _i32, _vp, _i32, _vp,
                                              _vp, _vp, _vp]),
        "xrd_vjp_coeffs": (ctypes.c_int, [_vp, G, D, _vp, _vp, _vp, _i32, _vp, _i32, _vp, _vp,
                                          _vp, _vp, _vp, _vp]),
        "xrd_poisson_nll": (ctypes.c_int, [_vp, _i64, _vp, _vp, _vp, _vp,
                                           ctypes.POINTER(_dbl), _vp]),
        "xrd_forward_ct": (ctypes.c_int, [_vp, G, D, _vp, _i32, _vp, _vp, _vp, _vp]),
        "xrd_vjp_ct": (ctypes.c_int, [_vp, G, D, _vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m xrd_tomography_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def check(rc: int, fn: str):
    if rc != XRD_OK:
        msg = lib().xrd_last_error_string().decode(errors="replace")
        raise XrdError(rc, fn, msg)
