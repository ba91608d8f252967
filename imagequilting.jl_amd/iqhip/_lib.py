"""ctypes binding of libiqhip.so (the C ABI declared in include/iqhip.h).

There is no fallback: if the HIP library is missing or no GPU is visible,
every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libiqhip.so")

IQ_OK = 0
IQ_F32 = 0
IQ_F64 = 1

STATUS = {
    -1: "IQ_ERR_ARG",
    -2: "IQ_ERR_HIP",
    -3: "IQ_ERR_NOCAND",
    -4: "IQ_ERR_NOTINDB",
    -5: "IQ_ERR_ITER",
    -6: "IQ_ERR_UNSUPPORTED",
    -7: "IQ_ERR_NOMEM",
}

# every symbol include/iqhip.h declares
EXPORTS = (
    "iq_device_count", "iq_ctx_create", "iq_ctx_destroy", "iq_last_error", "iq_ctx_synchronize",
    "iq_host_alloc", "iq_host_free",
    "iq_imfilter", "iq_fastdistance", "iq_band_pick", "iq_relaxation", "iq_taumodel", "iq_sample",
    "iq_graphcut", "iq_plan_create", "iq_plan_destroy", "iq_plan_run", "iq_plan_results",
    "iq_plan_create_multi", "iq_plan_last_time", "iq_plan_profile", "iq_plan_scan_stats", "iq_plan_scan_tuning", "iq_plan_cut_time", "iq_plan_cut_stats",
    "iq_simulate", "iq_prepare", "iq_prepare_soft", "iq_finddisabled",
    "iq_plan_learn", "iq_plan_set_proposal", "iq_plan_set_check_sample", "iq_plan_proposal_stats", "iq_plan_encoder", "iq_plan_select_stats",
    "iq_plan_validate", "iq_scan_error_bound", "iq_plan_bank_stats", "iq_graphcut_noalloc",
    "iq_plan_learn_conv", "iq_plan_conv_stats", "iq_plan_conv_params", "iq_plan_conv_grad", "iq_plan_bank_rows",
)

IQ_CONV_NPARAM = 29441  # include/iqhip.h

IQ_NOALLOC_MAX_NODES = 16384  # include/iqhip.h


class IQError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("ti", ctypes.c_void_p),
        ("ti_dtype", ctypes.c_int),
        ("tisize", ctypes.c_int64 * 3),
        ("tilesize", ctypes.c_int64 * 3),
        ("ovlsize", ctypes.c_int64 * 3),
        ("spacing", ctypes.c_int64 * 3),
        ("ntiles", ctypes.c_int64 * 3),
        ("padsize", ctypes.c_int64 * 3),
        ("disabled", ctypes.c_void_p),
        ("path", ctypes.c_void_p),
        ("npath", ctypes.c_int64),
        ("skipped", ctypes.c_void_p),
        ("hardgrid", ctypes.c_void_p),
        ("naux", ctypes.c_int),
        ("aux", ctypes.c_void_p),
        ("auxti", ctypes.c_void_p),
        ("tol", ctypes.c_double),
        ("nreal", ctypes.c_int64),
    ]


class ConvOpts(ctypes.Structure):
    _fields_ = [
        ("nqueries", ctypes.c_int64),
        ("topn", ctypes.c_int),
        ("nrandom", ctypes.c_int),
        ("steps", ctypes.c_int),
        ("batch", ctypes.c_int),
        ("lr", ctypes.c_float),
        ("lr_head", ctypes.c_float),
        ("seed", ctypes.c_uint64),
        ("val_k", ctypes.c_int64),
        ("nhard", ctypes.c_int),
    ]


_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH):
    """Load libiqhip.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise IQError(-2, f"{path} not built: run __graft_entry__.build() or make -C imagequilting.jl_amd/csrc")
        L = ctypes.CDLL(path)
        vp, i64, f64, c_int = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
        P = ctypes.POINTER
        sig = {
            "iq_device_count": [P(c_int)],
            "iq_ctx_create": [c_int, P(vp)],
            "iq_ctx_destroy": [vp],
            "iq_ctx_synchronize": [vp],
            "iq_host_alloc": [vp, ctypes.c_size_t, P(vp)],
            "iq_host_free": [vp, vp],
            "iq_imfilter": [vp, vp, vp, vp, vp, vp],
            "iq_fastdistance": [vp, vp, c_int, vp, vp, vp, vp, vp],
            "iq_band_pick": [vp, vp, i64, f64, vp, vp],
            "iq_relaxation": [vp, vp, vp, c_int, i64, f64, vp, vp],
            "iq_taumodel": [vp, vp, i64, vp, vp, c_int, i64, vp],
            "iq_sample": [vp, vp, i64, f64, vp],
            "iq_graphcut": [vp, vp, vp, c_int, vp, c_int, vp],
            "iq_graphcut_noalloc": [vp, vp, vp, c_int, vp, c_int, vp],
            "iq_plan_learn_conv": [vp, vp, vp],
            "iq_plan_conv_stats": [vp, vp],
            "iq_plan_conv_params": [vp, vp],
            "iq_plan_conv_grad": [vp, vp, c_int, vp, i64, i64, vp, vp, vp, vp, vp, vp],
            "iq_plan_bank_rows": [vp, vp, c_int, vp, i64, vp],
            "iq_plan_create": [vp, vp, P(vp)],
            "iq_plan_create_multi": [vp, vp, c_int, P(vp)],
            "iq_plan_validate": [vp, c_int],
            "iq_plan_bank_stats": [vp, vp, vp, vp],
            "iq_scan_error_bound": [i64, c_int, P(f64)],
            "iq_plan_destroy": [vp],
            "iq_plan_run": [vp, vp, vp, vp],
            "iq_plan_results": [vp, vp, vp, vp, vp],
            "iq_plan_last_time": [vp, P(f64)],
            "iq_plan_profile": [vp, c_int],
            "iq_plan_scan_stats": [vp, vp, vp, vp, vp],
            "iq_plan_scan_tuning": [vp, vp, vp],
            "iq_plan_cut_time": [vp, vp, vp],
            "iq_plan_cut_stats": [vp, vp, vp, vp],
            "iq_simulate": [vp, vp, vp, vp, vp, vp, vp],
            "iq_prepare": [vp, c_int, vp, vp, vp],
            "iq_plan_learn": [vp, vp, i64, c_int, ctypes.c_uint64],
            "iq_plan_set_proposal": [vp, c_int, i64, c_int, f64],
            "iq_plan_set_check_sample": [vp, c_int],
            "iq_plan_proposal_stats": [vp, vp],
            "iq_plan_encoder": [vp, c_int, vp, vp, vp, vp],
            "iq_plan_select_stats": [vp, vp, vp, vp],
            "iq_prepare_soft": [vp, c_int, vp, vp, vp, vp],
            "iq_finddisabled": [vp, c_int, vp, vp, vp, vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = c_int
        L.iq_last_error.argtypes = [vp]
        L.iq_last_error.restype = ctypes.c_char_p
        _lib = L
        return L


def check(rc: int, ctx=None) -> None:
    if rc != IQ_OK:
        L = load()
        msg = L.iq_last_error(ctx)
        raise IQError(rc, msg.decode() if msg else "")
