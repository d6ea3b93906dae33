"""ctypes binding of libtrident_hip.so (the C-ABI in include/trident_hip.h).

The product path is the HIP library; this module only loads it and declares
its signatures.  Loading fails loudly when the library is missing: there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtrident_hip.so")

TR_OK = 0
TR_ERR_INVALID = 1
TR_ERR_OOM = 2
TR_ERR_OVERFLOW = 3
TR_ERR_HIP = 4
TR_ROWS = 0
TR_COLS = 1

ERROR_NAMES = {
    TR_OK: "TR_OK",
    TR_ERR_INVALID: "TR_ERR_INVALID",
    TR_ERR_OOM: "TR_ERR_OOM",
    TR_ERR_OVERFLOW: "TR_ERR_OVERFLOW",
    TR_ERR_HIP: "TR_ERR_HIP",
}


class tr_csx_t(ctypes.Structure):
    """Mirror of `tr_csx_t` (and of the reference's mmio::CSX)."""

    _fields_ = [
        ("majordim", ctypes.c_int32),
        ("nrows", ctypes.c_int32),
        ("ncols", ctypes.c_int32),
        ("contig", ctypes.c_int32),
        ("nnz", ctypes.c_int64),
        ("ptr_vec", ctypes.c_void_p),
        ("idx_vec", ctypes.c_void_p),
        ("val", ctypes.c_void_p),
        ("buf", ctypes.c_void_p),
        ("buf_size", ctypes.c_size_t),
    ]


class tr_tuning_t(ctypes.Structure):
    """Mirror of `tr_tuning_t` (multiply cutovers chosen by the autotuner)."""

    _fields_ = [
        ("light_max_flops", ctypes.c_int32),
        ("window_entries", ctypes.c_int32),
        ("item_windows", ctypes.c_int32),
        ("heavy_table", ctypes.c_int32),
        ("dense_width", ctypes.c_int32),
    ]


# (name, restype, argtypes) for every entry point of include/trident_hip.h
_P = ctypes.POINTER
_csx = _P(tr_csx_t)
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t
SIGNATURES = [
    ("tr_ctx_create", ctypes.c_int, [_P(_vp), ctypes.c_int, _vp]),
    ("tr_ctx_destroy", ctypes.c_int, [_vp]),
    ("tr_last_error", ctypes.c_char_p, [_vp]),
    ("tr_ctx_synchronize", ctypes.c_int, [_vp]),
    ("tr_ctx_stream", _vp, [_vp]),
    ("tr_spgemm", ctypes.c_int, [_vp, _csx, _csx, _csx]),
    ("tr_spgemm_panels", ctypes.c_int, [_vp, _csx, _csx, _i64, _csx, _P(_i32), _i32, _P(_i32)]),
    ("tr_sp_mma", ctypes.c_int, [_vp, _csx, _csx, _csx]),
    ("tr_spadd", ctypes.c_int, [_vp, _csx, _csx]),
    ("tr_spadd_into", ctypes.c_int, [_vp, _csx, _csx, _csx, _i64, _P(_i64)]),
    ("tr_spadd_into_noalloc", ctypes.c_int, [_vp, _csx, _csx, _csx, _i64, _P(_i64)]),
    ("tr_symbolic", ctypes.c_int, [_vp, _csx, _csx, _P(_i64), _P(_sz)]),
    ("tr_spgemm_flops", ctypes.c_int, [_vp, _csx, _csx, _P(_i64)]),
    ("tr_tuning_default", ctypes.c_int, [_P(tr_tuning_t)]),
    ("tr_set_default_tuning", ctypes.c_int, [_P(tr_tuning_t)]),
    ("tr_ctx_set_tuning", ctypes.c_int, [_vp, _P(tr_tuning_t)]),
    ("tr_ctx_get_tuning", ctypes.c_int, [_vp, _P(tr_tuning_t)]),
    ("tr_batch_create", ctypes.c_int, [_vp, _i32, _csx, _csx, _csx, _i32, _P(_vp)]),
    ("tr_batch_run", ctypes.c_int, [_vp]),
    ("tr_batch_result", ctypes.c_int, [_vp, _i32, _csx]),
    ("tr_batch_result_panels", ctypes.c_int, [_vp, _i32, _i64, _csx, _P(_i32), _i32, _P(_i32)]),
    ("tr_batch_check", ctypes.c_int, [_vp]),
    ("tr_batch_stats", ctypes.c_int, [_vp, _P(_i64), _P(_i64), _P(_i64)]),
    ("tr_batch_destroy", ctypes.c_int, [_vp]),
    ("tr_server_create", ctypes.c_int, [ctypes.c_int, _sz, _i32, _i32, _P(_vp)]),
    ("tr_server_register", ctypes.c_int, [_vp, _csx, _P(_i32)]),
    ("tr_server_submit", ctypes.c_int, [_vp, _i32, _i32, _i32, _P(_i64)]),
    ("tr_server_wait", ctypes.c_int, [_vp, _i64, _csx, _P(ctypes.c_double)]),
    ("tr_server_release", ctypes.c_int, [_vp, _i64]),
    ("tr_server_stats", ctypes.c_int, [_vp, _P(_i64), _P(_i64), _P(ctypes.c_double), _P(ctypes.c_double),
                                       _P(ctypes.c_double)]),
    ("tr_server_reset_stats", ctypes.c_int, [_vp]),
    ("tr_server_destroy", ctypes.c_int, [_vp]),
    ("tr_bitmask_gen", ctypes.c_int, [_vp, _vp, _i32, _vp]),
    ("tr_bitmask_and", ctypes.c_int, [_vp, _vp, _vp, _i32, _vp]),
    ("tr_compress", ctypes.c_int, [_vp, _csx, _vp, _csx]),
    ("tr_csx_transpose", ctypes.c_int, [_vp, _csx, _csx]),
    ("tr_csc_to_csr", ctypes.c_int, [_vp, _csx, _csx]),
    ("tr_csc_to_csr_into", ctypes.c_int, [_vp, _csx, _vp, _vp, _vp]),
    ("tr_csx_sort_rows", ctypes.c_int, [_vp, _csx]),
    ("tr_csx_validate", ctypes.c_int, [_vp, _csx, _P(_i32)]),
    ("tr_csx_free", ctypes.c_int, [_vp, _csx]),
    ("tr_csx_upload", ctypes.c_int, [_vp, _i32, _i32, _i32, _i64, _vp, _vp, _vp, _csx]),
    ("tr_csx_download", ctypes.c_int, [_vp, _csx, _vp, _vp, _vp]),
    ("tr_malloc", ctypes.c_int, [_vp, _P(_vp), _sz]),
    ("tr_free", ctypes.c_int, [_vp, _vp]),
    ("tr_memcpy_h2d", ctypes.c_int, [_vp, _vp, _vp, _sz]),
    ("tr_memcpy_d2h", ctypes.c_int, [_vp, _vp, _vp, _sz]),
    ("tr_memcpy_d2d", ctypes.c_int, [_vp, _vp, _vp, _sz]),
    ("tr_memset", ctypes.c_int, [_vp, _vp, ctypes.c_int, _sz]),
    ("tr_device_malloc", ctypes.c_int, [ctypes.c_int, _P(_vp), _sz]),
    ("tr_device_free", ctypes.c_int, [_vp]),
    ("tr_event_create", ctypes.c_int, [_P(_vp)]),
    ("tr_event_record", ctypes.c_int, [_vp, _vp]),
    ("tr_ctx_wait_event", ctypes.c_int, [_vp, _vp]),
    ("tr_event_synchronize", ctypes.c_int, [_vp]),
    ("tr_event_destroy", ctypes.c_int, [_vp]),
    ("tr_pool_trim", ctypes.c_int, [_vp]),
    ("tr_pool_stats", ctypes.c_int, [_vp, _P(_sz), _P(_sz)]),
    ("tr_pool_set_limit", ctypes.c_int, [_vp, _sz]),
    ("tr_ctx_set_timing", ctypes.c_int, [_vp, ctypes.c_int]),
    ("tr_ctx_reset_timing", ctypes.c_int, [_vp]),
    ("tr_ctx_timing_count", ctypes.c_int, [_vp]),
    ("tr_ctx_timing_get", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _P(_i64),
                                         _P(ctypes.c_double)]),
    ("tr_calibration_copy", ctypes.c_int, [_vp, _vp, _vp, _sz]),
    ("tr_version", ctypes.c_char_p, []),
]

_lib = None


def load() -> ctypes.CDLL:
    """Load libtrident_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: the HIP library must be built "
                               "(__graft_entry__.build()); there is no CPU fallback")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class TridentError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code
