"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, and only as the checker; the product library never calls it.
The algorithms and their reference citations are in oracle.c / oracle.h.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run `make -C oracle`")
        lib = ctypes.CDLL(path)
        lib.or_bitmask_gen.argtypes = [_i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
        lib.or_bitmask_gen.restype = None
        lib.or_bitmask_and.argtypes = [_u64p, _u64p, ctypes.c_int32, _u64p]
        lib.or_bitmask_and.restype = None
        lib.or_compress.argtypes = [ctypes.c_int32, _i32p, _i32p, _f64p, _u64p, _i32p, _i32p, _f64p]
        lib.or_compress.restype = ctypes.c_int64
        lib.or_spgemm_rownnz.argtypes = [ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _i32p, _i32p, _i64p]
        lib.or_spgemm_rownnz.restype = ctypes.c_int64
        lib.or_spgemm_numeric.argtypes = [ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _f64p, _i32p, _i32p,
                                          _f64p, _i64p, _i32p, _f64p]
        lib.or_spgemm_numeric.restype = None
        lib.or_spgemm_flops.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p]
        lib.or_spgemm_flops.restype = ctypes.c_int64
        lib.or_spadd_rownnz.argtypes = [ctypes.c_int32, _i32p, _i32p, _i32p, _i32p, _i64p]
        lib.or_spadd_rownnz.restype = ctypes.c_int64
        lib.or_spadd_numeric.argtypes = [ctypes.c_int32, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p, _i64p,
                                         _i32p, _f64p]
        lib.or_spadd_numeric.restype = None
        lib.or_transpose.argtypes = [ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p]
        lib.or_transpose.restype = None
        lib.or_rows_sorted.argtypes = [ctypes.c_int32, _i32p, _i32p]
        lib.or_rows_sorted.restype = ctypes.c_int32
        lib.or_num_threads.argtypes = []
        lib.or_num_threads.restype = ctypes.c_int32
        _lib = lib
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def bitmask_gen(ptr, n: int, word_bits: int = 64) -> np.ndarray:
    ptr = _c(ptr, np.int32)
    dt = {8: np.uint8, 32: np.uint32, 64: np.uint64}[word_bits]
    out = np.zeros((n + word_bits - 1) // word_bits, dtype=dt)
    load().or_bitmask_gen(_p(ptr, _i32p), n, word_bits, out.ctypes.data_as(ctypes.c_void_p))
    return out


def bitmask_and(a, b) -> np.ndarray:
    a = _c(a, np.uint64)
    b = _c(b, np.uint64)
    out = np.zeros_like(a)
    load().or_bitmask_and(_p(a, _u64p), _p(b, _u64p), a.shape[0], _p(out, _u64p))
    return out


def compress(nseg: int, ptr, idx, val, mask):
    ptr, idx, val, mask = _c(ptr, np.int32), _c(idx, np.int32), _c(val, np.float64), _c(mask, np.uint64)
    lib = load()
    optr = np.zeros(nseg + 1, np.int32)
    n = lib.or_compress(nseg, _p(ptr, _i32p), _p(idx, _i32p), _p(val, _f64p), _p(mask, _u64p),
                        _p(optr, _i32p), None, None)
    oidx = np.zeros(n, np.int32)
    oval = np.zeros(n, np.float64)
    lib.or_compress(nseg, _p(ptr, _i32p), _p(idx, _i32p), _p(val, _f64p), _p(mask, _u64p),
                    _p(optr, _i32p), _p(oidx, _i32p), _p(oval, _f64p))
    return optr, oidx, oval


def spgemm_flops(A, B) -> int:
    ap, ai, bp = _c(A.ptr, np.int32), _c(A.idx, np.int32), _c(B.ptr, np.int32)
    return int(load().or_spgemm_flops(A.nrows, _p(ap, _i32p), _p(ai, _i32p), _p(bp, _i32p)))


def spgemm_rownnz(A, B) -> np.ndarray:
    ap, ai = _c(A.ptr, np.int32), _c(A.idx, np.int32)
    bp, bi = _c(B.ptr, np.int32), _c(B.idx, np.int32)
    rn = np.zeros(A.nrows, np.int64)
    load().or_spgemm_rownnz(A.nrows, B.ncols, _p(ap, _i32p), _p(ai, _i32p), _p(bp, _i32p), _p(bi, _i32p),
                            _p(rn, _i64p))
    return rn


def spgemm(A, B):
    """C = A*B in the canonical order; returns (ptr int64, idx int32, val f64)."""
    assert A.ncols == B.nrows
    ap, ai, av = _c(A.ptr, np.int32), _c(A.idx, np.int32), _c(A.val, np.float64)
    bp, bi, bv = _c(B.ptr, np.int32), _c(B.idx, np.int32), _c(B.val, np.float64)
    rn = spgemm_rownnz(A, B)
    cp = np.zeros(A.nrows + 1, np.int64)
    np.cumsum(rn, out=cp[1:])
    ci = np.zeros(int(cp[-1]), np.int32)
    cv = np.zeros(int(cp[-1]), np.float64)
    load().or_spgemm_numeric(A.nrows, B.ncols, _p(ap, _i32p), _p(ai, _i32p), _p(av, _f64p), _p(bp, _i32p),
                             _p(bi, _i32p), _p(bv, _f64p), _p(cp, _i64p), _p(ci, _i32p), _p(cv, _f64p))
    return cp, ci, cv


def spadd(A, B):
    """C = A + B for sorted operands; returns (ptr int64, idx, val)."""
    assert A.nrows == B.nrows and A.ncols == B.ncols
    ap, ai, av = _c(A.ptr, np.int32), _c(A.idx, np.int32), _c(A.val, np.float64)
    bp, bi, bv = _c(B.ptr, np.int32), _c(B.idx, np.int32), _c(B.val, np.float64)
    rn = np.zeros(A.nrows, np.int64)
    lib = load()
    lib.or_spadd_rownnz(A.nrows, _p(ap, _i32p), _p(ai, _i32p), _p(bp, _i32p), _p(bi, _i32p), _p(rn, _i64p))
    cp = np.zeros(A.nrows + 1, np.int64)
    np.cumsum(rn, out=cp[1:])
    ci = np.zeros(int(cp[-1]), np.int32)
    cv = np.zeros(int(cp[-1]), np.float64)
    lib.or_spadd_numeric(A.nrows, _p(ap, _i32p), _p(ai, _i32p), _p(av, _f64p), _p(bp, _i32p), _p(bi, _i32p),
                         _p(bv, _f64p), _p(cp, _i64p), _p(ci, _i32p), _p(cv, _f64p))
    return cp, ci, cv


def transpose(A):
    ap, ai, av = _c(A.ptr, np.int32), _c(A.idx, np.int32), _c(A.val, np.float64)
    tp = np.zeros(A.ncols + 1, np.int32)
    ti = np.zeros(A.nnz, np.int32)
    tv = np.zeros(A.nnz, np.float64)
    load().or_transpose(A.nrows, A.ncols, _p(ap, _i32p), _p(ai, _i32p), _p(av, _f64p), _p(tp, _i32p),
                        _p(ti, _i32p), _p(tv, _f64p))
    return tp, ti, tv


def rows_sorted(A) -> bool:
    ap, ai = _c(A.ptr, np.int32), _c(A.idx, np.int32)
    return bool(load().or_rows_sorted(A.nrows, _p(ap, _i32p), _p(ai, _i32p)))


def num_threads() -> int:
    return int(load().or_num_threads())
