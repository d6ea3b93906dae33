"""Seeded synthetic operands (ctypes front end of libtrident_gen.so).

Benchmark and test infrastructure: the C++ generators in csrc/gen.cpp build
int32 CSR matrices with sorted unique columns and fp64 values; this module
copies them into numpy arrays.  See include/trident_gen.h for the contract.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class _HostCsr(ctypes.Structure):
    _fields_ = [
        ("nrows", ctypes.c_int32),
        ("ncols", ctypes.c_int32),
        ("nnz", ctypes.c_int64),
        ("ptr", ctypes.POINTER(ctypes.c_int32)),
        ("idx", ctypes.POINTER(ctypes.c_int32)),
        ("val", ctypes.POINTER(ctypes.c_double)),
    ]


@dataclass
class Csr:
    """Host CSR matrix: int32 ptr/idx, float64 val."""

    nrows: int
    ncols: int
    ptr: np.ndarray
    idx: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.idx.shape[0])

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.nrows, self.ncols), dtype=np.float64)
        rows = np.repeat(np.arange(self.nrows), np.diff(self.ptr.astype(np.int64)))
        d[rows, self.idx] = self.val
        return d

    def row_ids(self) -> np.ndarray:
        return np.repeat(np.arange(self.nrows, dtype=np.int32), np.diff(self.ptr.astype(np.int64)))


_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libtrident_gen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        P = ctypes.POINTER(_HostCsr)
        lib.tr_gen_er.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, P]
        lib.tr_gen_rmat.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint64, P]
        lib.tr_gen_laplace27.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double, P]
        lib.tr_gen_sa_prolongator.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, P]
        lib.tr_host_transpose.argtypes = [P, P]
        lib.tr_host_csr_free.argtypes = [P]
        for f in ("tr_gen_er", "tr_gen_rmat", "tr_gen_laplace27", "tr_gen_sa_prolongator",
                  "tr_host_transpose"):
            getattr(lib, f).restype = ctypes.c_int
        lib.tr_host_csr_free.restype = None
        _lib = lib
    return _lib


def _take(h: _HostCsr) -> Csr:
    lib = _load()
    n = h.nrows
    nnz = h.nnz
    ptr = np.ctypeslib.as_array(h.ptr, shape=(n + 1,)).copy()
    if nnz:
        idx = np.ctypeslib.as_array(h.idx, shape=(nnz,)).copy()
        val = np.ctypeslib.as_array(h.val, shape=(nnz,)).copy()
    else:
        idx = np.zeros(0, np.int32)
        val = np.zeros(0, np.float64)
    out = Csr(h.nrows, h.ncols, ptr, idx, val)
    lib.tr_host_csr_free(ctypes.byref(h))
    return out


def _check(rc: int, what: str):
    if rc != 0:
        raise ValueError(f"{what} failed with code {rc}")


def erdos_renyi(n: int, per_row: int = 8, seed: int = 0) -> Csr:
    h = _HostCsr()
    _check(_load().tr_gen_er(n, per_row, seed, ctypes.byref(h)), "tr_gen_er")
    return _take(h)


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1) -> Csr:
    h = _HostCsr()
    _check(_load().tr_gen_rmat(scale, edge_factor, a, b, c, seed, ctypes.byref(h)), "tr_gen_rmat")
    return _take(h)


def laplace27(n: int, diag: float = 26.0, off: float = -1.0) -> Csr:
    h = _HostCsr()
    _check(_load().tr_gen_laplace27(n, diag, off, ctypes.byref(h)), "tr_gen_laplace27")
    return _take(h)


def sa_prolongator(n: int, agg: int = 3, diag: float = 26.0, off: float = -1.0,
                   omega: float = 2.0 / 3.0) -> Csr:
    h = _HostCsr()
    _check(_load().tr_gen_sa_prolongator(n, agg, diag, off, omega, ctypes.byref(h)),
           "tr_gen_sa_prolongator")
    return _take(h)


def from_arrays(nrows: int, ncols: int, ptr, idx, val) -> Csr:
    return Csr(int(nrows), int(ncols), np.ascontiguousarray(ptr, np.int32),
               np.ascontiguousarray(idx, np.int32), np.ascontiguousarray(val, np.float64))


def transpose(m: Csr) -> Csr:
    src = _HostCsr()
    src.nrows, src.ncols, src.nnz = m.nrows, m.ncols, m.nnz
    ptr = np.ascontiguousarray(m.ptr, np.int32)
    idx = np.ascontiguousarray(m.idx, np.int32)
    val = np.ascontiguousarray(m.val, np.float64)
    src.ptr = ptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    src.idx = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    src.val = val.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    h = _HostCsr()
    _check(_load().tr_host_transpose(ctypes.byref(src), ctypes.byref(h)), "tr_host_transpose")
    return _take(h)


def submatrix(m: Csr, r0: int, r1: int, c0: int, c1: int) -> Csr:
    """Tile m[r0:r1, c0:c1] with re-based column indices (host)."""
    ptr = m.ptr.astype(np.int64)
    s, e = ptr[r0], ptr[r1]
    rows = np.repeat(np.arange(r1 - r0), np.diff(ptr[r0:r1 + 1]))
    cols = m.idx[s:e]
    vals = m.val[s:e]
    keep = (cols >= c0) & (cols < c1)
    rows, cols, vals = rows[keep], cols[keep] - c0, vals[keep]
    cnt = np.bincount(rows, minlength=r1 - r0)
    nptr = np.zeros(r1 - r0 + 1, np.int64)
    np.cumsum(cnt, out=nptr[1:])
    return Csr(r1 - r0, c1 - c0, nptr.astype(np.int32), cols.astype(np.int32), vals.astype(np.float64))
