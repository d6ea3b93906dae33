"""Python front end of the local SpGEMM path (over the C-ABI).

`Context` owns a HIP stream and the library's device pool; `DeviceCsx`
wraps a `tr_csx_t`; `LocalMatrix` mirrors the reference's
KokkosWrap::LocalMatrix (include/KokkosWrap.hpp:43-77) with the same static
operations (`spgemm`, `sp_mma`, `spadd`) and `SpaComm` mirrors the
compression helpers of include/spacomm.cuh.  Every operation runs in the HIP
library; errors raise `TridentError` with the library's message.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Tuple

import numpy as np

from . import _lib
from ._lib import TR_COLS, TR_ROWS, TridentError, tr_csx_t
from .gen import Csr


def _np_ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


TUNING_FIELDS = ("light_max_flops", "window_entries", "item_windows", "heavy_table", "dense_width")


def _tuning_struct(tuning: dict) -> "_lib.tr_tuning_t":
    lib = _lib.load()
    t = _lib.tr_tuning_t()
    lib.tr_tuning_default(ctypes.byref(t))
    for k, v in tuning.items():
        if k not in TUNING_FIELDS:
            raise ValueError(f"unknown tuning field {k!r}")
        setattr(t, k, int(v))
    return t


def tuning_default() -> dict:
    """The library's built-in multiply tuning."""
    t = _lib.tr_tuning_t()
    _lib.load().tr_tuning_default(ctypes.byref(t))
    return {f: int(getattr(t, f)) for f in TUNING_FIELDS}


def set_default_tuning(tuning: Optional[dict]):
    """Tuning that contexts created afterwards start with (None: built-in)."""
    lib = _lib.load()
    t = _tuning_struct(tuning) if tuning is not None else None
    rc = lib.tr_set_default_tuning(ctypes.byref(t) if t is not None else None)
    if rc != _lib.TR_OK:
        raise TridentError(rc, f"invalid tuning {tuning}")


def load_tuning(path: str) -> dict:
    """The tuning chosen by trident_amd.autotune, from its JSON output."""
    import json
    with open(path) as f:
        doc = json.load(f)
    t = doc.get("tuning", doc)
    return {k: int(t[k]) for k in TUNING_FIELDS if k in t}


class Context:
    """A device context (one HIP stream, one caching pool)."""

    def __init__(self, device: int = 0, stream: Optional[int] = None):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.tr_ctx_create(ctypes.byref(h), device, stream)
        if rc != _lib.TR_OK:
            raise TridentError(rc, "tr_ctx_create failed (is a GPU visible?)")
        self.h = h

    # error plumbing
    def check(self, rc: int, what: str = ""):
        if rc != _lib.TR_OK:
            msg = self.lib.tr_last_error(self.h).decode()
            raise TridentError(rc, f"{what}: {msg}" if what else msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.tr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        self.check(self.lib.tr_ctx_synchronize(self.h), "synchronize")

    def set_tuning(self, tuning: Optional[dict]):
        """Multiply cutovers of this context (None: the process default)."""
        t = _tuning_struct(tuning) if tuning is not None else None
        self.check(self.lib.tr_ctx_set_tuning(self.h, ctypes.byref(t) if t is not None else None), "tuning")

    def get_tuning(self) -> dict:
        t = _lib.tr_tuning_t()
        self.check(self.lib.tr_ctx_get_tuning(self.h, ctypes.byref(t)), "tuning")
        return {f: int(getattr(t, f)) for f in TUNING_FIELDS}

    @property
    def stream(self) -> int:
        return int(self.lib.tr_ctx_stream(self.h) or 0)

    # raw memory
    def malloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        self.check(self.lib.tr_malloc(self.h, ctypes.byref(p), nbytes), "tr_malloc")
        return int(p.value or 0)

    def free(self, ptr: int):
        self.check(self.lib.tr_free(self.h, ctypes.c_void_p(ptr)), "tr_free")

    def to_device(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = self.malloc(max(arr.nbytes, 1))
        if arr.nbytes:
            self.check(self.lib.tr_memcpy_h2d(self.h, ctypes.c_void_p(p), _np_ptr(arr), arr.nbytes), "h2d")
        return p

    def to_host(self, ptr: int, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            self.check(self.lib.tr_memcpy_d2h(self.h, _np_ptr(out), ctypes.c_void_p(ptr), out.nbytes), "d2h")
        return out

    def pool_stats(self) -> Tuple[int, int]:
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        self.lib.tr_pool_stats(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def pool_trim(self):
        self.check(self.lib.tr_pool_trim(self.h), "tr_pool_trim")

    def set_pool_limit(self, nbytes: int):
        self.lib.tr_pool_set_limit(self.h, nbytes)

    # instrumentation
    def set_timing(self, on: bool):
        self.check(self.lib.tr_ctx_set_timing(self.h, 1 if on else 0), "timing")

    def reset_timing(self):
        self.check(self.lib.tr_ctx_reset_timing(self.h), "timing")

    def timings(self) -> dict:
        n = self.lib.tr_ctx_timing_count(self.h)
        out = {}
        buf = ctypes.create_string_buffer(128)
        for i in range(n):
            launches = ctypes.c_int64()
            ms = ctypes.c_double()
            self.check(self.lib.tr_ctx_timing_get(self.h, i, buf, 128, ctypes.byref(launches), ctypes.byref(ms)))
            out[buf.value.decode()] = (int(launches.value), float(ms.value))
        return out

    # matrices
    def upload(self, m: Csr, majordim: int = TR_ROWS) -> "DeviceCsx":
        ptr = np.ascontiguousarray(m.ptr, np.int32)
        idx = np.ascontiguousarray(m.idx, np.int32)
        val = np.ascontiguousarray(m.val, np.float64)
        d = tr_csx_t()
        self.check(self.lib.tr_csx_upload(self.h, majordim, m.nrows, m.ncols, m.nnz, _np_ptr(ptr), _np_ptr(idx),
                                          _np_ptr(val), ctypes.byref(d)), "tr_csx_upload")
        return DeviceCsx(self, d)


class DeviceCsx:
    """A device matrix: a `tr_csx_t` plus the context it belongs to."""

    def __init__(self, ctx: Context, d: Optional[tr_csx_t] = None):
        self.ctx = ctx
        self.d = d if d is not None else tr_csx_t()

    @property
    def initialized(self) -> bool:
        return bool(self.d.ptr_vec)

    @property
    def nnz(self) -> int:
        return int(self.d.nnz)

    @property
    def shape(self) -> Tuple[int, int]:
        return int(self.d.nrows), int(self.d.ncols)

    @property
    def majordim(self) -> int:
        return int(self.d.majordim)

    def ptrdim(self) -> int:
        return self.d.nrows if self.d.majordim == TR_ROWS else self.d.ncols

    def download(self) -> Csr:
        """Copy to host; for a CSC matrix the arrays are the CSC arrays."""
        P = self.ptrdim()
        ptr = np.zeros(P + 1, np.int32)
        idx = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz, np.float64)
        self.ctx.check(self.ctx.lib.tr_csx_download(self.ctx.h, ctypes.byref(self.d), _np_ptr(ptr), _np_ptr(idx),
                                                    _np_ptr(val)), "tr_csx_download")
        return Csr(int(self.d.nrows), int(self.d.ncols), ptr, idx, val)

    def free(self):
        if self.initialized:
            self.ctx.check(self.ctx.lib.tr_csx_free(self.ctx.h, ctypes.byref(self.d)), "tr_csx_free")

    def validate(self) -> int:
        st = ctypes.c_int32()
        self.ctx.check(self.ctx.lib.tr_csx_validate(self.ctx.h, ctypes.byref(self.d), ctypes.byref(st)),
                       "tr_csx_validate")
        return int(st.value)

    def sort_rows(self):
        self.ctx.check(self.ctx.lib.tr_csx_sort_rows(self.ctx.h, ctypes.byref(self.d)), "tr_csx_sort_rows")

    def transpose(self) -> "DeviceCsx":
        out = tr_csx_t()
        self.ctx.check(self.ctx.lib.tr_csx_transpose(self.ctx.h, ctypes.byref(self.d), ctypes.byref(out)),
                       "tr_csx_transpose")
        return DeviceCsx(self.ctx, out)

    def csc_to_csr(self) -> "DeviceCsx":
        out = tr_csx_t()
        self.ctx.check(self.ctx.lib.tr_csc_to_csr(self.ctx.h, ctypes.byref(self.d), ctypes.byref(out)),
                       "tr_csc_to_csr")
        return DeviceCsx(self.ctx, out)


# ---------------------------------------------------------------- operations
def spgemm(A: DeviceCsx, B: DeviceCsx) -> DeviceCsx:
    ctx = A.ctx
    out = tr_csx_t()
    ctx.check(ctx.lib.tr_spgemm(ctx.h, ctypes.byref(A.d), ctypes.byref(B.d), ctypes.byref(out)), "tr_spgemm")
    return DeviceCsx(ctx, out)


def spgemm_panels(A: DeviceCsx, B: DeviceCsx, max_panel_nnz: int = 2**31 - 1,
                  max_panels: int = 256) -> Tuple[List[DeviceCsx], List[int]]:
    ctx = A.ctx
    arr = (tr_csx_t * max_panels)()
    rb = (ctypes.c_int32 * (max_panels + 1))()
    npan = ctypes.c_int32()
    ctx.check(ctx.lib.tr_spgemm_panels(ctx.h, ctypes.byref(A.d), ctypes.byref(B.d), max_panel_nnz, arr, rb,
                                       max_panels, ctypes.byref(npan)), "tr_spgemm_panels")
    n = npan.value
    panels = []
    for p in range(n):
        d = tr_csx_t()
        ctypes.memmove(ctypes.byref(d), ctypes.byref(arr[p]), ctypes.sizeof(tr_csx_t))
        panels.append(DeviceCsx(ctx, d))
    return panels, [int(rb[p]) for p in range(n + 1)]


def sp_mma(A: DeviceCsx, B: DeviceCsx, C: DeviceCsx):
    ctx = A.ctx
    ctx.check(ctx.lib.tr_sp_mma(ctx.h, ctypes.byref(A.d), ctypes.byref(B.d), ctypes.byref(C.d)), "tr_sp_mma")


def spadd(A: DeviceCsx, C: DeviceCsx):
    ctx = A.ctx
    ctx.check(ctx.lib.tr_spadd(ctx.h, ctypes.byref(A.d), ctypes.byref(C.d)), "tr_spadd")


def symbolic(A: DeviceCsx, B: DeviceCsx) -> Tuple[int, int]:
    ctx = A.ctx
    nnz = ctypes.c_int64()
    nbytes = ctypes.c_size_t()
    ctx.check(ctx.lib.tr_symbolic(ctx.h, ctypes.byref(A.d), ctypes.byref(B.d), ctypes.byref(nnz),
                                  ctypes.byref(nbytes)), "tr_symbolic")
    return int(nnz.value), int(nbytes.value)


def spgemm_flops(A: DeviceCsx, B: DeviceCsx) -> int:
    ctx = A.ctx
    f = ctypes.c_int64()
    ctx.check(ctx.lib.tr_spgemm_flops(ctx.h, ctypes.byref(A.d), ctypes.byref(B.d), ctypes.byref(f)),
              "tr_spgemm_flops")
    return int(f.value)


class LocalMatrix:
    """Mirror of KokkosWrap::LocalMatrix<int32_t,int32_t,double>
    (include/KokkosWrap.hpp:43-77) over the HIP library."""

    def __init__(self, m: Optional[DeviceCsx] = None, ctx: Optional[Context] = None):
        self.m = m if m is not None else DeviceCsx(ctx)

    @property
    def initialized(self) -> bool:
        return self.m.initialized

    @staticmethod
    def spgemm(A: "LocalMatrix", B: "LocalMatrix") -> "LocalMatrix":
        return LocalMatrix(spgemm(A.m, B.m))

    @staticmethod
    def sp_mma(A: "LocalMatrix", B: "LocalMatrix", C: "LocalMatrix"):
        sp_mma(A.m, B.m, C.m)

    @staticmethod
    def spadd(A: "LocalMatrix", C: "LocalMatrix"):
        spadd(A.m, C.m)

    def get_csx(self) -> DeviceCsx:
        return self.m

    def freeBuffers(self):  # noqa: N802 - reference name
        self.m.free()


class SpaComm:
    """Bitmask helpers of include/spacomm.cuh on 64-bit words."""

    @staticmethod
    def gen_bitmask(ctx: Context, ptr_dev: int, n: int) -> int:
        """Returns a device pointer to ceil(n/64) uint64 words (pool-owned)."""
        words = (n + 63) // 64
        out = ctx.malloc(max(words, 1) * 8)
        ctx.check(ctx.lib.tr_bitmask_gen(ctx.h, ctypes.c_void_p(ptr_dev), n, ctypes.c_void_p(out)),
                  "tr_bitmask_gen")
        return out

    @staticmethod
    def intersect(ctx: Context, a: int, b: int, words: int) -> int:
        out = ctx.malloc(max(words, 1) * 8)
        ctx.check(ctx.lib.tr_bitmask_and(ctx.h, ctypes.c_void_p(a), ctypes.c_void_p(b), words, ctypes.c_void_p(out)),
                  "tr_bitmask_and")
        return out

    @staticmethod
    def compress(M: DeviceCsx, mask_dev: int) -> DeviceCsx:
        ctx = M.ctx
        out = tr_csx_t()
        ctx.check(ctx.lib.tr_compress(ctx.h, ctypes.byref(M.d), ctypes.c_void_p(mask_dev), ctypes.byref(out)),
                  "tr_compress")
        return DeviceCsx(ctx, out)


class Batch:
    """One stage of tile products as a single hipGraph launch
    (tr_batch_create / tr_batch_run): C_out_t = C_in_t + A_t * B_t."""

    def __init__(self, ctx: Context, A: List[DeviceCsx], B: List[DeviceCsx],
                 C_in: Optional[List[Optional[DeviceCsx]]] = None, compress: bool = False, graph: bool = True):
        n = len(A)
        assert len(B) == n and (C_in is None or len(C_in) == n)
        self.ctx = ctx
        self.n = n
        arrA = (tr_csx_t * n)(*[a.d for a in A])
        arrB = (tr_csx_t * n)(*[b.d for b in B])
        arrC = None
        if C_in is not None:
            arrC = (tr_csx_t * n)(*[(c.d if c is not None else tr_csx_t()) for c in C_in])
        h = ctypes.c_void_p()
        flags = (1 if compress else 0) | (0 if graph else 2)
        ctx.check(ctx.lib.tr_batch_create(ctx.h, n, arrA, arrB, arrC, flags, ctypes.byref(h)), "tr_batch_create")
        self.h = h
        self._keep = (A, B, C_in)  # operands must outlive the batch

    def run(self):
        self.ctx.check(self.ctx.lib.tr_batch_run(self.h), "tr_batch_run")

    def check(self):
        """Raises if the last run met operands whose pattern differs from the plan."""
        self.ctx.check(self.ctx.lib.tr_batch_check(self.h), "tr_batch_check")

    def result(self, t: int) -> DeviceCsx:
        d = tr_csx_t()
        self.ctx.check(self.ctx.lib.tr_batch_result(self.h, t, ctypes.byref(d)), "tr_batch_result")
        return DeviceCsx(self.ctx, d)

    def result_panels(self, t: int, max_panel_nnz: int = 2**31 - 1,
                      max_panels: int = 256) -> Tuple[List[DeviceCsx], List[int]]:
        """Task t's result as int32 row panels (views valid until close)."""
        arr = (tr_csx_t * max_panels)()
        rb = (ctypes.c_int32 * (max_panels + 1))()
        npan = ctypes.c_int32()
        self.ctx.check(self.ctx.lib.tr_batch_result_panels(self.h, t, max_panel_nnz, arr, rb, max_panels,
                                                           ctypes.byref(npan)), "tr_batch_result_panels")
        out = []
        for p in range(npan.value):
            d = tr_csx_t()
            ctypes.memmove(ctypes.byref(d), ctypes.byref(arr[p]), ctypes.sizeof(tr_csx_t))
            out.append(DeviceCsx(self.ctx, d))
        return out, [int(rb[p]) for p in range(npan.value + 1)]

    def stats(self) -> Tuple[int, int, int]:
        f, p, o = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self.ctx.lib.tr_batch_stats(self.h, ctypes.byref(f), ctypes.byref(p), ctypes.byref(o))
        return int(f.value), int(p.value), int(o.value)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.tr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


TR_REQ_SPGEMM = 0
TR_REQ_GALERKIN = 1
TR_REQ_DISCARD = 0x100


class Server:
    """Continuous-batching request loop (tr_server_*): SpGEMM and Galerkin
    requests on resident operands, admitted by their symbolic memory bound."""

    def __init__(self, mem_budget: int, workers: int = 2, max_batch: int = 16, device: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self.lib.tr_server_create(device, mem_budget, workers, max_batch, ctypes.byref(h))
        if rc != _lib.TR_OK:
            raise TridentError(rc, "tr_server_create failed")
        self.h = h
        self._ops = []

    def register(self, m: DeviceCsx) -> int:
        op = ctypes.c_int32()
        rc = self.lib.tr_server_register(self.h, ctypes.byref(m.d), ctypes.byref(op))
        if rc != _lib.TR_OK:
            raise TridentError(rc, "tr_server_register failed")
        self._ops.append(m)
        return int(op.value)

    def submit(self, kind: int, a: int, b: int, discard: bool = False) -> int:
        rid = ctypes.c_int64()
        rc = self.lib.tr_server_submit(self.h, kind | (TR_REQ_DISCARD if discard else 0), a, b, ctypes.byref(rid))
        if rc != _lib.TR_OK:
            raise TridentError(rc, "tr_server_submit failed")
        return int(rid.value)

    def wait(self, rid: int, want_result: bool = True):
        d = tr_csx_t()
        lat = ctypes.c_double()
        rc = self.lib.tr_server_wait(self.h, rid, ctypes.byref(d) if want_result else None, ctypes.byref(lat))
        if rc != _lib.TR_OK:
            raise TridentError(rc, f"request {rid} failed")
        return (d if want_result else None), float(lat.value)

    def release(self, rid: int):
        self.lib.tr_server_release(self.h, rid)

    def stats(self) -> dict:
        c, b = ctypes.c_int64(), ctypes.c_int64()
        p50, p99, gf = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.lib.tr_server_stats(self.h, ctypes.byref(c), ctypes.byref(b), ctypes.byref(p50), ctypes.byref(p99),
                                 ctypes.byref(gf))
        return {"completed": c.value, "batches": b.value, "p50_ms": p50.value, "p99_ms": p99.value,
                "GFLOP/s": gf.value}

    def reset_stats(self):
        self.lib.tr_server_reset_stats(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.tr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
