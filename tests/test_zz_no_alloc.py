"""tr_spadd_into_noalloc (the add with its temporaries in a fixed per-GPU
device array) against tr_spadd_into and the oracle, through the C-ABI:
bit-identical results on two calls in a row (the second with a smaller
input, so stale scratch from the first call must not leak), no pool
allocation during the calls, the capacity overflow report, and a clear
rejection of operands whose temporaries exceed the fixed array."""
import ctypes

import numpy as np
import pytest

import oracle
from trident_amd import _lib, gen
from trident_amd.device import Context

pytestmark = pytest.mark.gpu


def _perturbed(A, seed):
    """Same shape as A, a different pattern: every other entry of A plus a
    random ER matrix, merged and sorted."""
    rng = np.random.default_rng(seed)
    keep = rng.random(A.nnz) < 0.5
    rows = A.row_ids()[keep]
    cols = A.idx[keep]
    E = gen.erdos_renyi(A.nrows, 3, seed=seed) if A.nrows == A.ncols else None
    if E is not None:
        rows = np.concatenate([rows, E.row_ids()])
        cols = np.concatenate([cols, E.idx])
    key = np.unique(rows.astype(np.int64) * A.ncols + cols)
    r = (key // A.ncols).astype(np.int32)
    c = (key % A.ncols).astype(np.int32)
    ptr = np.zeros(A.nrows + 1, np.int64)
    np.add.at(ptr, r + 1, 1)
    v = rng.uniform(-1, 1, key.size)
    return gen.from_arrays(A.nrows, A.ncols, np.cumsum(ptr), c, v)


class _Out:
    """Caller-owned output arrays of an add."""

    def __init__(self, ctx, rows, cap):
        self.ctx = ctx
        self.p = ctx.malloc(4 * (rows + 1))
        self.i = ctx.malloc(4 * max(cap, 1))
        self.v = ctx.malloc(8 * max(cap, 1))
        self.d = _lib.tr_csx_t()
        self.d.ptr_vec, self.d.idx_vec, self.d.val = self.p, self.i, self.v
        self.rows = rows

    def download(self):
        n = int(self.d.nnz)
        return (self.ctx.to_host(self.p, self.rows + 1, np.int32), self.ctx.to_host(self.i, n, np.int32),
                self.ctx.to_host(self.v, n, np.float64))

    def free(self):
        for q in (self.p, self.i, self.v):
            self.ctx.free(q)


def _add(ctx, fn, dA, dB, out, cap):
    need = ctypes.c_int64(-1)
    rc = fn(ctx.h, ctypes.byref(dA.d), ctypes.byref(dB.d), ctypes.byref(out.d), cap, ctypes.byref(need))
    return rc, int(need.value)


def test_noalloc_add_matches_spadd_into_twice():
    lib = _lib.load()
    with Context(0) as ctx:
        for step, (scale, seed) in enumerate([(13, 11), (9, 12)]):  # second call smaller
            A = gen.rmat(scale, 8, seed=seed)
            B = _perturbed(A, seed + 100)
            print(f"noalloc add call {step}: R-MAT s{scale} nnz(A)={A.nnz} nnz(B)={B.nnz}", flush=True)
            dA, dB = ctx.upload(A), ctx.upload(B)
            ep, ei, ev = oracle.spadd(A, B)
            cap = int(ei.size) + 17
            ref, got = _Out(ctx, A.nrows, cap), _Out(ctx, A.nrows, cap)
            rc, need = _add(ctx, lib.tr_spadd_into, dA, dB, ref, cap)
            ctx.check(rc, "tr_spadd_into")
            ctx.synchronize()
            used0 = ctx.pool_stats()
            rc, need2 = _add(ctx, lib.tr_spadd_into_noalloc, dA, dB, got, cap)
            ctx.check(rc, "tr_spadd_into_noalloc")
            ctx.synchronize()
            assert ctx.pool_stats() == used0, "the no-alloc add touched the context pool"
            assert need2 == need == ei.size
            rp, ri, rv = ref.download()
            gp, gi, gv = got.download()
            assert np.array_equal(gp, rp) and np.array_equal(gi, ri)
            assert np.array_equal(gv.view(np.uint64), rv.view(np.uint64))
            assert np.array_equal(gp.astype(np.int64), ep) and np.array_equal(gi, ei)
            assert np.array_equal(gv.view(np.uint64), ev.view(np.uint64))
            # too small a capacity: reported, nothing written
            small = _Out(ctx, A.nrows, 4)
            rc, need3 = _add(ctx, lib.tr_spadd_into_noalloc, dA, dB, small, 4)
            assert rc == _lib.TR_ERR_OVERFLOW and need3 == ei.size
            for o in (ref, got, small):
                o.free()
            dA.free()
            dB.free()


def test_noalloc_add_rejects_oversized_operands():
    lib = _lib.load()
    with Context(0) as ctx:
        n = 3 << 20  # 28 B of temporaries per row: more than the 64 MiB array
        print(f"noalloc add rejection: {n} empty rows", flush=True)
        Z = gen.from_arrays(n, 8, np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0))
        dA, dB = ctx.upload(Z), ctx.upload(Z)
        out = _Out(ctx, n, 1)
        rc, _ = _add(ctx, lib.tr_spadd_into_noalloc, dA, dB, out, 1)
        assert rc == _lib.TR_ERR_INVALID
        assert b"scratch" in lib.tr_last_error(ctx.h)
        # the context is still usable: the regular add takes the same operands
        rc, need = _add(ctx, lib.tr_spadd_into, dA, dB, out, 1)
        ctx.check(rc, "tr_spadd_into")
        assert need == 0
