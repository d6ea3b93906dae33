"""GPU parity of bitmask generation / intersection / tile compression, the
sorted sparse add, sp_mma's accumulation semantics and the transpose,
against the CPU oracle (bit-exact: integer work and exact fp64 copies/sums)."""
import ctypes

import numpy as np
import pytest

import oracle
from trident_amd import gen
from trident_amd._lib import TR_COLS
from trident_amd.device import DeviceCsx, SpaComm, sp_mma, spadd, spgemm

pytestmark = pytest.mark.gpu


def dev_mask(ctx, ptr, n):
    dptr = ctx.to_device(np.asarray(ptr, np.int32))
    m = SpaComm.gen_bitmask(ctx, dptr, n)
    host = ctx.to_host(m, ((n + 63) // 64,), np.uint64)
    ctx.free(dptr)
    return m, host


def test_bitmask_known_answers(ctx):
    # test/test_compression.cu:550-589 (MASK_SIZE=8 words equal the low byte)
    _, t = dev_mask(ctx, [0, 0, 1, 1, 3, 3, 6, 6], 7)
    ma, a = dev_mask(ctx, [0, 1, 2, 2, 3, 4, 4, 5, 6], 8)
    mb, b = dev_mask(ctx, [0, 0, 1, 2, 2, 3, 4, 4, 5], 8)
    assert t.tolist() == [42] and a.tolist() == [219] and b.tolist() == [182]
    mi = SpaComm.intersect(ctx, ma, mb, 1)
    assert ctx.to_host(mi, (1,), np.uint64).tolist() == [146]


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 100003])
def test_bitmask_random(ctx, n):
    rng = np.random.default_rng(n)
    lens = rng.integers(0, 3, n) * (rng.random(n) < 0.4)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    _, got = dev_mask(ctx, ptr, n)
    assert got.tolist() == oracle.bitmask_gen(ptr, n, 64).tolist()


def _compress_check(ctx, M, mask, majordim=0):
    dM = ctx.upload(M, majordim)
    dmask = ctx.to_device(mask)
    out = SpaComm.compress(dM, dmask)
    assert out.majordim == majordim and out.shape == dM.shape
    H = out.download()
    nseg = M.nrows if majordim == 0 else M.ncols
    p, i, v = oracle.compress(nseg, M.ptr, M.idx, M.val, mask)
    assert np.array_equal(H.ptr, p) and np.array_equal(H.idx, i) and np.array_equal(H.val, v)
    # contiguous tile layout [val | idx | ptr] (include/tile_holder.cuh:80-89)
    d = out.d
    assert d.contig == 1 and d.val == d.buf
    assert d.idx_vec == d.buf + 8 * d.nnz and d.ptr_vec == d.buf + 12 * d.nnz
    out.free()
    dM.free()
    ctx.free(dmask)
    return H


def test_compress_tmp_test1(ctx):
    val = np.arange(1, 28, dtype=np.float64)
    col = np.array([0, 1, 3, 7, 0, 1, 4, 7, 8, 10, 0, 2, 4, 6, 8, 1, 3, 5, 7, 9, 2, 8, 1, 5, 7, 0, 10], np.int32)
    ptr = np.array([0, 1, 4, 10, 15, 20, 22, 25, 27], np.int32)
    M = gen.from_arrays(8, 11, ptr, col, val)
    H = _compress_check(ctx, M, np.array([3], np.uint64))
    assert H.ptr.tolist() == [0, 1, 4, 4, 4, 4, 4, 4, 4] and H.idx.tolist() == [0, 1, 3, 7]
    H = _compress_check(ctx, M, np.array([0b10010011], np.uint64))
    assert H.ptr.tolist() == [0, 1, 4, 4, 4, 9, 9, 9, 11]


@pytest.mark.parametrize("scale", [8, 12, 16])
def test_compress_random_csr_and_csc(ctx, scale):
    A = gen.rmat(scale, 16, seed=scale)
    n = A.nrows
    rng = np.random.default_rng(scale)
    words = (n + 63) // 64
    for mask in (rng.integers(0, 2**63, size=words, dtype=np.uint64),
                 np.full(words, np.uint64(2**64 - 1)), np.zeros(words, np.uint64)):
        _compress_check(ctx, A, mask, 0)
        At = gen.transpose(A)  # CSC arrays of A
        Mcsc = gen.from_arrays(A.nrows, A.ncols, At.ptr, At.idx, At.val)
        _compress_check(ctx, Mcsc, mask, TR_COLS)


def test_compress_long_segment(ctx):
    # one segment far longer than a wave's chunk
    rng = np.random.default_rng(2)
    lens = np.array([5, 200000, 0, 3] + [7] * 60 + [100000], np.int64)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = rng.integers(0, 1000, ptr[-1]).astype(np.int32)
    M = gen.from_arrays(len(lens), 1000, ptr, idx, rng.uniform(-1, 1, ptr[-1]))
    _compress_check(ctx, M, np.array([0xAAAAAAAAAAAAAAAB, 1], np.uint64))


def test_compress_preserves_product(ctx):
    A = gen.rmat(11, 8, seed=11)
    B = gen.rmat(11, 8, seed=12)
    At = gen.transpose(A)
    k = A.ncols
    ma = oracle.bitmask_gen(At.ptr, k, 64)
    mb = oracle.bitmask_gen(B.ptr, k, 64)
    mask = oracle.bitmask_and(ma, mb)
    dmask = ctx.to_device(mask)
    dAcsc = ctx.upload(gen.from_arrays(A.nrows, A.ncols, At.ptr, At.idx, At.val), TR_COLS)
    dB = ctx.upload(B)
    cA = SpaComm.compress(dAcsc, dmask)
    cB = SpaComm.compress(dB, dmask)
    cA_csr = cA.csc_to_csr()
    C = spgemm(cA_csr, cB).download()
    cp, ci, cv = oracle.spgemm(A, B)
    assert np.array_equal(C.ptr.astype(np.int64), cp) and np.array_equal(C.idx, ci) and np.array_equal(C.val, cv)


@pytest.mark.parametrize("scale", [6, 10, 14])
def test_spadd_random(ctx, scale):
    A = gen.rmat(scale, 8, seed=1)
    B = gen.rmat(scale, 8, seed=2)
    dA = ctx.upload(A)
    dC = ctx.upload(B)
    spadd(dA, dC)
    H = dC.download()
    cp, ci, cv = oracle.spadd(A, B)
    assert np.array_equal(H.ptr.astype(np.int64), cp) and np.array_equal(H.idx, ci) and np.array_equal(H.val, cv)


def test_spadd_long_rows_and_semantics(ctx):
    rng = np.random.default_rng(0)
    n = 3000
    rowsA = [np.sort(rng.choice(100000, size=s, replace=False)) for s in (50000, 0, 3, 70000)]
    rowsB = [np.sort(rng.choice(100000, size=s, replace=False)) for s in (30000, 5, 0, 70000)]
    def mk(rows):
        ptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
        return gen.from_arrays(len(rows), 100000, ptr, np.concatenate(rows), rng.uniform(-1, 1, ptr[-1]))
    A, B = mk(rowsA), mk(rowsB)
    dA = ctx.upload(A)
    dC = ctx.upload(B)
    spadd(dA, dC)
    cp, ci, cv = oracle.spadd(A, B)
    H = dC.download()
    assert np.array_equal(H.ptr.astype(np.int64), cp) and np.array_equal(H.idx, ci) and np.array_equal(H.val, cv)
    # A.nnz == 0 leaves C untouched (KokkosWrap.hpp:400-402)
    E = gen.from_arrays(4, 100000, np.zeros(5), [], [])
    before = dC.download()
    spadd(ctx.upload(E), dC)
    after = dC.download()
    assert np.array_equal(before.idx, after.idx) and np.array_equal(before.val, after.val)
    # uninitialised C becomes a copy of A (KokkosWrap.hpp:403-407)
    U = DeviceCsx(ctx)
    spadd(dA, U)
    assert U.initialized and np.array_equal(U.download().val, A.val)
    _ = n


def test_sp_mma_accumulates_stages(ctx):
    # C = sum_s A_s * B_s over four stages, as the stage loop of
    # hns_spgemm_async drives sp_mma_hybrid (src/hns_spgemm.cu:953-969)
    rng = np.random.default_rng(9)
    As = [gen.rmat(10, 8, seed=100 + s) for s in range(4)]
    Bs = [gen.rmat(10, 8, seed=200 + s) for s in range(4)]
    C = DeviceCsx(ctx)
    ref = None
    for A, B in zip(As, Bs):
        sp_mma(ctx.upload(A), ctx.upload(B), C)
        prod = oracle.spgemm(A, B)
        P = gen.from_arrays(A.nrows, B.ncols, prod[0], prod[1], prod[2])
        if ref is None:
            ref = P
        else:
            cp, ci, cv = oracle.spadd(P, ref)
            ref = gen.from_arrays(A.nrows, B.ncols, cp, ci, cv)
    H = C.download()
    assert np.array_equal(H.ptr, ref.ptr) and np.array_equal(H.idx, ref.idx) and np.array_equal(H.val, ref.val)
    # an empty product leaves C as it was (cusparse_helpers.cuh:719-722)
    Z = gen.from_arrays(1024, 1024, np.zeros(1025), [], [])
    sp_mma(ctx.upload(Z), ctx.upload(Bs[0]), C)
    assert np.array_equal(C.download().val, ref.val)
    _ = rng


def test_transpose_and_csc_to_csr(ctx):
    A = gen.rmat(12, 16, seed=3)
    dA = ctx.upload(A)
    T = dA.transpose().download()
    tp, ti, tv = oracle.transpose(A)
    assert np.array_equal(T.ptr, tp) and np.array_equal(T.idx, ti) and np.array_equal(T.val, tv)
    # CSC arrays of A (= CSR of A^T) back to CSR of A
    Mcsc = ctx.upload(gen.from_arrays(A.nrows, A.ncols, tp, ti, tv), TR_COLS)
    R = Mcsc.csc_to_csr().download()
    assert np.array_equal(R.ptr, A.ptr) and np.array_equal(R.idx, A.idx) and np.array_equal(R.val, A.val)


def test_sort_rows_long(ctx):
    rng = np.random.default_rng(4)
    lens = [1, 0, 70, 3000, 20000, 40000]
    rows = [rng.permutation(50000)[:s] for s in lens]
    ptr = np.concatenate([[0], np.cumsum(lens)])
    M = gen.from_arrays(len(lens), 50000, ptr, np.concatenate(rows), rng.uniform(-1, 1, ptr[-1]))
    d = ctx.upload(M)
    assert d.validate() & 4
    d.sort_rows()
    assert d.validate() == 0
    H = d.download()
    for r in range(len(lens)):
        s, e = ptr[r], ptr[r + 1]
        order = np.argsort(M.idx[s:e], kind="stable")
        assert np.array_equal(H.idx[s:e], M.idx[s:e][order])
        assert np.array_equal(H.val[s:e], M.val[s:e][order])
    _ = ctypes
