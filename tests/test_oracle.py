"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against something other than its own
formula:

* or_bitmask_gen / or_bitmask_and — the known answers of the reference's
  test/test_compression.cu:550-589 (42, 219, 182, AND 146 with MASK_SIZE=8),
  the mod-k patterns of its synthetic generator (:301-518, check_mod_pattern
  :452-468, expectedA_fn/expectedB_fn :489-518), and the byte identity
  between 8-bit and 64-bit mask words.
* or_compress — the hand-derived result of tmp_test1 (:106-238) under both
  MASK_SIZE=8 (production, include/common.h:65) and the MASK_SIZE=4 its
  comment assumes; invariants all-ones = identity, zero mask = empty,
  idempotence, and Compress(A) * Compress(B) == A * B under intersection
  masks.
* or_spgemm_* — dense numpy products on integer-valued inputs (exact in any
  order), the boolean-product structure, closed forms (identity,
  permutation, diagonal), structural-zero retention, the sequential fma chain
  evaluated with exact rationals (pins the canonical order and the use of
  fma), the reference fixtures n16_full.bmtx / n1024.bmtx, and scipy on
  Erdos-Renyi n=1024.
* or_spadd_* — dense sums on integer values and a dict merge.
* or_transpose — numpy dense transpose.
"""
import os
import random
from fractions import Fraction

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from trident_amd import gen
from trident_amd.formats import read_bmtx


def fma_exact(a: float, b: float, c: float) -> float:
    # int/int true division is correctly rounded in CPython
    f = Fraction(a) * Fraction(b) + Fraction(c)
    return f.numerator / f.denominator


def csr_from_dense(d):
    rows, cols = np.nonzero(d)
    ptr = np.zeros(d.shape[0] + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=d.shape[0]), out=ptr[1:])
    return gen.from_arrays(d.shape[0], d.shape[1], ptr, cols, d[rows, cols])


def random_int_csr(rng, m, n, density, lo=-3, hi=4):
    d = rng.integers(lo, hi, size=(m, n)).astype(np.float64)
    d[rng.random((m, n)) > density] = 0.0
    return csr_from_dense(d)


# ---------------------------------------------------------------- bitmasks
def test_bitmask_known_answers_mask8():
    # test/test_compression.cu:550-589
    t = oracle.bitmask_gen([0, 0, 1, 1, 3, 3, 6, 6], 7, 8)
    a = oracle.bitmask_gen([0, 1, 2, 2, 3, 4, 4, 5, 6], 8, 8)
    b = oracle.bitmask_gen([0, 0, 1, 2, 2, 3, 4, 4, 5], 8, 8)
    assert list(t) == [42]
    assert list(a) == [219]
    assert list(b) == [182]
    assert int(a[0]) & int(b[0]) == 146


def test_bitmask_known_answers_mask64_and():
    a = oracle.bitmask_gen([0, 1, 2, 2, 3, 4, 4, 5, 6], 8, 64)
    b = oracle.bitmask_gen([0, 0, 1, 2, 2, 3, 4, 4, 5], 8, 64)
    assert list(a) == [219] and list(b) == [182]
    assert list(oracle.bitmask_and(a, b)) == [146]


def test_bitmask_mask8_bytes_equal_mask64_words():
    rng = np.random.default_rng(5)
    for n in (1, 63, 64, 65, 640, 1000):
        lens = rng.integers(0, 3, n) * (rng.random(n) < 0.5)
        ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        m8 = oracle.bitmask_gen(ptr, n, 8)
        m64 = oracle.bitmask_gen(ptr, n, 64)
        pad = np.zeros(len(m64) * 8, np.uint8)
        pad[:len(m8)] = m8
        assert pad.view(np.uint64).tolist() == m64.tolist()


def _synthetic_ptr(cond, n):
    """Segment pointer of the reference's gen_syntetic_matrix for one rank
    (test_compression.cu:301-450): segment i is non-empty iff cond(i)."""
    lens = [((i + 1) if i < n // 2 else (n - i)) if cond(i) else 0 for i in range(n)]
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)


def _mod_pattern(n, k, m):
    # check_mod_pattern (:452-468): bit i set iff i % m == k
    return oracle.bitmask_gen(np.concatenate([[0], np.cumsum([1 if i % m == k else 0 for i in range(n)])]),
                              n, 64)


@pytest.mark.parametrize("n", [16, 64, 96, 256])
def test_bitmask_mod_patterns_and_crt(n):
    conds = {"even": lambda i: i % 2 == 0, "odd": lambda i: i % 2 == 1,
             "twoOthree": lambda i: i % 3 == 2, "oneOthree": lambda i: i % 3 == 1,
             "zeroOthree": lambda i: i % 3 == 0}
    masks = {k: oracle.bitmask_gen(_synthetic_ptr(c, n), n, 64) for k, c in conds.items()}
    assert masks["even"].tolist() == _mod_pattern(n, 0, 2).tolist()
    assert masks["odd"].tolist() == _mod_pattern(n, 1, 2).tolist()
    assert masks["twoOthree"].tolist() == _mod_pattern(n, 2, 3).tolist()
    # expectedA_fn / expectedB_fn (:489-518): intersections are mod-6 classes
    pairs = {("even", "twoOthree"): 2, ("even", "oneOthree"): 4, ("even", "zeroOthree"): 0,
             ("odd", "twoOthree"): 5, ("odd", "oneOthree"): 1, ("odd", "zeroOthree"): 3}
    for (a, b), k in pairs.items():
        assert oracle.bitmask_and(masks[a], masks[b]).tolist() == _mod_pattern(n, k, 6).tolist()


# ---------------------------------------------------------------- compression
TMP1_VAL = np.arange(1, 28, dtype=np.float64)
TMP1_COL = np.array([0, 1, 3, 7, 0, 1, 4, 7, 8, 10, 0, 2, 4, 6, 8, 1, 3, 5, 7, 9, 2, 8, 1, 5, 7, 0, 10], np.int32)
TMP1_PTR = np.array([0, 1, 4, 10, 15, 20, 22, 25, 27], np.int32)


def test_compress_tmp_test1_mask8():
    # mask words {3, 9} with MASK_SIZE=8: word 0 = rows 0,1; word 1 is past
    # the 8 rows, so rows 0 and 1 are kept
    optr, oidx, oval = oracle.compress(8, TMP1_PTR, TMP1_COL, TMP1_VAL, np.array([3], np.uint64))
    assert optr.tolist() == [0, 1, 4, 4, 4, 4, 4, 4, 4]
    assert oidx.tolist() == [0, 1, 3, 7]
    assert oval.tolist() == [1.0, 2.0, 3.0, 4.0]


def test_compress_tmp_test1_mask4():
    # the comment's MASK_SIZE=4 reading: word0 = 3 -> rows 0,1; word1 = 9 ->
    # rows 4,7; as one 64-bit mask: bits {0,1,4,7}
    mask = np.array([(1 << 0) | (1 << 1) | (1 << 4) | (1 << 7)], np.uint64)
    optr, oidx, oval = oracle.compress(8, TMP1_PTR, TMP1_COL, TMP1_VAL, mask)
    assert optr.tolist() == [0, 1, 4, 4, 4, 9, 9, 9, 11]
    assert oidx.tolist() == [0, 1, 3, 7, 1, 3, 5, 7, 9, 0, 10]
    assert oval.tolist() == [1, 2, 3, 4, 16, 17, 18, 19, 20, 26, 27]


def test_compress_invariants():
    rng = np.random.default_rng(1)
    A = gen.rmat(8, 8, seed=3)
    n = A.nrows
    ones = np.full((n + 63) // 64, np.uint64(0xFFFFFFFFFFFFFFFF))
    p, i, v = oracle.compress(n, A.ptr, A.idx, A.val, ones)
    assert p.tolist() == A.ptr.tolist() and i.tolist() == A.idx.tolist() and v.tolist() == A.val.tolist()
    p, i, v = oracle.compress(n, A.ptr, A.idx, A.val, np.zeros_like(ones))
    assert p.tolist() == [0] * (n + 1) and len(i) == 0
    mask = rng.integers(0, 2**63, size=len(ones), dtype=np.uint64)
    p1, i1, v1 = oracle.compress(n, A.ptr, A.idx, A.val, mask)
    p2, i2, v2 = oracle.compress(n, p1, i1, v1, mask)
    assert p1.tolist() == p2.tolist() and i1.tolist() == i2.tolist() and v1.tolist() == v2.tolist()
    # dropped rows are exactly the masked-off ones
    keep = np.array([(int(mask[r >> 6]) >> (r & 63)) & 1 for r in range(n)], bool)
    assert np.all((np.diff(p1) == np.diff(A.ptr)) | ~keep)
    assert np.all(np.diff(p1)[~keep] == 0)


def test_compress_preserves_product_under_intersection_masks():
    # A compressed by columns (as CSC) and B by rows with the AND of the
    # non-empty masks: removed entries contribute nothing to A*B
    A = gen.rmat(9, 4, seed=11)
    B = gen.rmat(9, 4, seed=12)
    At = gen.transpose(A)  # CSC of A = CSR of A^T
    k = A.ncols
    ma = oracle.bitmask_gen(At.ptr, k, 64)   # non-empty columns of A
    mb = oracle.bitmask_gen(B.ptr, k, 64)    # non-empty rows of B
    mask = oracle.bitmask_and(ma, mb)
    pa, ia, va = oracle.compress(k, At.ptr, At.idx, At.val, mask)
    pb, ib, vb = oracle.compress(k, B.ptr, B.idx, B.val, mask)
    Ac = gen.transpose(gen.from_arrays(k, A.nrows, pa, ia, va))
    Bc = gen.from_arrays(k, B.ncols, pb, ib, vb)
    c1 = oracle.spgemm(A, B)
    c2 = oracle.spgemm(Ac, Bc)
    for x, y in zip(c1, c2):
        assert np.array_equal(x, y)
    assert len(ia) < A.nnz or len(ib) < B.nnz  # something was dropped


# ---------------------------------------------------------------- spgemm
def _dense_of(m, cp, ci, cv, ncols):
    d = np.zeros((m, ncols))
    rows = np.repeat(np.arange(m), np.diff(cp))
    d[rows, ci] = cv
    return d


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spgemm_matches_dense_integer_product(seed):
    rng = np.random.default_rng(seed)
    A = random_int_csr(rng, 37, 23, 0.3)
    B = random_int_csr(rng, 23, 41, 0.3)
    cp, ci, cv = oracle.spgemm(A, B)
    D = A.to_dense() @ B.to_dense()
    assert np.array_equal(_dense_of(37, cp, ci, cv, 41), D)
    # structure = boolean product, structural zeros kept
    S = (A.to_dense() != 0).astype(np.int64) @ (B.to_dense() != 0).astype(np.int64)
    assert int(cp[-1]) == int((S > 0).sum())
    for r in range(37):
        cols = ci[cp[r]:cp[r + 1]]
        assert np.all(np.diff(cols) > 0)
        assert set(cols.tolist()) == set(np.nonzero(S[r])[0].tolist())


def test_spgemm_structural_zero_kept():
    A = gen.from_arrays(1, 2, [0, 2], [0, 1], [1.0, 1.0])
    B = gen.from_arrays(2, 3, [0, 2, 3], [0, 2, 0], [2.5, 1.0, -2.5])
    cp, ci, cv = oracle.spgemm(A, B)
    assert cp.tolist() == [0, 2] and ci.tolist() == [0, 2]
    assert cv[0] == 0.0 and not np.signbit(cv[0]) and cv[1] == 1.0


def test_spgemm_closed_forms():
    A = gen.rmat(9, 8, seed=21)
    n = A.nrows
    I = gen.from_arrays(n, n, np.arange(n + 1), np.arange(n), np.ones(n))
    for X, Y in ((I, A), (A, I)):
        cp, ci, cv = oracle.spgemm(X, Y)
        assert cp.tolist() == A.ptr.tolist() and ci.tolist() == A.idx.tolist() and cv.tolist() == A.val.tolist()
    perm = np.random.default_rng(4).permutation(n)
    P = gen.from_arrays(n, n, np.arange(n + 1), perm, np.ones(n))
    cp, ci, cv = oracle.spgemm(P, A)
    for r in range(n):
        s, e = A.ptr[perm[r]], A.ptr[perm[r] + 1]
        assert ci[cp[r]:cp[r + 1]].tolist() == A.idx[s:e].tolist()
        assert cv[cp[r]:cp[r + 1]].tolist() == A.val[s:e].tolist()
    dvals = np.random.default_rng(5).uniform(-2, 2, n)
    D = gen.from_arrays(n, n, np.arange(n + 1), np.arange(n), dvals)
    cp, ci, cv = oracle.spgemm(D, A)
    assert np.array_equal(cv, A.val * np.repeat(dvals, np.diff(A.ptr)))
    Z = gen.from_arrays(n, n, np.zeros(n + 1), [], [])
    cp, ci, cv = oracle.spgemm(A, Z)
    assert int(cp[-1]) == 0


def test_spgemm_canonical_fma_order_exact():
    # every output equals the fma chain over A's stored order, evaluated with
    # exact rationals; a reordered sum or mul+add rounding fails here
    rng = random.Random(7)
    for trial in range(40):
        m, k, n = 3, 12, 4
        rows = []
        for _ in range(m):
            ks = rng.sample(range(k), rng.randint(1, k))  # stored (unsorted) order
            rows.append(ks)
        aptr = [0]
        aidx, aval = [], []
        for ks in rows:
            aidx += ks
            aval += [rng.uniform(-1, 1) for _ in ks]
            aptr.append(len(aidx))
        bptr = [0]
        bidx, bval = [], []
        for _ in range(k):
            cs = sorted(rng.sample(range(n), rng.randint(0, n)))
            bidx += cs
            bval += [rng.uniform(-1, 1) for _ in cs]
            bptr.append(len(bidx))
        A = gen.from_arrays(m, k, aptr, aidx, aval)
        B = gen.from_arrays(k, n, bptr, bidx, bval)
        cp, ci, cv = oracle.spgemm(A, B)
        for i in range(m):
            acc = {}
            for p in range(aptr[i], aptr[i + 1]):
                kk, a = aidx[p], aval[p]
                for q in range(bptr[kk], bptr[kk + 1]):
                    acc[bidx[q]] = fma_exact(a, bval[q], acc.get(bidx[q], 0.0))
            got = dict(zip(ci[cp[i]:cp[i + 1]].tolist(), cv[cp[i]:cp[i + 1]].tolist()))
            assert got == acc


def test_spgemm_reference_fixtures():
    # small_matrices/*.bmtx from the reference; A*A sizes worked out by hand
    # for the dense 16x16 and by the boolean product for n1024
    A = read_bmtx(os.path.join(GOLDEN, "n16_full.bmtx"))
    assert (A.nrows, A.ncols, A.nnz) == (16, 16, 256)
    assert oracle.spgemm_flops(A, A) == 16 * 16 * 16
    cp, ci, cv = oracle.spgemm(A, A)
    assert int(cp[-1]) == 256
    assert np.allclose(_dense_of(16, cp, ci, cv, 16), A.to_dense() @ A.to_dense(), rtol=1e-14, atol=0)
    B = read_bmtx(os.path.join(GOLDEN, "n1024.bmtx"))
    assert (B.nrows, B.ncols, B.nnz) == (1024, 1024, 2048)
    assert oracle.spgemm_flops(B, B) == 4144
    cp, ci, cv = oracle.spgemm(B, B)
    assert int(cp[-1]) == 4136
    S = (B.to_dense() != 0).astype(np.int64) @ (B.to_dense() != 0).astype(np.int64)
    assert int((S > 0).sum()) == 4136


def test_spgemm_er1024_against_scipy():
    sp = pytest.importorskip("scipy.sparse")
    A = gen.erdos_renyi(1024, 8, seed=0)
    cp, ci, cv = oracle.spgemm(A, A)
    S = sp.csr_matrix((A.val, A.idx, A.ptr), shape=(1024, 1024))
    R = (S @ S).tocsr()
    R.sort_indices()
    # U[-1,1) values: no exact cancellation expected, so scipy's pattern is
    # the full boolean product here
    assert int(cp[-1]) == R.nnz
    assert np.array_equal(ci, R.indices) and np.array_equal(cp, R.indptr)
    assert np.allclose(cv, R.data, rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------- spadd / transpose
def test_spadd_matches_dense_and_merge():
    rng = np.random.default_rng(3)
    A = random_int_csr(rng, 50, 40, 0.2)
    B = random_int_csr(rng, 50, 40, 0.2)
    cp, ci, cv = oracle.spadd(A, B)
    D = A.to_dense() + B.to_dense()
    assert np.array_equal(_dense_of(50, cp, ci, cv, 40), D)
    for r in range(50):
        want = {}
        for c, v in zip(A.idx[A.ptr[r]:A.ptr[r + 1]], A.val[A.ptr[r]:A.ptr[r + 1]]):
            want[int(c)] = v
        for c, v in zip(B.idx[B.ptr[r]:B.ptr[r + 1]], B.val[B.ptr[r]:B.ptr[r + 1]]):
            want[int(c)] = want[int(c)] + v if int(c) in want else v
        got = dict(zip(ci[cp[r]:cp[r + 1]].tolist(), cv[cp[r]:cp[r + 1]].tolist()))
        assert got == want


def test_transpose_dense():
    A = gen.rmat(8, 6, seed=9)
    tp, ti, tv = oracle.transpose(A)
    T = gen.from_arrays(A.ncols, A.nrows, tp, ti, tv)
    assert np.array_equal(T.to_dense(), A.to_dense().T)
    assert oracle.rows_sorted(T)


def test_rows_sorted_rejects():
    # strictly increasing columns within each row, rows independent: a swap,
    # a repeated column or a descent anywhere in any row makes it false, and
    # a descent across a row boundary does not
    ok = gen.from_arrays(3, 9, [0, 3, 3, 6], [1, 4, 8, 0, 2, 5], np.ones(6))
    assert oracle.rows_sorted(ok)
    for idx in ([4, 1, 8, 0, 2, 5], [1, 4, 4, 0, 2, 5], [1, 4, 8, 0, 5, 2], [1, 4, 8, 2, 2, 5]):
        assert not oracle.rows_sorted(gen.from_arrays(3, 9, [0, 3, 3, 6], idx, np.ones(6))), idx
    assert oracle.rows_sorted(gen.from_arrays(2, 9, [0, 2, 4], [7, 8, 0, 1], np.ones(4)))
    # one descent deep in a large random matrix
    R = gen.rmat(10, 8, seed=4)
    idx = R.idx.copy()
    r = int(np.argmax(np.diff(R.ptr) >= 3))
    p = int(R.ptr[r]) + 1
    idx[p], idx[p + 1] = idx[p + 1], idx[p]
    assert oracle.rows_sorted(R) and not oracle.rows_sorted(gen.from_arrays(R.nrows, R.ncols, R.ptr, idx, R.val))


# ---------------------------------------------------------------- generators
def test_generators_shapes_and_sortedness():
    A = gen.erdos_renyi(1024, 8, seed=0)
    assert A.nrows == 1024 and 8000 < A.nnz <= 8192 and oracle.rows_sorted(A)
    R = gen.rmat(12, 16, seed=1)
    assert R.nrows == 4096 and oracle.rows_sorted(R)
    assert gen.rmat(12, 16, seed=1).idx.tolist() == R.idx.tolist()  # deterministic
    L = gen.laplace27(8)
    assert L.nnz == sum((1 + (x > 0) + (x < 7)) for x in range(8)) ** 3
    P = gen.sa_prolongator(9)
    assert (P.nrows, P.ncols) == (729, 27) and oracle.rows_sorted(P)
    # P = (I - w D^-1 A) P0 checked densely
    Ad = gen.laplace27(9).to_dense()
    agg = np.array([(x // 3) + 3 * ((y // 3) + 3 * (z // 3)) for z in range(9) for y in range(9) for x in range(9)])
    P0 = np.zeros((729, 27))
    P0[np.arange(729), agg] = 1.0
    Pd = P0 - (2.0 / 3.0) / 26.0 * (Ad @ P0)
    assert np.allclose(P.to_dense(), Pd, rtol=0, atol=1e-15)
