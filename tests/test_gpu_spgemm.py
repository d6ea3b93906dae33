"""GPU parity of the local multiply (tr_spgemm, tr_spgemm_panels,
tr_symbolic, tr_spgemm_flops) against the CPU oracle.

The bar: row pointers and column indices bit-exact (sorted columns) and
values bit-exact — the HIP kernels apply every product in the oracle's
canonical order, so max relative error is 0 (the contract allows 1e-12).
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from trident_amd import gen
from trident_amd.device import spgemm, spgemm_flops, spgemm_panels, symbolic
from trident_amd.formats import read_bmtx
from trident_amd._lib import TridentError, TR_ERR_INVALID, TR_ERR_OVERFLOW

pytestmark = pytest.mark.gpu


def assert_same(C, ref, label=""):
    cp, ci, cv = ref
    assert C.nrows == len(cp) - 1, label
    assert np.array_equal(C.ptr.astype(np.int64), cp), f"{label}: row pointers differ"
    assert np.array_equal(C.idx, ci), f"{label}: column indices differ"
    if not np.array_equal(C.val, cv):
        rel = np.abs(C.val - cv) / np.maximum(np.abs(cv), 1e-300)
        pytest.fail(f"{label}: values differ, max rel err {rel.max():.3e}")


def run(ctx, A, B, label=""):
    dA = ctx.upload(A)
    dB = ctx.upload(B) if B is not A else dA
    dC = spgemm(dA, dB)
    C = dC.download()
    ref = oracle.spgemm(A, B)
    assert_same(C, ref, label)
    dC.free()
    dA.free()
    if B is not A:
        dB.free()
    return C


def test_er1024_plain(ctx):
    # the CPU configuration: Erdos-Renyi n=1024, 8 columns per row, seed 0
    A = gen.erdos_renyi(1024, 8, seed=0)
    run(ctx, A, A, "er1024")


@pytest.mark.parametrize("scale,ef,seed", [(8, 4, 1), (10, 8, 2), (12, 16, 1), (14, 16, 3)])
def test_rmat(ctx, scale, ef, seed):
    A = gen.rmat(scale, ef, seed=seed)
    run(ctx, A, A, f"rmat s{scale}")


TUNINGS = [
    dict(heavy_table=4096, window_entries=1920),
    dict(heavy_table=4096, window_entries=1920, dense_width=4096),
    dict(heavy_table=4096, window_entries=64, item_windows=1),
    dict(dense_width=0),
    dict(dense_width=512, window_entries=1409),
    dict(window_entries=64, item_windows=64, light_max_flops=64),
    dict(window_entries=384, item_windows=4, light_max_flops=256),
    dict(window_entries=1409, item_windows=2),
]


@pytest.mark.parametrize("tuning", TUNINGS, ids=lambda t: "-".join(f"{k}{v}" for k, v in t.items()))
def test_tuning_does_not_change_results(ctx, tuning):
    # every cutover setting routes rows through different kernels; results
    # stay bit-identical to the oracle
    A = gen.rmat(13, 16, seed=6)
    ctx.set_tuning(dict(ctx.get_tuning(), **tuning))
    try:
        run(ctx, A, A, f"tuning {tuning}")
        B = gen.submatrix(gen.rmat(15, 16, seed=8), 0, 2000, 0, 32768)
        C = gen.rmat(15, 16, seed=9)
        run(ctx, B, C, f"tuning {tuning} wide")
    finally:
        ctx.set_tuning(None)


def test_rmat_rectangular_pair(ctx):
    A = gen.submatrix(gen.rmat(13, 16, seed=4), 0, 3000, 0, 8192)
    B = gen.submatrix(gen.rmat(13, 16, seed=5), 0, 8192, 100, 5000)
    run(ctx, A, B, "rect")


def test_reference_fixtures(ctx):
    for name in ("n16_full.bmtx", "n1024.bmtx"):
        A = read_bmtx(os.path.join(GOLDEN, name))
        run(ctx, A, A, name)


def test_closed_forms(ctx):
    A = gen.rmat(11, 8, seed=21)
    n = A.nrows
    I = gen.from_arrays(n, n, np.arange(n + 1), np.arange(n), np.ones(n))
    C = run(ctx, I, A, "I*A")
    assert np.array_equal(C.idx, A.idx) and np.array_equal(C.val, A.val)
    C = run(ctx, A, I, "A*I")
    assert np.array_equal(C.idx, A.idx) and np.array_equal(C.val, A.val)
    Z = gen.from_arrays(n, n, np.zeros(n + 1), [], [])
    C = run(ctx, A, Z, "A*0")
    assert C.nnz == 0


def test_heavy_dense_rows_multiwindow(ctx):
    # a few rows whose products cover most columns: exercises the heavy
    # symbolic bitmap, multi-window items and the in-step conflict ordering
    rng = np.random.default_rng(0)
    n = 20000
    dense_rows = 6
    B = gen.erdos_renyi(n, 40, seed=3)
    rows = []
    for r in range(dense_rows):
        rows.append(np.sort(rng.choice(n, size=3000 + 1000 * r, replace=False)))
    rows += [np.sort(rng.choice(n, size=3, replace=False)) for _ in range(100)]
    ptr = np.concatenate([[0], np.cumsum([len(x) for x in rows])])
    A = gen.from_arrays(len(rows), n, ptr, np.concatenate(rows), rng.uniform(-1, 1, ptr[-1]))
    run(ctx, A, B, "dense rows")


def test_heavy_multipass_wide(ctx):
    # more than 2^20 columns: the heavy bitmap runs several column passes
    rng = np.random.default_rng(1)
    ncols = (1 << 20) + 5000
    k = 300
    bl = []
    for r in range(k):
        cnt = 20 if r % 7 else 4000
        bl.append(np.sort(rng.choice(ncols, size=cnt, replace=False)))
    bptr = np.concatenate([[0], np.cumsum([len(x) for x in bl])])
    B = gen.from_arrays(k, ncols, bptr, np.concatenate(bl), rng.uniform(-1, 1, bptr[-1]))
    al = [np.sort(rng.choice(k, size=int(rng.integers(1, 200)), replace=False)) for _ in range(64)]
    aptr = np.concatenate([[0], np.cumsum([len(x) for x in al])])
    A = gen.from_arrays(64, k, aptr, np.concatenate(al), rng.uniform(-1, 1, aptr[-1]))
    run(ctx, A, B, "multipass")


@pytest.mark.parametrize("ncols", [1 << 19, (1 << 20) + 777])
def test_heavy_sparse_rows(ctx, ncols):
    # heavy rows whose few thousand outputs spread over all columns: their
    # column passes are sparse, so windows are cut by count only (wider than
    # the bitmap) and ranked by key comparison; dense rows in the same matrix
    # keep bitmap windows; overlapping B rows give collisions inside windows
    rng = np.random.default_rng(11)
    k = 400
    bl = []
    for r in range(k):
        cnt = 3000 if r % 50 == 0 else int(rng.integers(30, 200))
        bl.append(np.sort(rng.choice(ncols, size=cnt, replace=False)))
    bptr = np.concatenate([[0], np.cumsum([len(x) for x in bl])])
    B = gen.from_arrays(k, ncols, bptr, np.concatenate(bl), rng.uniform(-1, 1, bptr[-1]))
    al = []
    for i in range(96):
        size = int(rng.integers(8, 60)) if i % 3 else int(rng.integers(100, 300))
        al.append(np.sort(rng.choice(k, size=size, replace=False)))
    aptr = np.concatenate([[0], np.cumsum([len(x) for x in al])])
    A = gen.from_arrays(96, k, aptr, np.concatenate(al), rng.uniform(-1, 1, aptr[-1]))
    run(ctx, A, B, f"sparse heavy {ncols}")


@pytest.mark.parametrize("tuning", [None, dict(heavy_table=4096, window_entries=2816)],
                         ids=["table2048", "table4096"])
def test_heavy_window_bitmap_pass_edges(ctx, tuning):
    # windows up to 2^19 columns wide are ranked one 32768-column presence
    # bitmap pass at a time: entries on both sides of every pass edge, at
    # the window-width quantum and at the matrix's last column must keep
    # their sorted places.  Rows of 40..400 A entries put both row classes
    # (2^19- and 2^20-column symbolic passes) over windows of 3 to 16 passes,
    # with both heavy table sizes.
    if tuning:
        ctx.set_tuning(dict(ctx.get_tuning(), **tuning))
    try:
        _pass_edges(ctx)
    finally:
        ctx.set_tuning(None)


def _pass_edges(ctx):
    rng = np.random.default_rng(23)
    ncols = (1 << 20) + 300
    edges = np.arange(32768, ncols, 32768)
    special = np.unique(np.concatenate([edges - 1, edges, [0, ncols - 1]]))
    k = 500
    bl = []
    for _ in range(k):
        extra = rng.choice(ncols, size=int(rng.integers(10, 60)), replace=False)
        pick = special[rng.random(special.size) < 0.3]
        bl.append(np.unique(np.concatenate([extra, pick])))
    bptr = np.concatenate([[0], np.cumsum([len(x) for x in bl])])
    bval = rng.uniform(-1, 1, bptr[-1]) * np.exp2(rng.integers(-20, 21, bptr[-1]))
    B = gen.from_arrays(k, ncols, bptr, np.concatenate(bl), bval)
    al = [np.sort(rng.choice(k, size=int(rng.integers(40, 200 if i % 2 else 400)), replace=False))
          for i in range(48)]
    aptr = np.concatenate([[0], np.cumsum([len(x) for x in al])])
    A = gen.from_arrays(48, k, aptr, np.concatenate(al), rng.uniform(-1, 1, aptr[-1]))
    C = run(ctx, A, B, "bitmap pass edges")
    assert np.isin(special, C.idx).all()


def test_galerkin_operators(ctx):
    # the AMG products A*P and P^T*(A*P): light rows whose B rows hold about
    # 4.5 entries, multiplied with their products flattened over the lanes
    n = 20
    A, P = gen.laplace27(n), gen.sa_prolongator(n)
    AP = run(ctx, A, P, "A*P")
    run(ctx, gen.transpose(P), AP, "P^T*(A*P)")


@pytest.mark.parametrize("per_row", [64, 40, 20], ids=["owner512", "mask256", "mask128"])
def test_light_short_rows_ordered(ctx, per_row):
    # per_row A entries per row over B rows of 3 entries in 16 columns: a
    # round of up to 64 flattened products meets each column several times,
    # and the values span 2^-20..2^20, so any application out of A-order
    # changes bits.
    # 192 products per row take 512-slot tables (per-slot owner rounds); 120
    # and 60 take 256- and 128-slot tables (per-slot lane masks)
    rng = np.random.default_rng(17)
    k, ncols = 400, 16
    bl = [np.sort(rng.choice(ncols, size=3, replace=False)) for _ in range(k)]
    bptr = np.concatenate([[0], np.cumsum([len(x) for x in bl])])
    bval = rng.uniform(-1, 1, bptr[-1]) * np.exp2(rng.integers(-20, 21, bptr[-1]))
    B = gen.from_arrays(k, ncols, bptr, np.concatenate(bl), bval)
    m = 2000
    al = [np.sort(rng.choice(k, size=per_row, replace=False)) for _ in range(m)]
    aptr = np.concatenate([[0], np.cumsum([len(x) for x in al])])
    aval = rng.uniform(-1, 1, aptr[-1]) * np.exp2(rng.integers(-20, 21, aptr[-1]))
    A = gen.from_arrays(m, k, aptr, np.concatenate(al), aval)
    run(ctx, A, B, "short B rows")


def test_heavy_ordered_collisions(ctx):
    # Heavy rows of 1500 A entries over B rows that all hold four hub
    # columns, a few of 32 semi-hub columns and random ones: each 256-product
    # step of a row's first window meets every hub slot about 17 times and
    # the semi-hub slots once to three times, so the step's slot lists take
    # every apply path (one, two, three and longer products).  Values span
    # 2^-20..2^20.  The same products summed in the reverse A-order (A's
    # columns relabelled k -> K-1-k, B's rows reversed to match) differ in
    # most hub and semi-hub entries, so the bit-exact match pins the order.
    rng = np.random.default_rng(29)
    k, ncols, m = 3000, 40000, 48
    bl = []
    for _ in range(k):
        semi = np.nonzero(rng.random(32) < 0.1)[0] + 4
        rand = rng.choice(np.arange(36, ncols), size=8, replace=False)
        bl.append(np.sort(np.concatenate([np.arange(4), semi, rand])))
    bptr = np.concatenate([[0], np.cumsum([len(x) for x in bl])])
    bval = rng.uniform(-1, 1, bptr[-1]) * np.exp2(rng.integers(-20, 21, bptr[-1]))
    B = gen.from_arrays(k, ncols, bptr, np.concatenate(bl), bval)
    al = [np.sort(rng.choice(k, size=1500, replace=False)) for _ in range(m)]
    aptr = np.concatenate([[0], np.cumsum([len(x) for x in al])])
    aidx = np.concatenate(al)
    aval = rng.uniform(-1, 1, aptr[-1]) * np.exp2(rng.integers(-20, 21, aptr[-1]))
    A = gen.from_arrays(m, k, aptr, aidx, aval)
    C = run(ctx, A, B, "heavy ordered collisions")
    ridx = np.concatenate([(k - 1 - aidx[aptr[i]:aptr[i + 1]])[::-1] for i in range(m)])
    rval = np.concatenate([aval[aptr[i]:aptr[i + 1]][::-1] for i in range(m)])
    Ar = gen.from_arrays(m, k, aptr, ridx, rval)
    rl = bl[::-1]
    rptr = np.concatenate([[0], np.cumsum([len(x) for x in rl])])
    rbval = np.concatenate([bval[bptr[i]:bptr[i + 1]] for i in range(k - 1, -1, -1)])
    Br = gen.from_arrays(k, ncols, rptr, np.concatenate(rl), rbval)
    rp, ri, rv = oracle.spgemm(Ar, Br)
    assert np.array_equal(rp, C.ptr.astype(np.int64)) and np.array_equal(ri, C.idx)
    multi = C.idx < 36  # hub and semi-hub columns: the ones with several products
    frac = float(np.mean(rv[multi] != C.val[multi]))
    print(f"reverse A-order changes {frac:.1%} of the hub and semi-hub entries", flush=True)
    assert frac > 0.5


@pytest.mark.parametrize("tuning", [dict(dense_width=0), dict(dense_width=2048),
                                    dict(heavy_table=4096, window_entries=2816, dense_width=4096)],
                         ids=["hashed", "dense2048", "dense4096"])
def test_dense_accumulator_windows(ctx, tuning):
    # Heavy rows whose output is dense over a few thousand columns: windows
    # of about 1150 (or 2800) entries span fewer columns than the table has
    # slots, so with dense_width set they take the direct-addressed
    # accumulator; every eighth B row adds a column far out, so each row also
    # has a wide sparse window that keeps hashing.  Hub columns 0..3 sit in
    # every B row (long slot lists in every step).  Hashed and direct runs
    # must both match the oracle bit for bit.
    rng = np.random.default_rng(41)
    k, ncols, m = 2500, 60000, 40
    bl = []
    for i in range(k):
        cols = np.concatenate([np.arange(4), rng.choice(np.arange(4, 3000), size=36, replace=False)])
        if i % 8 == 0:
            cols = np.concatenate([cols, [rng.integers(3000, ncols)]])
        bl.append(np.sort(cols))
    bptr = np.concatenate([[0], np.cumsum([len(x) for x in bl])])
    bval = rng.uniform(-1, 1, bptr[-1]) * np.exp2(rng.integers(-20, 21, bptr[-1]))
    B = gen.from_arrays(k, ncols, bptr, np.concatenate(bl), bval)
    al = [np.sort(rng.choice(k, size=int(rng.integers(300, 1200)), replace=False)) for _ in range(m)]
    aptr = np.concatenate([[0], np.cumsum([len(x) for x in al])])
    aval = rng.uniform(-1, 1, aptr[-1]) * np.exp2(rng.integers(-20, 21, aptr[-1]))
    A = gen.from_arrays(m, k, aptr, np.concatenate(al), aval)
    ctx.set_tuning(dict(ctx.get_tuning(), **tuning))
    try:
        run(ctx, A, B, f"dense accumulator {tuning}")
    finally:
        ctx.set_tuning(None)


def test_degenerate_shapes(ctx):
    # zero rows, an empty inner dimension, rows with nothing, 1 x 1
    rng = np.random.default_rng(31)
    B = gen.rmat(8, 8, seed=5)
    A0 = gen.from_arrays(0, B.nrows, [0], [], [])
    C = run(ctx, A0, B, "0-row A")
    assert C.nrows == 0 and C.nnz == 0
    Ak = gen.from_arrays(5, 0, np.zeros(6), [], [])
    Bk = gen.from_arrays(0, 7, [0], [], [])
    C = run(ctx, Ak, Bk, "empty inner dimension")
    assert (C.nrows, C.ncols, C.nnz) == (5, 7, 0)
    ptr = np.zeros(B.nrows + 1, np.int64)
    ptr[101:] = 3
    Ae = gen.from_arrays(B.nrows, B.nrows, ptr, [3, 7, 200], rng.uniform(-1, 1, 3))  # one non-empty row
    run(ctx, Ae, B, "all rows empty but one")
    run(ctx, gen.from_arrays(1, 1, [0, 1], [0], [3.0]), gen.from_arrays(1, 1, [0, 1], [0], [-0.5]), "1 x 1")


@pytest.mark.parametrize("per_row", [40, 300, 3000], ids=["light", "mid", "heavy"])
def test_single_column_chains(ctx, per_row):
    # every B row holds only column 0, so each C row is one entry: the fma
    # chain of all its products in A-order (per_row of them, values over
    # 2^-20..2^20), taken by the light path, the wave-flattened short rows
    # and the compact heavy rows
    rng = np.random.default_rng(per_row)
    k = 4000
    B = gen.from_arrays(k, 1, np.arange(k + 1), np.zeros(k, np.int32),
                        rng.uniform(-1, 1, k) * np.exp2(rng.integers(-20, 21, k)))
    m = 64
    al = [np.sort(rng.choice(k, size=per_row, replace=False)) for _ in range(m)]
    aptr = np.concatenate([[0], np.cumsum([len(x) for x in al])])
    aval = rng.uniform(-1, 1, aptr[-1]) * np.exp2(rng.integers(-20, 21, aptr[-1]))
    A = gen.from_arrays(m, k, aptr, np.concatenate(al), aval)
    C = run(ctx, A, B, f"single column, {per_row} products per entry")
    assert C.nnz == m


def _device_bytes_equal(ctx, pa, pb, nbytes, chunk=1 << 29):
    """Byte equality of two device ranges, copied to the host in chunks."""
    ha = np.empty(max(1, min(chunk, nbytes)), np.uint8)
    hb = np.empty_like(ha)
    off = 0
    while off < nbytes:
        n = min(chunk, nbytes - off)
        ctx.check(ctx.lib.tr_memcpy_d2h(ctx.h, ha.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(pa + off), n))
        ctx.check(ctx.lib.tr_memcpy_d2h(ctx.h, hb.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(pb + off), n))
        if not np.array_equal(ha[:n], hb[:n]):
            return False
        off += n
    return True


def test_heavy_items_beyond_grid_cap(ctx):
    # R-MAT scale 19 with 64-entry windows, one window per heavy item.  A
    # window holds fewer than 384 entries, so (heavy-row entries) / 384
    # bounds the item count from below; it exceeds 2^22, the cap of the heavy
    # numeric grid (one workgroup per item, and a dispatch counts its
    # work-items in 32 bits), so workgroups stride over the items.  Every
    # cutover keeps the result bit-identical: it must equal the default
    # tuning's panels, and the oracle on sampled rows (the 16 heaviest
    # among them).
    A = gen.rmat(19, 16, seed=1)
    dA = ctx.upload(A)
    nnz, _ = symbolic(dA, dA)
    light_entries_max = A.nrows * 1024  # a light row has <= 1024 products
    assert (nnz - light_entries_max) // 384 > 2 ** 22
    ref, rb_ref = spgemm_panels(dA, dA, max_panels=64)
    ctx.set_tuning(dict(ctx.get_tuning(), window_entries=64, item_windows=1))
    try:
        got, rb_got = spgemm_panels(dA, dA, max_panels=64)
    finally:
        ctx.set_tuning(None)
    assert rb_got == rb_ref and sum(p.nnz for p in got) == nnz
    for p, q in zip(got, ref):
        assert (p.d.nrows, p.nnz) == (q.d.nrows, q.nnz)
        assert _device_bytes_equal(ctx, p.d.ptr_vec, q.d.ptr_vec, 4 * (p.d.nrows + 1)), "row pointers differ"
        assert _device_bytes_equal(ctx, p.d.idx_vec, q.d.idx_vec, 4 * p.nnz), "column indices differ"
        assert _device_bytes_equal(ctx, p.d.val, q.d.val, 8 * p.nnz), "values differ"
    # and the oracle agrees on sampled rows (heavy ones included)
    rng = np.random.default_rng(19)
    heavy = np.argsort(np.diff(A.ptr.astype(np.int64)))[-16:]
    rows = np.unique(np.concatenate([heavy, rng.choice(A.nrows, 240, replace=False)]))
    lens = np.diff(A.ptr.astype(np.int64))[rows]
    sp = np.zeros(len(rows) + 1, np.int64)
    np.cumsum(lens, out=sp[1:])
    sel = np.concatenate([np.arange(A.ptr[r], A.ptr[r + 1]) for r in rows])
    cp, ci, cv = oracle.spgemm(gen.from_arrays(len(rows), A.ncols, sp, A.idx[sel], A.val[sel]), A)
    for t, r in enumerate(rows):
        pn = int(np.searchsorted(rb_got, r, side="right") - 1)
        P = got[pn]
        pp = P.ctx.to_host(P.d.ptr_vec + 4 * int(r - rb_got[pn]), (2,), np.int32)
        s0, e0 = int(pp[0]), int(pp[1])
        gi = P.ctx.to_host(P.d.idx_vec + 4 * s0, (e0 - s0,), np.int32)
        gv = P.ctx.to_host(P.d.val + 8 * s0, (e0 - s0,), np.float64)
        assert np.array_equal(gi, ci[cp[t]:cp[t + 1]]), f"row {r}: columns differ from the oracle"
        assert np.array_equal(gv.view(np.uint64), cv[cp[t]:cp[t + 1]].view(np.uint64)), f"row {r}: values differ"
    for p in got + ref:
        p.free()
    dA.free()


def test_structural_zero_and_unsorted_A(ctx):
    # products cancelling to exactly 0 keep their entry; A's stored order
    # (unsorted within a row) defines the accumulation order
    A = gen.from_arrays(2, 3, [0, 2, 5], [0, 1, 2, 0, 1], [1.0, 1.0, 0.1, 0.7, -0.3])
    B = gen.from_arrays(3, 3, [0, 2, 3, 5], [0, 2, 0, 1, 2], [2.5, 1.0, -2.5, 0.25, 3.0])
    dA = ctx.upload(A)
    dB = ctx.upload(B)
    with pytest.raises(TridentError) as e:
        spgemm(dA, dB)  # A unsorted: rejected with a clear error
    assert e.value.code == TR_ERR_INVALID
    dA.sort_rows()
    A_sorted = dA.download()
    dC = spgemm(dA, dB)
    C = dC.download()
    assert_same(C, oracle.spgemm(A_sorted, B), "zeros")
    assert C.idx[:2].tolist() == [0, 2] and C.val[0] == 0.0


def test_flops_and_symbolic(ctx):
    A = gen.rmat(13, 16, seed=8)
    dA = ctx.upload(A)
    assert spgemm_flops(dA, dA) == oracle.spgemm_flops(A, A)
    nnz, nbytes = symbolic(dA, dA)
    assert nnz == int(oracle.spgemm_rownnz(A, A).sum())
    assert nbytes >= nnz * 12


def test_panels(ctx):
    A = gen.rmat(12, 16, seed=6)
    ref = oracle.spgemm(A, A)
    total = int(ref[0][-1])
    dA = ctx.upload(A)
    cap = total // 5 + 1
    panels, rb = spgemm_panels(dA, dA, max_panel_nnz=cap)
    assert len(panels) >= 5 and rb[0] == 0 and rb[-1] == A.nrows
    ptrs, idxs, vals = [], [], []
    base = 0
    for p, P in enumerate(panels):
        H = P.download()
        assert H.nrows == rb[p + 1] - rb[p] and H.nnz <= cap
        ptrs.append(H.ptr[:-1].astype(np.int64) + base)
        base += H.nnz
        idxs.append(H.idx)
        vals.append(H.val)
        P.free()
    ptr = np.concatenate(ptrs + [np.array([base])])
    assert np.array_equal(ptr, ref[0])
    assert np.array_equal(np.concatenate(idxs), ref[1])
    assert np.array_equal(np.concatenate(vals), ref[2])
    # a single tile cannot hold it when the cap is below nnz
    with pytest.raises(TridentError) as e:
        spgemm_panels(dA, dA, max_panel_nnz=cap, max_panels=2)
    assert e.value.code == TR_ERR_OVERFLOW


def test_invalid_operands_rejected(ctx):
    A = gen.rmat(8, 4, seed=1)
    B = gen.rmat(9, 4, seed=1)
    dA = ctx.upload(A)
    dB = ctx.upload(B)
    with pytest.raises(TridentError) as e:
        spgemm(dA, dB)  # 256 cols vs 512 rows
    assert e.value.code == TR_ERR_INVALID
    bad = gen.from_arrays(A.nrows, A.ncols, A.ptr, np.where(A.idx == A.idx.max(), 10**6, A.idx), A.val)
    dbad = ctx.upload(bad)
    with pytest.raises(TridentError) as e:
        spgemm(dbad, dA)
    assert e.value.code == TR_ERR_INVALID


@pytest.mark.parametrize("width", [24, 60, 120, 136, 400])
def test_heavy_rows_with_few_entries(ctx, width):
    # Rows with thousands of multiplications but few distinct columns: a
    # one-pass heavy row with <= 128 entries goes to the light numeric (table
    # of twice its entries, lane-mask ordering), wider ones keep windows.
    # Every B row draws 40 columns from a band of `width`, so each output sums
    # about 40 * 96 / width products in A-order.
    rng = np.random.default_rng(width)
    n, band = 4096, width
    def rows_from(cols_per_row, lo, hi, nrows):
        idx = [np.sort(rng.choice(np.arange(lo, hi), size=cols_per_row, replace=False)) for _ in range(nrows)]
        ptr = np.arange(0, (nrows + 1) * cols_per_row, cols_per_row, dtype=np.int64)
        return ptr, np.concatenate(idx).astype(np.int32)
    ap, ai = rows_from(96, 0, n, 256)
    A = gen.from_arrays(256, n, ap, ai, rng.uniform(-1, 1, ai.size))
    bp, bi = rows_from(min(40, band), 1000, 1000 + band, n)
    B = gen.from_arrays(n, n, bp, bi, rng.uniform(-1, 1, bi.size))
    print(f"few-entry heavy rows: band {width}", flush=True)
    run(ctx, A, B, f"band {width}")
