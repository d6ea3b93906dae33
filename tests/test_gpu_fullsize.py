"""Full-size parity through a size-independent property: exact linear
sketches.  With small integer values every product and partial sum below is
an integer far below 2^53, so fp64 computes it exactly in any order, and
C V must equal A (B V) bit for bit, structure included.  V is an n x 64
sketch with one random integer in [-3, 3] per row j, in column j mod 64, so
every value and every column of C enters the check; two sketches per
product, at the benchmark's full sizes, where the CPU oracle would take
minutes:
  - C = A A, A = R-MAT scale 20 (edge factor 16, seed 1), 9.7e9 entries
    returned as five int32 row panels;
  - C = P^T (A P), the 27-point Laplacian 128^3 (26 / -1) with its SA
    prolongator's structure;
  - a batched stage (R-MAT scale 20 cut 4 x 4, stage 3) added into partial-C
    tiles: C_out V = C_in V + A_ik (A_kj V).
The sketches are products through the library itself."""
import numpy as np
import pytest

from oracle import spadd
from trident_amd import gen
from trident_amd.device import Batch, spgemm, spgemm_panels

pytestmark = pytest.mark.gpu


def _integer_values(M, seed):
    """M's structure with values in {-2, -1, 1, 2}."""
    rng = np.random.default_rng(seed)
    v = rng.choice(np.array([-2.0, -1.0, 1.0, 2.0]), size=M.nnz)
    return gen.from_arrays(M.nrows, M.ncols, M.ptr, M.idx, v)


def _sketch(n, seed):
    """n x 64 CSR: row j holds one integer in [-3, 3] (never 0) at column j mod 64."""
    rng = np.random.default_rng(seed)
    v = rng.choice(np.array([-3.0, -2.0, -1.0, 1.0, 2.0, 3.0]), size=n)
    return gen.from_arrays(n, 64, np.arange(n + 1, dtype=np.int64), (np.arange(n) % 64).astype(np.int32), v)


def _host(ctxs):
    """Concatenate device CSR row blocks into one host CSR (ptr int64)."""
    ptr, idx, val = [np.zeros(1, np.int64)], [], []
    for M in ctxs:
        H = M.download()
        ptr.append(ptr[-1][-1] + H.ptr[1:].astype(np.int64))
        idx.append(H.idx)
        val.append(H.val)
    return np.concatenate(ptr), np.concatenate(idx), np.concatenate(val)


def _same(x, y, label):
    assert np.array_equal(x[0], y[0]), f"{label}: row pointers differ"
    assert np.array_equal(x[1], y[1]), f"{label}: columns differ"
    bad = int(np.sum(x[2] != y[2]))
    assert bad == 0, f"{label}: {bad} values differ"


def test_rmat20_product_sketches(ctx):
    A = _integer_values(gen.rmat(20, 16, seed=1), seed=21)
    print(f"full size: R-MAT s20 A*A, nnz(A) {A.nnz}", flush=True)
    dA = ctx.upload(A)
    panels, rb = spgemm_panels(dA, dA, max_panels=64)
    assert len(panels) >= 2 and rb[-1] == A.nrows
    for seed in (1, 2):
        dV = ctx.upload(_sketch(A.ncols, seed))
        CV = [spgemm(P, dV) for P in panels]
        AV = spgemm(dA, dV)
        AAV = spgemm(dA, AV)
        _same(_host(CV), _host([AAV]), f"C V vs A (A V), sketch {seed}")
        for M in CV + [AV, AAV, dV]:
            M.free()
    for P in panels:
        P.free()


def test_galerkin128_sketches(ctx):
    A = gen.laplace27(128)
    P = _integer_values(gen.sa_prolongator(128), seed=22)
    print(f"full size: Galerkin 128^3, nnz(A) {A.nnz}, nnz(P) {P.nnz}", flush=True)
    dA, dP = ctx.upload(A), ctx.upload(P)
    dPt = dP.transpose()
    AP = spgemm(dA, dP)
    C = spgemm(dPt, AP)
    for seed in (3, 4):
        dV = ctx.upload(_sketch(P.ncols, seed))
        CV = spgemm(C, dV)
        PV = spgemm(dP, dV)
        APV = spgemm(dA, PV)
        PtAPV = spgemm(dPt, APV)
        _same(_host([CV]), _host([PtAPV]), f"C V vs P^T (A (P V)), sketch {seed}")
        for M in (CV, PV, APV, PtAPV, dV):
            M.free()


def test_stage_batch_sketches(ctx):
    A = _integer_values(gen.rmat(20, 16, seed=2), seed=23)
    g, k = 4, 3
    n = A.nrows // g
    T = {(i, j): gen.submatrix(A, i * n, (i + 1) * n, j * n, (j + 1) * n) for i in range(g) for j in range(g)}
    del A
    print("full size: R-MAT s20 4x4 stage 3 batch with partial C", flush=True)
    dT = {key: ctx.upload(m) for key, m in T.items()}
    keys = [(i, j) for i in range(g) for j in range(g)]
    Cin = [spgemm(dT[(i, k - 1)], dT[(k - 1, j)]) for i, j in keys]
    b = Batch(ctx, [dT[(i, k)] for i, j in keys], [dT[(k, j)] for i, j in keys], Cin)
    b.run()
    b.check()
    for seed in (5, 6):
        dV = ctx.upload(_sketch(n, seed))
        for t, (i, j) in enumerate(keys):
            OV = spgemm(b.result(t), dV)
            BV = spgemm(dT[(k, j)], dV)
            ABV = spgemm(dT[(i, k)], BV)
            CV = spgemm(Cin[t], dV)
            x = _host([OV])
            ya, yc = _host([ABV]), _host([CV])
            # C_in V + A (B V), merged per row with the oracle's exact sum
            s_ptr, s_idx, s_val = spadd(gen.from_arrays(n, 64, *ya), gen.from_arrays(n, 64, *yc))
            _same(x, (s_ptr, s_idx, s_val), f"tile {(i, j)}, sketch {seed}")
            for M in (OV, BV, ABV, CV):
                M.free()
        dV.free()
    b.close()
