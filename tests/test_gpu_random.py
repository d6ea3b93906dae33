"""Randomised GPU parity sweep of the multiply and the multiply-add against
the CPU oracle, bit for bit.

Each seed draws the shapes (1..2000 rows; inner dimension up to 3000;
output widths from a handful of columns to past 2^20, which takes the heavy
symbolic's multi-pass path), heavy-tailed row lengths with empty rows, values
spread over 2^-20..2^20 (so any change of the fma order shows), and one
multiply tuning.  The sp_mma chain accumulates three products of different
patterns into one C, as a stage loop does."""
import numpy as np
import pytest

import oracle
from trident_amd import gen
from trident_amd.device import DeviceCsx, sp_mma, spgemm

pytestmark = pytest.mark.gpu

TUNINGS = [
    None,
    dict(dense_width=0),
    dict(light_max_flops=64, window_entries=64, item_windows=1),
    dict(light_max_flops=256, window_entries=384, item_windows=4, dense_width=512),
    dict(heavy_table=4096, window_entries=1920, dense_width=4096),
    dict(heavy_table=4096, window_entries=64, item_windows=64, dense_width=0),
]


def _values(rng, n):
    return rng.uniform(-1.0, 1.0, n) * np.exp2(rng.integers(-20, 21, n))


def _random_csr(rng, m, n, mean_len, heavy_frac):
    """m x n CSR, sorted unique columns: lognormal row lengths around
    mean_len, a share heavy_frac of rows 20x longer, a fifth of rows empty."""
    lens = rng.lognormal(np.log(max(mean_len, 1.0)), 1.0, m)
    lens[rng.random(m) < heavy_frac] *= 20
    lens[rng.random(m) < 0.2] = 0
    lens = np.minimum(lens.astype(np.int64), n)
    rows = []
    for ln in lens:
        if ln == 0:
            rows.append(np.empty(0, np.int64))
        elif ln * 4 >= n:
            rows.append(np.sort(rng.choice(n, size=int(ln), replace=False)))
        else:
            rows.append(np.unique(rng.integers(0, n, size=int(ln))))
    ptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    idx = np.concatenate(rows) if ptr[-1] else np.empty(0, np.int64)
    return gen.from_arrays(m, n, ptr, idx, _values(rng, int(ptr[-1])))


def _shapes(rng):
    m = int(rng.integers(1, 2001))
    k = int(rng.integers(1, 3001))
    n = int(rng.choice([int(rng.integers(1, 65)), int(rng.integers(1000, 5001)),
                        int(rng.integers(200_000, (1 << 20) + 5000))]))
    return m, k, n


def _check(C, ref, label):
    cp, ci, cv = ref
    assert np.array_equal(C.ptr.astype(np.int64), cp), f"{label}: row pointers differ"
    assert np.array_equal(C.idx, ci), f"{label}: column indices differ"
    assert np.array_equal(C.val, cv), f"{label}: values differ"


@pytest.mark.parametrize("seed", range(16))
def test_random_products(ctx, seed):
    rng = np.random.default_rng(1000 + seed)
    m, k, n = _shapes(rng)
    A = _random_csr(rng, m, k, float(rng.uniform(1, 60)), 0.05)
    B = _random_csr(rng, k, n, float(rng.uniform(1, 40)), 0.05)
    tuning = TUNINGS[seed % len(TUNINGS)]
    label = f"seed {seed}: {m}x{k} * {k}x{n}, nnz {A.nnz}/{B.nnz}, tuning {tuning}"
    print(label, flush=True)
    if tuning:
        ctx.set_tuning(dict(ctx.get_tuning(), **tuning))
    try:
        dA, dB = ctx.upload(A), ctx.upload(B)
        dC = spgemm(dA, dB)
        _check(dC.download(), oracle.spgemm(A, B), label)
        dC.free()
        dA.free()
        dB.free()
    finally:
        ctx.set_tuning(None)


@pytest.mark.parametrize("seed", range(6))
def test_random_sp_mma_chain(ctx, seed):
    rng = np.random.default_rng(2000 + seed)
    m, _, n = _shapes(rng)
    C = DeviceCsx(ctx)
    ref = None
    for s in range(3):
        k = int(rng.integers(1, 2001))
        A = _random_csr(rng, m, k, float(rng.uniform(1, 40)), 0.05)
        B = _random_csr(rng, k, n, float(rng.uniform(1, 30)), 0.05)
        print(f"seed {seed} stage {s}: {m}x{k} * {k}x{n}, nnz {A.nnz}/{B.nnz}", flush=True)
        dA, dB = ctx.upload(A), ctx.upload(B)
        sp_mma(dA, dB, C)
        dA.free()
        dB.free()
        pp, pi, pv = oracle.spgemm(A, B)
        P = gen.from_arrays(m, n, pp, pi, pv)
        if ref is None:
            ref = P  # an uninitialised C becomes the first product, empty or not
        elif pp[-1] > 0:  # an empty product leaves C as it was
            cp, ci, cv = oracle.spadd(P, ref)
            ref = gen.from_arrays(m, n, cp, ci, cv)
    _check(C.download(), (ref.ptr.astype(np.int64), ref.idx, ref.val), f"sp_mma chain seed {seed}")
