"""The reference's stage loop written against the drop-in C++ headers
(tests/cpp/test_stage_loop.cpp: TileHolder, SpaCommHandler::Compress,
CusparseCSX(&handle, csx, conversion_buffs), LocalMatrix,
sp_mma_hybrid / AccumThreadHandle), run for every rank of a 4 x 4 grid in
Cannon order and checked bit-exactly against the oracle: each rank's partial
C equals the oracle's stage products summed in the same stage order
(C = P_s + C, the add of sp_mma_hybrid), with empty products skipped."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT
from trident_amd import gen

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "test_stage_loop")


def read_ranks(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            hdr = f.read(16)
            if not hdr:
                break
            r, c, m, nnz = np.frombuffer(hdr, np.int32)
            ptr = np.frombuffer(f.read(4 * (m + 1)), np.int32)
            idx = np.frombuffer(f.read(4 * nnz), np.int32)
            val = np.frombuffer(f.read(8 * nnz), np.float64)
            out[(int(r), int(c))] = (ptr, idx, val)
    return out


def expected_rank(A, P, r, c):
    """Oracle: stage products of rank (r, c) in Cannon order, summed as the
    loop does (C_accum = C_prod + C_local)."""
    m = A.nrows // P
    acc = None
    for it in range(P):
        k = (r + c + it) % P
        a = gen.submatrix(A, r * m, (r + 1) * m, k * m, (k + 1) * m)
        b = gen.submatrix(A, k * m, (k + 1) * m, c * m, (c + 1) * m)
        if a.nnz == 0 or b.nnz == 0:
            continue
        pp, pi, pv = oracle.spgemm(a, b)
        if pi.size == 0:
            continue
        prod = gen.from_arrays(m, m, pp, pi, pv)
        if acc is None:
            acc = prod
        else:
            sp, si, sv = oracle.spadd(prod, acc)
            acc = gen.from_arrays(m, m, sp, si, sv)
    if acc is None:
        return np.zeros(m + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float64)
    return acc.ptr.astype(np.int64), acc.idx, acc.val


@pytest.mark.parametrize("mode", [0, 1, 2], ids=["sp_mma_hybrid", "accum_thread_csx", "accum_thread_localmatrix"])
def test_stage_loop_matches_oracle(tmp_path, mode):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built; run __graft_entry__.build()")
    scale, ef, seed, P = 12, 8, 5, 4
    out = str(tmp_path / f"ranks_{mode}.bin")
    print(f"stage loop: R-MAT s{scale} ef{ef} seed {seed}, {P}x{P} grid, mode {mode}", flush=True)
    r = subprocess.run([BIN, str(scale), str(ef), str(seed), str(P), str(mode), out], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_ranks(out)
    assert len(got) == P * P
    A = gen.rmat(scale, ef, seed=seed)
    total = 0
    for rr in range(P):
        for cc in range(P):
            ep, ei, ev = expected_rank(A, P, rr, cc)
            gp, gi, gv = got[(rr, cc)]
            assert np.array_equal(gp.astype(np.int64), ep), f"rank {(rr, cc)}: row pointers differ"
            assert np.array_equal(gi, ei), f"rank {(rr, cc)}: column indices differ"
            assert np.array_equal(gv.view(np.uint64), ev.view(np.uint64)), f"rank {(rr, cc)}: values differ"
            total += gi.size
    assert total > 0
