"""GPU parity of the batched stage executor (tr_batch_*): one SUMMA stage of
a 4x4 tile grid, C_ij <- C_ij + A_ik * A_kj for all 16 (i, j), as one
hipGraph launch, against the CPU oracle — bit-exact, with and without
bitmask compression, across repeated graph replays; a replay with a changed
sparsity pattern is caught by the device-side guard."""
import numpy as np
import pytest

import oracle
from trident_amd import gen
from trident_amd._lib import TR_ERR_INVALID, TR_ERR_OVERFLOW, TridentError
from trident_amd.device import Batch, sp_mma, DeviceCsx

pytestmark = pytest.mark.gpu


def tiles_of(A, g):
    n = A.nrows // g
    return {(i, k): gen.submatrix(A, i * n, (i + 1) * n, k * n, (k + 1) * n) for i in range(g) for k in range(g)}


def ref_add(P, C):
    if C is None:
        return P
    cp, ci, cv = oracle.spadd(P, C)
    return gen.from_arrays(P.nrows, P.ncols, cp, ci, cv)


def prod(A, B):
    cp, ci, cv = oracle.spgemm(A, B)
    return gen.from_arrays(A.nrows, B.ncols, cp, ci, cv)


def same(H, R):
    return (np.array_equal(H.ptr.astype(np.int64), R.ptr.astype(np.int64)) and np.array_equal(H.idx, R.idx)
            and np.array_equal(H.val, R.val))


@pytest.mark.parametrize("compress", [False, True])
def test_stage_batch_matches_oracle(ctx, compress):
    A = gen.rmat(12, 16, seed=2)
    g = 4
    T = tiles_of(A, g)
    dT = {key: ctx.upload(m) for key, m in T.items()}
    k_prev, k = 3, 1
    # partial C from a previous stage (k_prev), some tiles left empty
    Cin_h, Cin_d = {}, {}
    for i in range(g):
        for j in range(g):
            if (i + j) % 3 == 0:
                Cin_h[(i, j)], Cin_d[(i, j)] = None, None
            else:
                P = prod(T[(i, k_prev)], T[(k_prev, j)])
                Cin_h[(i, j)] = P
                Cin_d[(i, j)] = ctx.upload(P)
    keys = [(i, j) for i in range(g) for j in range(g)]
    b = Batch(ctx, [dT[(i, k)] for i, j in keys], [dT[(k, j)] for i, j in keys], [Cin_d[key] for key in keys],
              compress=compress)
    refs = {(i, j): ref_add(prod(T[(i, k)], T[(k, j)]), Cin_h[(i, j)]) for i, j in keys}
    for rep in range(3):
        if rep:
            b.run()
            # unrelated eager work between replays
            x = ctx.upload(T[(0, 0)])
            x.free()
        ctx.synchronize()
        for t, key in enumerate(keys):
            assert same(b.result(t).download(), refs[key]), f"replay {rep} task {key}"
    flops, nnz_p, nnz_o = b.stats()
    assert flops == sum(oracle.spgemm_flops(T[(i, k)], T[(k, j)]) for i, j in keys)
    assert nnz_o == sum(r.nnz for r in refs.values())
    b.close()


def test_batch_equals_sequential_sp_mma(ctx):
    # the batched form and 16 plain sp_mma calls give identical tiles
    A = gen.rmat(11, 16, seed=5)
    T = tiles_of(A, 4)
    dT = {key: ctx.upload(m) for key, m in T.items()}
    keys = [(i, j) for i in range(4) for j in range(4)]
    Cs = []
    for i, j in keys:
        C = DeviceCsx(ctx)
        sp_mma(dT[(i, 2)], dT[(2, j)], C)
        sp_mma(dT[(i, 3)], dT[(3, j)], C)
        Cs.append(C)
    first = [DeviceCsx(ctx) for _ in keys]
    for t, (i, j) in enumerate(keys):
        sp_mma(dT[(i, 2)], dT[(2, j)], first[t])
    b = Batch(ctx, [dT[(i, 3)] for i, j in keys], [dT[(3, j)] for i, j in keys], first, compress=True)
    b.run()
    ctx.synchronize()
    for t in range(len(keys)):
        H, R = b.result(t).download(), Cs[t].download()
        assert same(H, R)
    b.close()


@pytest.mark.parametrize("compress", [False, True])
def test_eager_batch_replays_match_graph(ctx, compress):
    # a batch created without a graph re-runs the same launch sequence
    # eagerly; its results equal the captured batch's on every run
    A = gen.rmat(11, 16, seed=9)
    T = tiles_of(A, 4)
    dT = {key: ctx.upload(m) for key, m in T.items()}
    keys = [(i, j) for i in range(4) for j in range(4)]
    As, Bs = [dT[(i, 1)] for i, j in keys], [dT[(1, j)] for i, j in keys]
    bg = Batch(ctx, As, Bs, None, compress=compress)
    be = Batch(ctx, As, Bs, None, compress=compress, graph=False)
    for rep in range(2):
        bg.run()
        be.run()
        ctx.synchronize()
        for t in range(len(keys)):
            assert same(be.result(t).download(), bg.result(t).download()), f"run {rep} task {keys[t]}"
    bg.close()
    be.close()


def _narrow_and_spread(n, per_row, seed):
    """Two CSR matrices with the same row pointers (so the same nnz): one whose
    rows all use columns 0..per_row-1, one whose rows use random columns."""
    rng = np.random.default_rng(seed)
    ptr = np.arange(0, (n + 1) * per_row, per_row, dtype=np.int64)
    narrow = np.tile(np.arange(per_row, dtype=np.int32), n)
    spread = np.concatenate([np.sort(rng.choice(n, per_row, replace=False)) for _ in range(n)]).astype(np.int32)
    val = rng.uniform(-1, 1, n * per_row)
    return gen.from_arrays(n, n, ptr, narrow, val), gen.from_arrays(n, n, ptr, spread, val)


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("accum", [False, True])
def test_replay_guard_rejects_changed_pattern(ctx, graph, accum):
    """A replay reuses the plan's sizes, so only values may change between
    create and run.  Changed values: the replay is checked and exact.  A
    changed pattern (same nnz, columns spread out so the product grows past
    the planned arena): the run writes nothing out of bounds and
    tr_batch_check reports it; restoring the pattern makes runs valid again."""
    n = 512
    A = gen.erdos_renyi(n, 8, seed=5)
    Bn, Bs = _narrow_and_spread(n, 4, seed=6)
    C0 = prod(A, gen.erdos_renyi(n, 4, seed=7)) if accum else None
    print(f"replay guard: graph={graph} accum={accum}", flush=True)
    dA, dB = ctx.upload(A), ctx.upload(Bn)
    dC = [ctx.upload(C0)] if accum else None
    b = Batch(ctx, [dA], [dB], dC, graph=graph)
    b.check()
    assert same(b.result(0).download(), ref_add(prod(A, Bn), C0))
    # values only: B's values change, its pattern does not
    Bv = gen.from_arrays(n, n, Bn.ptr, Bn.idx, -2.0 * Bn.val)
    ctx.check(ctx.lib.tr_memcpy_h2d(ctx.h, dB.d.val, Bv.val.ctypes.data, Bv.val.nbytes))
    b.run()
    b.check()
    assert same(b.result(0).download(), ref_add(prod(A, Bv), C0))
    # pattern change: the planned product arena is too small for it
    ctx.check(ctx.lib.tr_memcpy_h2d(ctx.h, dB.d.idx_vec, Bs.idx.ctypes.data, Bs.idx.nbytes))
    b.run()
    with pytest.raises(TridentError) as e:
        b.check()
    assert e.value.code == TR_ERR_INVALID
    # back to the planned pattern: valid again
    ctx.check(ctx.lib.tr_memcpy_h2d(ctx.h, dB.d.idx_vec, Bn.idx.ctypes.data, Bn.idx.nbytes))
    b.run()
    b.check()
    assert same(b.result(0).download(), ref_add(prod(A, Bv), C0))
    b.close()


@pytest.mark.parametrize("accum", [False, True])
def test_result_panels_match_oracle(ctx, accum):
    """A batch result read as int32 row panels (the way a result larger than
    an int32 tile is read) under a forced small panel cap: the panels
    concatenate to the oracle's result and each holds at most the cap."""
    A = gen.rmat(11, 16, seed=4)
    T = tiles_of(A, 2)
    keys = [(0, 0), (1, 1), (0, 1)]
    dT = {key: ctx.upload(m) for key, m in T.items()}
    Cin_h = [prod(T[(i, 1)], T[(1, j)]) if accum else None for i, j in keys]
    Cin_d = [ctx.upload(c) for c in Cin_h] if accum else None
    print(f"result panels: accum={accum}", flush=True)
    b = Batch(ctx, [dT[(i, 0)] for i, j in keys], [dT[(0, j)] for i, j in keys], Cin_d)
    for t, (i, j) in enumerate(keys):
        R = ref_add(prod(T[(i, 0)], T[(0, j)]), Cin_h[t])
        lens = np.diff(R.ptr.astype(np.int64))
        cap = max(int(lens.max()), R.nnz // 5)
        panels, rb = b.result_panels(t, max_panel_nnz=cap)
        assert rb[0] == 0 and rb[-1] == R.nrows and len(panels) >= 2
        ptr, idx, val = [np.zeros(1, np.int64)], [], []
        for P in panels:
            H = P.download()
            assert H.nnz <= cap
            ptr.append(ptr[-1][-1] + H.ptr[1:].astype(np.int64))
            idx.append(H.idx)
            val.append(H.val)
        assert np.array_equal(np.concatenate(ptr), R.ptr.astype(np.int64))
        assert np.array_equal(np.concatenate(idx), R.idx)
        assert np.array_equal(np.concatenate(val).view(np.uint64), R.val.view(np.uint64))
        if lens.max() > 1:
            with pytest.raises(TridentError) as e:
                b.result_panels(t, max_panel_nnz=int(lens.max()) - 1)
            assert e.value.code == TR_ERR_OVERFLOW
    b.close()
