"""Device memory contracts of the C-ABI: tr_symbolic's byte count bounds
what tr_spgemm takes from a fresh context's pool, and an injected
allocation limit (tr_pool_set_limit) makes tr_spgemm, tr_sp_mma and
tr_batch_create fail with TR_ERR_OOM, after which the same context still
computes correct products."""
import numpy as np
import pytest

import oracle
from trident_amd import gen
from trident_amd._lib import TR_ERR_OOM, TridentError
from trident_amd.device import Batch, Context, DeviceCsx, sp_mma, spgemm, symbolic

pytestmark = pytest.mark.gpu


def _same_as_oracle(C, A, B):
    cp, ci, cv = oracle.spgemm(A, B)
    H = C.download()
    return (np.array_equal(H.ptr.astype(np.int64), cp) and np.array_equal(H.idx, ci)
            and np.array_equal(H.val.view(np.uint64), cv.view(np.uint64)))


@pytest.mark.parametrize("case", ["rmat16", "galerkin16"])
def test_symbolic_bound_covers_spgemm(case):
    if case == "rmat16":
        A = B = gen.rmat(16, 16, seed=1)
    else:
        A, B = gen.laplace27(16), gen.sa_prolongator(16)
    print(f"memory bound: {case}", flush=True)
    with Context(0) as c1, Context(0) as c2:
        nnz, bound = symbolic(c1.upload(A), c1.upload(B))
        dA, dB = c2.upload(A), c2.upload(B)
        base, _ = c2.pool_stats()  # the operands themselves
        C = spgemm(dA, dB)
        c2.synchronize()
        used, _ = c2.pool_stats()
        print(f"  nnz {nnz} bound {bound} used {used - base}", flush=True)
        assert C.nnz == nnz
        assert used - base <= bound, f"tr_spgemm took {used - base} bytes, bound {bound}"
        assert _same_as_oracle(C, A, B)


def test_pool_limit_makes_calls_fail_with_oom_and_context_recovers():
    A = gen.rmat(14, 16, seed=3)
    with Context(0) as ctx:
        dA = ctx.upload(A)
        base, _ = ctx.pool_stats()
        nnz, _ = symbolic(dA, dA)
        for limit in (base + 4096, base + (nnz * 12) // 2):
            print(f"pool limit {limit}", flush=True)
            ctx.set_pool_limit(limit)
            with pytest.raises(TridentError) as e:
                spgemm(dA, dA)
            assert e.value.code == TR_ERR_OOM
            C = DeviceCsx(ctx)
            with pytest.raises(TridentError) as e:
                sp_mma(dA, dA, C)
            assert e.value.code == TR_ERR_OOM
            assert not C.initialized
            with pytest.raises(TridentError) as e:
                Batch(ctx, [dA, dA], [dA, dA])
            assert e.value.code == TR_ERR_OOM
            ctx.set_pool_limit(0)
            # the context is usable again and results are unchanged
            C = spgemm(dA, dA)
            assert _same_as_oracle(C, A, A)
            C.free()
            b = Batch(ctx, [dA], [dA])
            assert _same_as_oracle(b.result(0), A, A)
            b.close()
