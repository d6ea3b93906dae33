"""GPU tests of the serving loop (tr_server_*): mixed SpGEMM and Galerkin
requests, continuously batched, give the same results as each product run
alone (bit-exact), and the CPU oracle agrees; admission rejects a request
whose symbolic memory bound exceeds the budget."""
import numpy as np
import pytest

import oracle
from trident_amd import gen
from trident_amd._lib import TR_ERR_OOM, TridentError
from trident_amd.device import (DeviceCsx, Server, TR_REQ_GALERKIN, TR_REQ_SPGEMM, spgemm)

pytestmark = pytest.mark.gpu


def same(X, Y):
    return (np.array_equal(X.ptr, Y.ptr) and np.array_equal(X.idx, Y.idx) and np.array_equal(X.val, Y.val))


def test_mixed_requests_same_alone_and_in_mix(ctx):
    rng = np.random.default_rng(3)
    mats = {f"rmat{s}_{ef}": gen.rmat(s, ef, seed=s * 10 + ef) for s in (9, 10, 11) for ef in (8, 16)}
    lap = {n: (gen.laplace27(n), gen.sa_prolongator(n)) for n in (6, 9, 12)}
    dev = {k: ctx.upload(v) for k, v in mats.items()}
    dlap = {n: (ctx.upload(A), ctx.upload(P)) for n, (A, P) in lap.items()}
    srv = Server(mem_budget=8 << 30, workers=2, max_batch=8)
    ops = {k: srv.register(v) for k, v in dev.items()}
    lops = {n: (srv.register(A), srv.register(P)) for n, (A, P) in dlap.items()}
    reqs = []
    for _ in range(40):
        if rng.random() < 0.6:
            k = list(mats)[rng.integers(len(mats))]
            reqs.append(("spgemm", k, srv.submit(TR_REQ_SPGEMM, ops[k], ops[k])))
        else:
            n = list(lap)[rng.integers(len(lap))]
            reqs.append(("galerkin", n, srv.submit(TR_REQ_GALERKIN, lops[n][0], lops[n][1])))
    # alone: plain calls on the test context
    alone = {}
    for kind, key, rid in reqs:
        if (kind, key) in alone:
            continue
        if kind == "spgemm":
            alone[(kind, key)] = spgemm(dev[key], dev[key]).download()
        else:
            A, P = dlap[key]
            Pt = P.transpose()
            alone[(kind, key)] = spgemm(Pt, spgemm(A, P)).download()
    for kind, key, rid in reqs:
        d, lat = srv.wait(rid)
        got = DeviceCsx(ctx, d).download()
        assert same(got, alone[(kind, key)]), (kind, key)
        assert lat >= 0.0
        srv.release(rid)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] >= 1
    # the oracle agrees with the served results
    cp, ci, cv = oracle.spgemm(mats["rmat10_16"], mats["rmat10_16"])
    X = alone[("spgemm", "rmat10_16")] if ("spgemm", "rmat10_16") in alone else spgemm(
        dev["rmat10_16"], dev["rmat10_16"]).download()
    assert np.array_equal(X.ptr.astype(np.int64), cp) and np.array_equal(X.idx, ci) and np.array_equal(X.val, cv)
    A, P = lap[9]
    AP = oracle.spgemm(A, P)
    APm = gen.from_arrays(A.nrows, P.ncols, *AP)
    Pt = gen.transpose(P)
    cp, ci, cv = oracle.spgemm(Pt, APm)
    Y = alone[("galerkin", 9)] if ("galerkin", 9) in alone else None
    if Y is not None:
        assert np.array_equal(Y.ptr.astype(np.int64), cp) and np.array_equal(Y.idx, ci) and np.array_equal(Y.val, cv)
    srv.close()


def test_admission_rejects_oversized_request(ctx):
    A = ctx.upload(gen.rmat(10, 16, seed=1))
    srv = Server(mem_budget=4096, workers=1, max_batch=4)
    op = srv.register(A)
    rid = srv.submit(TR_REQ_SPGEMM, op, op)
    with pytest.raises(TridentError) as e:
        srv.wait(rid)
    assert e.value.code == TR_ERR_OOM
    srv.close()
