"""GPU tests of the serving loop (tr_server_*): mixed SpGEMM and Galerkin
requests, continuously batched, give the same results as each product run
alone (bit-exact), and the CPU oracle agrees; admission rejects a request
whose symbolic memory bound exceeds the budget; a Galerkin request whose
two steps fit the budget only one at a time completes; destroying a server
with queued requests returns."""
import threading

import numpy as np
import pytest

import oracle
from trident_amd import gen
from trident_amd._lib import TR_ERR_OOM, TridentError
from trident_amd.device import (DeviceCsx, Server, TR_REQ_GALERKIN, TR_REQ_SPGEMM, spgemm, symbolic)

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


def _galerkin_bounds(ctx, A, P):
    """Admission bounds of the two steps of P^T (A P): (b0, b1)."""
    _, b0 = symbolic(A, P)
    AP = spgemm(A, P)
    _, b1 = symbolic(P.transpose(), AP)
    return b0, b1


def _wait_in_thread(srv, rid, timeout=120.0):
    """srv.wait(rid) on a helper thread: (result, error) or a failure if the
    call does not return within `timeout` seconds."""
    out = {}

    def run():
        try:
            out["r"] = srv.wait(rid)
        except TridentError as e:  # noqa: PERF203
            out["e"] = e

    th = threading.Thread(target=run, daemon=True)
    th.start()
    th.join(timeout)
    assert not th.is_alive(), f"request {rid} did not finish within {timeout} s (server stalled)"
    return out.get("r"), out.get("e")


@pytest.mark.parametrize("n", [6, 10])
def test_galerkin_completes_when_steps_fit_only_one_at_a_time(ctx, n):
    """A budget that holds either step's bound but not both: the second step
    must still be admitted once the first has finished (its AP then counts
    only its real size), so the request completes (or is rejected), never
    waits forever."""
    A, P = gen.laplace27(n), gen.sa_prolongator(n)
    dA, dP = ctx.upload(A), ctx.upload(P)
    b0, b1 = _galerkin_bounds(ctx, dA, dP)
    budget = max(b0, b1) + min(b0, b1) // 2
    print(f"galerkin {n}^3: bounds {b0} + {b1}, budget {budget}", flush=True)
    assert max(b0, b1) <= budget < b0 + b1
    srv = Server(mem_budget=budget, workers=1, max_batch=4)
    oa, op = srv.register(dA), srv.register(dP)
    rids = [srv.submit(TR_REQ_GALERKIN, oa, op) for _ in range(3)]
    expect = spgemm(dP.transpose(), spgemm(dA, dP)).download()
    for rid in rids:
        res, err = _wait_in_thread(srv, rid)
        if err is not None:
            assert err.code == TR_ERR_OOM
            continue
        d, _ = res
        assert same(DeviceCsx(ctx, d).download(), expect)
        srv.release(rid)
    srv.close()


def test_destroy_with_queued_requests_returns(ctx):
    """Destroying a server with requests still queued fails them and returns
    (the workers do not wait for the queue to drain)."""
    dA = ctx.upload(gen.rmat(12, 16, seed=4))
    srv = Server(mem_budget=8 << 30, workers=1, max_batch=2)
    op = srv.register(dA)
    for _ in range(64):
        srv.submit(TR_REQ_SPGEMM, op, op, discard=True)
    th = threading.Thread(target=srv.close, daemon=True)
    th.start()
    th.join(120)
    assert not th.is_alive(), "tr_server_destroy did not return"
