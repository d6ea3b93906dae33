#!/usr/bin/env python3
"""Benchmark of the local SpGEMM path on one MI355X.

Default workload (one step = one plain C = A*A through tr_spgemm_panels):
  A = R-MAT scale 20 (2^20 rows, edge factor 16, a,b,c,d = .57,.19,.19,.05,
  duplicates merged, seed 1), int32 CSR, fp64 values U[-1,1).
  C (~9.7e9 entries, ~116 GB) does not fit one int32 tile, so the product is
  returned as int32 row panels; all panels stay resident in HBM.

Prints ONE JSON line: GFLOP/s (2 * multiplications / step time), the
dominant kernel's roofline (algorithmic HBM bytes / its HIP-event time vs
8 TB/s), and a CPU baseline (the repository's FP64 oracle on a bounded row
sample, which doubles as a bit-exactness check of those rows).

Other workloads: --workload er1024 | rmat:<scale> (parity-sized cases).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpGEMM GFLOP/s + % HBM roofline; batched/served p50/p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
LIGHT_MAX_FLOPS = 1024  # row class cut used by the library (flops <= cut: light kernels)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name):
    from trident_amd import gen
    if name == "er1024":
        A = gen.erdos_renyi(1024, 8, seed=0)
        desc = "C = A*A, A = Erdos-Renyi n=1024, 8 uniform columns/row (seed 0), int32 CSR, fp64 U[-1,1)"
    elif name.startswith("rmat:"):
        s = int(name.split(":")[1])
        A = gen.rmat(s, 16, seed=1)
        desc = (f"C = A*A, A = R-MAT scale {s} (edge factor 16, a,b,c,d=.57,.19,.19,.05, dups merged, seed 1), "
                "int32 CSR, fp64 U[-1,1)")
    else:
        raise SystemExit(f"unknown workload {name}")
    return A, desc


def row_flops(A):
    rl = np.diff(A.ptr).astype(np.int64)
    f = np.zeros(A.nrows, np.int64)
    np.add.at(f, A.row_ids(), rl[A.idx])
    return f


def download_rows(ctx, panels, rb, hptrs, rows):
    """Copy selected rows of the panelled C to host: list of (idx, val)."""
    out = []
    for r in rows:
        p = int(np.searchsorted(rb, r, side="right") - 1)
        P = panels[p]
        lr = r - rb[p]
        s, e = int(hptrs[p][lr]), int(hptrs[p][lr + 1])
        n = e - s
        idx = np.empty(n, np.int32)
        val = np.empty(n, np.float64)
        if n:
            ctx.check(ctx.lib.tr_memcpy_d2h(ctx.h, idx.ctypes.data_as(ctypes.c_void_p),
                                            ctypes.c_void_p(P.d.idx_vec + 4 * s), 4 * n))
            ctx.check(ctx.lib.tr_memcpy_d2h(ctx.h, val.ctypes.data_as(ctypes.c_void_p),
                                            ctypes.c_void_p(P.d.val + 8 * s), 8 * n))
        out.append((idx, val))
    return out


def cpu_baseline_and_check(A, f, ctx, panels, rb, hptrs, target_mults, seed=7):
    """Time the oracle on a random row sample of ~target_mults multiplications
    and compare those rows of C with the GPU result."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from trident_amd import gen
    rng = np.random.default_rng(seed)
    order = rng.permutation(A.nrows)
    cum = np.cumsum(f[order])
    k = int(np.searchsorted(cum, target_mults)) + 1
    rows = np.sort(order[:k])
    sub_ptr = np.zeros(len(rows) + 1, np.int64)
    lens = np.diff(A.ptr.astype(np.int64))[rows]
    np.cumsum(lens, out=sub_ptr[1:])
    sel = np.concatenate([np.arange(A.ptr[r], A.ptr[r + 1]) for r in rows]) if len(rows) else np.zeros(0, np.int64)
    S = gen.from_arrays(len(rows), A.ncols, sub_ptr, A.idx[sel], A.val[sel])
    mults = int(f[rows].sum())
    t0 = time.perf_counter()
    cp, ci, cv = oracle.spgemm(S, A)
    t_cpu = time.perf_counter() - t0
    got = download_rows(ctx, panels, rb, hptrs, rows)
    same_struct = True
    max_rel = 0.0
    for t, (gi, gv) in enumerate(got):
        s, e = int(cp[t]), int(cp[t + 1])
        if gi.shape[0] != e - s or not np.array_equal(gi, ci[s:e]):
            same_struct = False
            continue
        ref = cv[s:e]
        if not np.array_equal(gv, ref):
            rel = np.abs(gv - ref) / np.maximum(np.abs(ref), np.finfo(np.float64).tiny)
            max_rel = max(max_rel, float(rel.max()))
    return {
        "value": round(2.0 * mults / t_cpu / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": oracle.num_threads(),
        "kind": "port",
        "sample": (f"{len(rows)} random rows of A (seed {seed}) times A: {mults:.3e} multiplications, "
                   f"{int(cp[-1]):.3e} outputs, {t_cpu:.2f} s; the repository's FP64 OpenMP Gustavson oracle "
                   "(the Kokkos OpenMP KokkosWrap path cannot be built here)"),
    }, {"rows_checked": int(len(rows)), "structure_equal": same_struct, "max_rel_err": max_rel}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat:20")
    ap.add_argument("--cpu-mults", type=float, default=1.5e9,
                    help="multiplications in the CPU-baseline row sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU's share of the path; --gpus {args.gpus} is not supported "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    from trident_amd.device import Context, spgemm_panels

    t = time.perf_counter()
    A, desc = build_workload(args.workload)
    f = row_flops(A)
    mults = int(f.sum())
    log(f"[bench] workload ready in {time.perf_counter() - t:.1f}s: n={A.nrows} nnz(A)={A.nnz} mults={mults:.4e}")

    ctx = Context(0)
    dA = ctx.upload(A)

    def step():
        panels, rb = spgemm_panels(dA, dA, max_panel_nnz=2**31 - 1, max_panels=1024)
        return panels, rb

    for w in range(args.warmup):
        t = time.perf_counter()
        panels, rb = step()
        ctx.synchronize()
        log(f"[bench] warmup {w}: {time.perf_counter() - t:.3f}s, {len(panels)} panels")
        if w + 1 < args.warmup or args.steps > 0:
            for P in panels:
                P.free()

    ctx.reset_timing()
    ctx.set_timing(True)
    ctx.synchronize()
    times = []
    t_all = time.perf_counter()
    for s in range(args.steps):
        t = time.perf_counter()
        panels, rb = step()
        ctx.synchronize()
        times.append(time.perf_counter() - t)
        log(f"[bench] step {s}: {times[-1]:.3f}s")
        if s + 1 < args.steps:
            for P in panels:
                P.free()
    ctx.synchronize()
    total = time.perf_counter() - t_all
    ctx.set_timing(False)
    kt = ctx.timings()
    K = max(args.steps, 1)
    ms_per_step = 1000.0 * total / K
    gflops = 2.0 * mults / (total / K) / 1e9

    # C structure (panel row pointers) for accounting and the sample check
    hptrs = []
    for P in panels:
        pp = np.empty(P.d.nrows + 1, np.int32)
        ctx.check(ctx.lib.tr_memcpy_d2h(ctx.h, pp.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(P.d.ptr_vec),
                                        4 * (P.d.nrows + 1)))
        hptrs.append(pp)
    rownnz = np.concatenate([np.diff(p.astype(np.int64)) for p in hptrs])
    nnzC = int(rownnz.sum())

    # dominant kernel and its algorithmic bytes per launch
    dom = max(kt.items(), key=lambda kv: kv[1][1]) if kt else ("none", (1, 0.0))
    dname, (dlaunch, dms) = dom
    heavy = f > LIGHT_MAX_FLOPS
    light = (f > 0) & ~heavy
    rl = np.diff(A.ptr.astype(np.int64))
    B_bytes = 12 * A.nnz + 4 * (A.nrows + 1)
    if dname == "num_heavy":
        alg = 12 * int(rownnz[heavy].sum()) + 12 * int(rl[heavy].sum()) + B_bytes
        alg_desc = "12 B per C entry of heavy rows written + 12 B per A entry of those rows + B read once"
    elif dname.startswith("num_light"):
        alg = 12 * int(rownnz[light].sum()) + 12 * int(rl[light].sum()) + B_bytes
        alg_desc = "12 B per C entry of light rows written + 12 B per A entry of those rows + B read once"
    elif dname == "sym_heavy":
        alg = 8 * int(rl[heavy].sum()) + 4 * A.nnz + 4 * (A.nrows + 1) + 8 * int(heavy.sum())
        alg_desc = "A entries of heavy rows + B indices read once + per-row counts written"
    else:
        alg = 12 * nnzC + 2 * B_bytes
        alg_desc = "whole product: A and B read once, C written once"
    per_launch_s = dms / 1000.0 / max(dlaunch, 1)
    achieved = alg / per_launch_s / 1e9 if per_launch_s > 0 else 0.0
    bytes_alg_total = 12 * nnzC + 4 * (A.nrows + 1) + 2 * B_bytes
    roofline = {
        "bound": "hbm",
        "kernel": dname,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": None,
        "alg_bytes_per_launch": int(alg),
        "alg_bytes_def": alg_desc,
        "kernel_ms_per_launch": round(1000 * per_launch_s, 3),
        "step_alg_GBs": round(bytes_alg_total / (total / K) / 1e9, 2),
        "step_frac": round(bytes_alg_total / (total / K) / 1e9 / HBM_PEAK_GBS, 5),
    }
    kernels = {k: {"launches": v[0] // K, "ms_per_step": round(v[1] / K, 3)} for k, v in sorted(kt.items())}

    cpu = None
    check = None
    if not args.no_cpu:
        log("[bench] cpu baseline + sample check ...")
        t = time.perf_counter()
        cpu, check = cpu_baseline_and_check(A, f, ctx, panels, rb, hptrs, args.cpu_mults)
        log(f"[bench] cpu baseline done in {time.perf_counter() - t:.1f}s: {cpu['value']} GFLOP/s, check {check}")

    out = {
        "metric": METRIC,
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "form": "plain drop-in (tr_spgemm_panels)",
            "nrows": A.nrows,
            "nnz_A": A.nnz,
            "mults": mults,
            "nnz_C": nnzC,
            "panels": len(panels),
            "step_times_s": [round(x, 4) for x in times],
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_sample": check,
        "kernels": kernels,
    }
    for P in panels:
        P.free()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
