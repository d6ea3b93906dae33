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
sample, which doubles as a bit-exactness check of those rows, plus the
Erdos-Renyi n=1024 product).  The same line carries the other forms, each
plain vs batched/served with its dominant kernel's roofline and an oracle
check on sampled rows:
  batched   one SUMMA stage of R-MAT scale 22 cut 4x4 (stage 3), 16 tile
            products added into partial-C tiles, one hipGraph launch;
  batched_first_stage  stage 2 of the same grid for ranks that start there
            (no partial C; tile products beyond int32 read as row panels);
  galerkin  P^T A P on the 27-point Laplacian 128^3 with its SA prolongator;
  served    2000 mixed requests at 1x and 4x the plain rate, p50/p99.

Other workloads: --workload er1024 | rmat:<scale> (parity-sized cases).
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpGEMM GFLOP/s + % HBM roofline; batched/served p50/p99 latency vs plain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name):
    from trident_amd import workloads
    if name == "er1024":
        desc = "C = A*A, A = Erdos-Renyi n=1024, 8 uniform columns/row (seed 0), int32 CSR, fp64 U[-1,1)"
    elif name.startswith("rmat:"):
        s = int(name.split(":")[1])
        desc = (f"C = A*A, A = R-MAT scale {s} (edge factor 16, a,b,c,d=.57,.19,.19,.05, dups merged, seed 1), "
                "int32 CSR, fp64 U[-1,1)")
    else:
        raise SystemExit(f"unknown workload {name}")
    return workloads.product_operand(name), desc


def apply_tuning(args):
    """Tuning of the multiply for this run: the autotuner's JSON for the
    workload when present (or --tuning PATH), the built-in values otherwise.
    Set as the process default so every context (server workers included)
    uses it.  Returns (tuning dict, source)."""
    from trident_amd import autotune
    from trident_amd.device import load_tuning, set_default_tuning, tuning_default
    if args.tuning == "none":
        return tuning_default(), "built-in"
    path = autotune.default_path(args.workload) if args.tuning == "auto" else args.tuning
    if not os.path.exists(path):
        return tuning_default(), "built-in"
    t = dict(tuning_default(), **load_tuning(path))
    set_default_tuning(t)
    return t, os.path.relpath(path, ROOT)


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 --pmc summary (profiles/pmc_traffic.txt): FETCH_SIZE and
    WRITE_SIZE in KiB per dispatch, each scaled by what the same counter read
    for a plain copy of known size.  None when no measurement is recorded."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if not line.startswith("traffic "):
                continue
            kv = dict(tok.split("=", 1) for tok in line.split()[1:])
            if kv.get("kernel") == kernel and kv.get("workload") == workload:
                b = 1024.0 * (float(kv["fetch_kib"]) * float(kv["fetch_scale"]) +
                              float(kv["write_kib"]) * float(kv["write_scale"]))
                return b, "profiles/pmc_traffic.txt"
    return None, None


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


def _cpu_topology():
    """(physical cores, logical CPUs) among the CPUs this process may run on,
    from the sysfs topology (unique package/core pairs)."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            with open(base + "physical_package_id") as f1, open(base + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            cores.add(("cpu", str(c)))
    return max(1, len(cores)), len(cpus)


def _cpu_model() -> str:
    """The host CPU's model name as lscpu reports it."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                return ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        names = [ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name")]
        return names[0] if names else "unknown"
    except OSError:
        return "unknown"


def cpu_env():
    """Environment of the CPU-baseline child process: OpenMP threads = the
    physical cores this process may use, bound close (OMP_PROC_BIND=close).
    Binding is confined to the child, so it never pins this
    process's own threads (the HIP runtime's, the server's workers)."""
    phys, logical = _cpu_topology()
    env = dict(os.environ, OMP_NUM_THREADS=str(phys), OMP_PROC_BIND="close")
    return env, phys, logical


def cpu_c1_baseline(oracle, repeats=10):
    """The reference's CPU-runnable case: C = A*A, A = Erdos-Renyi n=1024 with
    8 columns per row (seed 0), on the oracle: median wall time of `repeats`
    runs after 2 warm-ups (the multiply only)."""
    from trident_amd import workloads
    A = workloads.product_operand("er1024")
    mults = int(row_flops(A).sum())
    ts = []
    for r in range(repeats + 2):
        t = time.perf_counter()
        oracle.spgemm(A, A)
        if r >= 2:
            ts.append(time.perf_counter() - t)
    t = float(np.median(ts))
    return {"workload": "C = A*A, A = Erdos-Renyi n=1024, 8 columns/row (seed 0)", "mults": mults,
            "ms": round(1000 * t, 4), "GFLOP/s": round(2 * mults / t / 1e9, 3)}


def cpu_child():
    """CPU-baseline child (bench.py --cpu-child).  It is started before the
    parent touches the GPU (a process that has initialised the GPU does not
    start other programs) and waits for one line on stdin: PATH, whose
    PATH.npz holds the sampled rows of A, B = A and the GPU's rows of C.  It
    times the oracle on them and on the Erdos-Renyi n=1024 product, checks
    the GPU rows, and writes PATH.json; an empty line ends it."""
    path = sys.stdin.readline().strip()
    if not path:
        return 0
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from trident_amd import gen
    with np.load(path + ".npz") as npz:
        z = {k: npz[k] for k in npz.files}  # each item access of an NpzFile reads the array again
    S = gen.from_arrays(int(z["s_n"]), int(z["a_n"]), z["s_ptr"], z["s_idx"], z["s_val"])
    A = gen.from_arrays(int(z["a_n"]), int(z["a_n"]), z["a_ptr"], z["a_idx"], z["a_val"])
    t0 = time.perf_counter()
    cp, ci, cv = oracle.spgemm(S, A)
    t_cpu = time.perf_counter() - t0
    gp, gi, gv = z["g_ptr"], z["g_idx"], z["g_val"]
    got = [(gi[gp[i]:gp[i + 1]], gv[gp[i]:gp[i + 1]]) for i in range(len(gp) - 1)]
    same, max_rel = compare_rows(got, cp, ci, cv)
    out = {"t_cpu": t_cpu, "outputs": int(cp[-1]), "threads": oracle.num_threads(), "structure_equal": same,
           "max_rel_err": max_rel, "c1": cpu_c1_baseline(oracle)}
    with open(path + ".json", "w") as f:
        json.dump(out, f)
    return 0


def start_cpu_child():
    """The CPU-baseline child process (see cpu_child), with OpenMP threads =
    the physical cores, bound close; call before any GPU work."""
    env, phys, logical = cpu_env()
    proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-child"], env=env,
                            stdin=subprocess.PIPE, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    return proc, env, phys, logical


def cpu_baseline_and_check(child, A, f, ctx, panels, rb, hptrs, target_mults, seed=7):
    """Time the oracle on a random row sample of ~target_mults multiplications
    (in the child process, OpenMP bound to the physical cores) and compare
    those rows of C with the GPU result."""
    import tempfile
    proc, env, phys, logical = child
    rng = np.random.default_rng(seed)
    order = rng.permutation(A.nrows)
    cum = np.cumsum(f[order])
    k = int(np.searchsorted(cum, target_mults)) + 1
    rows = np.sort(order[:k])
    S = take_rows(A, rows)
    mults = int(f[rows].sum())
    got = download_rows(ctx, panels, rb, hptrs, rows)
    g_ptr = np.zeros(len(got) + 1, np.int64)
    np.cumsum([g[0].shape[0] for g in got], out=g_ptr[1:])
    with tempfile.TemporaryDirectory() as td:
        base = os.path.join(td, "cpu")
        np.savez(base + ".npz", s_n=S.nrows, s_ptr=S.ptr, s_idx=S.idx, s_val=S.val, a_n=A.nrows, a_ptr=A.ptr,
                 a_idx=A.idx, a_val=A.val, g_ptr=g_ptr,
                 g_idx=np.concatenate([g[0] for g in got]) if got else np.zeros(0, np.int32),
                 g_val=np.concatenate([g[1] for g in got]) if got else np.zeros(0))
        _, err = proc.communicate(base + "\n", timeout=600)
        if proc.returncode != 0 or not os.path.exists(base + ".json"):
            raise RuntimeError(f"CPU baseline child failed (exit {proc.returncode}): {err[-2000:]}")
        with open(base + ".json") as fh:
            res = json.load(fh)
    t_cpu = res["t_cpu"]
    return {
        "value": round(2.0 * mults / t_cpu / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": res["threads"],
        "cpu": f"{_cpu_model()} ({phys} physical cores / {logical} logical CPUs available to the process; "
               f"OMP_NUM_THREADS={env['OMP_NUM_THREADS']}, OMP_PROC_BIND={env['OMP_PROC_BIND']})",
        "kind": "port",
        "sample": (f"{len(rows)} random rows of A (seed {seed}) times A: {mults:.3e} multiplications, "
                   f"{res['outputs']:.3e} outputs, {t_cpu:.2f} s; the repository's FP64 OpenMP Gustavson oracle, a "
                   "substitute for the Kokkos OpenMP KokkosWrap path, which cannot be built here"),
        "c1": res["c1"],
    }, {"rows_checked": int(len(rows)), "structure_equal": res["structure_equal"], "max_rel_err": res["max_rel_err"]}


def take_rows(A, rows):
    """Host CSR of the given rows of A (all columns)."""
    from trident_amd import gen
    ptr = A.ptr.astype(np.int64)
    lens = np.diff(ptr)[rows]
    sub_ptr = np.zeros(len(rows) + 1, np.int64)
    np.cumsum(lens, out=sub_ptr[1:])
    sel = np.concatenate([np.arange(ptr[r], ptr[r + 1]) for r in rows]) if len(rows) else np.zeros(0, np.int64)
    return gen.from_arrays(len(rows), A.ncols, sub_ptr, A.idx[sel], A.val[sel])


def scatter_rows(n, ncols, rows, cp, ci, cv):
    """Host CSR with n rows whose given rows hold (cp, ci, cv) and the rest
    are empty."""
    from trident_amd import gen
    lens = np.zeros(n, np.int64)
    lens[rows] = np.diff(cp)
    ptr = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=ptr[1:])
    return gen.from_arrays(n, ncols, ptr, ci, cv)


def compare_rows(got, cp, ci, cv):
    """got: list of (idx, val) per sampled row; (cp, ci, cv) the oracle's.
    Returns (structure equal, max relative value error)."""
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
    return same_struct, max_rel


def device_rows(ctx, M, rows):
    """Rows of a device CSR tile (int32 ptr) as a list of (idx, val)."""
    hp = ctx.to_host(M.d.ptr_vec, (M.d.nrows + 1,), np.int32).astype(np.int64)
    out = []
    for r in rows:
        s, e = int(hp[r]), int(hp[r + 1])
        out.append((ctx.to_host(M.d.idx_vec + 4 * s, (e - s,), np.int32),
                    ctx.to_host(M.d.val + 8 * s, (e - s,), np.float64)))
    return out


def kernel_roofline(kt, steps, alg_bytes):
    """Roofline entry of the dominant kernel in a per-kernel timing table
    (name -> (launches, total ms)); alg_bytes(name) gives its algorithmic
    bytes per launch and their definition, or None."""
    if not kt:
        return None
    name, (launches, ms) = max(kt.items(), key=lambda kv: kv[1][1])
    per_launch_s = ms / 1000.0 / max(launches, 1)
    ab = alg_bytes(name, launches)
    if ab is None:
        return {"kernel": name, "kernel_ms_per_launch": round(1000 * per_launch_s, 3), "launches": launches}
    alg, desc = ab
    achieved = alg / per_launch_s / 1e9 if per_launch_s > 0 else 0.0
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "alg_bytes_per_launch": int(alg), "alg_bytes_def": desc,
            "kernel_ms_per_launch": round(1000 * per_launch_s, 3), "launches": launches}


def use_tuning(workload, choice="auto"):
    """Make the tuning of `workload` (the autotuner's JSON) the process default
    for contexts created afterwards; returns (tuning, source)."""
    from trident_amd import autotune
    from trident_amd.device import load_tuning, set_default_tuning, tuning_default
    set_default_tuning(None)
    if choice == "none":
        return tuning_default(), "built-in"
    path = autotune.default_path(workload) if choice == "auto" else choice
    if not os.path.exists(path):
        return tuning_default(), "built-in"
    t = dict(tuning_default(), **load_tuning(path))
    set_default_tuning(t)
    return t, os.path.relpath(path, ROOT)


def class_bytes(alen, flops, out_nnz, b_bytes, cut):
    """Algorithmic bytes of the light (flops <= cut) and heavy rows of one
    product: 12 B per output entry written and per A entry read, plus B read
    once by each class that has rows."""
    light = (flops > 0) & (flops <= cut)
    heavy = flops > cut
    lb = 12 * int(out_nnz[light].sum()) + 12 * int(alen[light].sum()) + (b_bytes if light.any() else 0)
    hb = 12 * int(out_nnz[heavy].sum()) + 12 * int(alen[heavy].sum()) + (b_bytes if heavy.any() else 0)
    return lb, hb


def stage_section(scale, k, steps, tuning="auto", with_compress=False, g=4):
    """Batched stage.  R-MAT scale S (edge factor 16, seed 2) cut into a g x g
    tile grid; SUMMA stage k updates every partial-C tile,
    C_ij <- C_ij + A_ik * A_kj (16 uneven tile products).  The partial-C
    tiles entering the stage are the products of stage k-1, A_i(k-1) A_(k-1)j:
    the partial C of the ranks whose Cannon order reaches stage k right after
    stage k-1; a tile whose sum would pass INT32_MAX entries (an int32 tile
    cannot hold it) enters empty.  (The sum over all earlier stages does not
    fit one GPU at scale 22: stage 0 alone makes about 5e10 entries.)  Plain: 16 sp_mma calls;
    batched: one hipGraph launch.  Checked against the oracle on sampled rows
    of three tiles."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from trident_amd import gen
    from trident_amd.device import Batch, Context, DeviceCsx, sp_mma, spgemm, spgemm_flops, symbolic
    tun, tsrc = use_tuning(f"stage:{scale}:{k}", tuning)
    t0 = time.perf_counter()
    A = gen.rmat(scale, 16, seed=2)
    n = A.nrows // g
    need = {(i, kk) for i in range(g) for kk in ({k, k - 1} if k > 0 else {k})}
    need |= {(kk, j) for j in range(g) for kk in ({k, k - 1} if k > 0 else {k})}
    T = {key: gen.submatrix(A, key[0] * n, (key[0] + 1) * n, key[1] * n, (key[1] + 1) * n) for key in need}
    del A
    log(f"[stage] R-MAT s{scale} tiles ready in {time.perf_counter() - t0:.1f}s")
    ctx = Context(0)
    dT = {key: ctx.upload(m) for key, m in T.items()}
    keys = [(i, j) for i in range(g) for j in range(g)]
    # a partial-C tile is an int32 tile (the reference's size_type): a tile
    # whose previous-stage product plus this stage's would pass INT32_MAX
    # entries enters the stage empty
    Cin = []
    for i, j in keys:
        c = DeviceCsx(ctx)
        if k > 0:
            n_prev, _ = symbolic(dT[(i, k - 1)], dT[(k - 1, j)])
            n_cur, _ = symbolic(dT[(i, k)], dT[(k, j)])
            if n_prev + n_cur <= 2**31 - 1:
                c = spgemm(dT[(i, k - 1)], dT[(k - 1, j)])
        Cin.append(c)
    ctx.synchronize()
    mults = sum(spgemm_flops(dT[(i, k)], dT[(k, j)]) for i, j in keys)
    nnz_cin = sum(c.nnz for c in Cin)
    log(f"[stage] stage {k}: {mults:.4e} multiplications, partial C in {nnz_cin:.4e} entries")

    def plain_step():
        outs = []
        for t, (i, j) in enumerate(keys):
            C = DeviceCsx(ctx, type(Cin[t].d).from_buffer_copy(Cin[t].d))
            C.d.contig = 0  # borrowed view: sp_mma replaces it, the original stays
            C.d.buf = None
            sp_mma(dT[(i, k)], dT[(k, j)], C)
            outs.append(C)
        ctx.synchronize()
        return outs

    def timed(fn, reps):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    res = {}
    for C in plain_step():  # warm-up
        C.free()
    holder = []

    def plain_once():
        while holder:
            holder.pop().free()
        holder.extend(plain_step())
    t_plain = timed(plain_once, steps)
    nnz_out = sum(c.nnz for c in holder)
    while holder:
        holder.pop().free()
    res["plain"] = {"ms": round(1000 * t_plain, 2), "GFLOP/s": round(2 * mults / t_plain / 1e9, 2)}
    log(f"[stage] plain: {res['plain']}")
    modes = ["batched"] + (["batched+compress"] if with_compress else [])
    rng = np.random.default_rng(11)
    for mode in modes:
        b = Batch(ctx, [dT[(i, k)] for i, j in keys], [dT[(k, j)] for i, j in keys], Cin,
                  compress=(mode == "batched+compress"))

        def run_b():
            b.run()
            ctx.synchronize()
        run_b()
        t_b = timed(run_b, steps)
        b.check()
        res[mode] = {"ms": round(1000 * t_b, 2), "GFLOP/s": round(2 * mults / t_b / 1e9, 2)}
        log(f"[stage] {mode}: {res[mode]}")
        if mode == "batched":
            # oracle on sampled rows of three tiles: the largest product, the
            # smallest non-empty one and one more
            fl = [spgemm_flops(dT[(i, k)], dT[(k, j)]) for i, j in keys]
            order = [t for t in np.argsort(fl) if fl[t] > 0]
            picks = sorted({int(order[-1]), int(order[0]), int(order[len(order) // 2])})
            ok, max_rel, checked = True, 0.0, 0
            for t in picks:
                i, j = keys[t]
                Ak, Bk = T[(i, k)], T[(k, j)]
                rows = np.sort(rng.choice(Ak.nrows, 48, replace=False))
                P = oracle.spgemm(take_rows(Ak, rows), Bk)
                if Cin[t].initialized:
                    Q = oracle.spgemm(take_rows(T[(i, k - 1)], rows), T[(k - 1, j)])
                    Pm = gen.from_arrays(len(rows), Bk.ncols, *P)
                    Qm = gen.from_arrays(len(rows), Bk.ncols, *Q)
                    P = oracle.spadd(Pm, Qm)
                same, rel = compare_rows(device_rows(ctx, b.result(t), rows), *P)
                ok &= same
                max_rel = max(max_rel, rel)
                checked += len(rows)
            res["oracle_sample"] = {"tiles": [list(keys[t]) for t in picks], "rows_checked": checked,
                                    "structure_equal": bool(ok), "max_rel_err": max_rel}
            log(f"[stage] oracle sample: {res['oracle_sample']}")
            _, nnz_p, _ = b.stats()
        b.close()
    res["speedup_batched_vs_plain"] = round(res["plain"]["ms"] / res["batched"]["ms"], 3)
    # per-kernel HIP-event times of one eager pass of the batch
    b = Batch(ctx, [dT[(i, k)] for i, j in keys], [dT[(k, j)] for i, j in keys], Cin, compress=False, graph=False)
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    b.run()
    ctx.synchronize()
    ctx.set_timing(False)
    kt = ctx.timings()
    b.close()
    rows_total = g * n
    ptr_b = 4 * 3 * (rows_total + g * g)

    def alg(name, launches):
        if name == "add_write":
            return (12 * (nnz_p + nnz_cin + nnz_out) + ptr_b,
                    "12 B per product entry and per partial-C entry read, per output entry written, + row pointers")
        if name == "add_count":
            return (4 * (nnz_p + nnz_cin) + ptr_b, "4 B per product and partial-C column index read + row pointers")
        return None
    res["kernels_ms"] = {kk: round(v[1], 3) for kk, v in sorted(kt.items())}
    res["roofline"] = kernel_roofline(kt, 1, alg)
    tile_bytes = sum(12 * m.nnz + 4 * (m.nrows + 1) for key, m in T.items() if k in key)
    step_bytes = tile_bytes + 12 * (nnz_cin + nnz_out) + 8 * (rows_total + g * g)
    res["step_alg_GBs"] = round(step_bytes / (res["batched"]["ms"] / 1000) / 1e9, 2)
    res["step_frac"] = round(res["step_alg_GBs"] / HBM_PEAK_GBS, 5)
    res.update({"workload": f"R-MAT scale {scale} (ef 16, seed 2) cut {g}x{g}; SUMMA stage k={k}: {g * g} tile "
                            f"products added into partial-C tiles (the stage {k - 1} products)",
                "mults": mults, "nnz_partial_C_in": nnz_cin, "tiles_with_partial_C": sum(c.initialized for c in Cin),
                "nnz_products": nnz_p, "nnz_out": nnz_out,
                "tuning": tun, "tuning_source": tsrc})
    for c in Cin:
        if c.initialized:
            c.free()
    ctx.close()
    return res


def first_stage_section(scale, k, steps, tuning="auto", g=4):
    """Batched first stage.  SUMMA stage k of R-MAT scale S (edge factor 16,
    seed 2) cut g x g, for ranks that start at that stage (no partial C yet):
    the 16 tile products alone.  At scale 22, stages 1 and 2 hold tile
    products of up to 2.9e9 entries, beyond an int32 tile, so both forms
    return them as int32 row panels: plain, 16 tr_spgemm_panels calls;
    batched, one hipGraph launch read with tr_batch_result_panels.  Checked
    against the oracle on sampled rows of three tiles."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from trident_amd import gen
    from trident_amd.device import Batch, Context, spgemm_flops, spgemm_panels
    tun, tsrc = use_tuning(f"stage:{scale}:{k}", tuning)
    t0 = time.perf_counter()
    A = gen.rmat(scale, 16, seed=2)
    n = A.nrows // g
    need = {(i, k) for i in range(g)} | {(k, j) for j in range(g)}
    T = {key: gen.submatrix(A, key[0] * n, (key[0] + 1) * n, key[1] * n, (key[1] + 1) * n) for key in need}
    del A
    log(f"[first stage] R-MAT s{scale} tiles ready in {time.perf_counter() - t0:.1f}s")
    ctx = Context(0)
    dT = {key: ctx.upload(m) for key, m in T.items()}
    keys = [(i, j) for i in range(g) for j in range(g)]
    fl = [spgemm_flops(dT[(i, k)], dT[(k, j)]) for i, j in keys]
    mults = sum(fl)

    def plain_step():
        outs = []
        for i, j in keys:
            panels, _ = spgemm_panels(dT[(i, k)], dT[(k, j)], max_panels=64)
            outs.extend(panels)
        ctx.synchronize()
        return outs

    for P in plain_step():
        P.free()
    ts = []
    nnz_out = 0
    for _ in range(steps):
        t = time.perf_counter()
        outs = plain_step()
        ts.append(time.perf_counter() - t)
        nnz_out = sum(P.nnz for P in outs)
        for P in outs:
            P.free()
    t_plain = float(np.median(ts))
    res = {"plain": {"ms": round(1000 * t_plain, 2), "GFLOP/s": round(2 * mults / t_plain / 1e9, 2)}}
    log(f"[first stage] plain: {res['plain']}")
    b = Batch(ctx, [dT[(i, k)] for i, j in keys], [dT[(k, j)] for i, j in keys])
    ts = []
    for _ in range(steps):
        t = time.perf_counter()
        b.run()
        ctx.synchronize()
        ts.append(time.perf_counter() - t)
    b.check()
    t_b = float(np.median(ts))
    res["batched"] = {"ms": round(1000 * t_b, 2), "GFLOP/s": round(2 * mults / t_b / 1e9, 2)}
    res["speedup_batched_vs_plain"] = round(t_plain / t_b, 3)
    log(f"[first stage] batched: {res['batched']}")
    # oracle on sampled rows of three tiles (the largest product among them)
    rng = np.random.default_rng(12)
    order = [t for t in np.argsort(fl) if fl[t] > 0]
    picks = sorted({int(order[-1]), int(order[0]), int(order[len(order) // 2])})
    ok, max_rel, checked, big = True, 0.0, 0, 0
    per_row_out = {}
    for t in range(len(keys)):
        panels, rb = b.result_panels(t, max_panels=64)
        big = max(big, len(panels))
        hp = np.concatenate([[0]] + [np.diff(ctx.to_host(P.d.ptr_vec, (P.d.nrows + 1,), np.int32).astype(np.int64))
                                     for P in panels])
        per_row_out[t] = hp[1:]
        if t not in picks:
            continue
        i, j = keys[t]
        rows = np.sort(rng.choice(T[(i, k)].nrows, 48, replace=False))
        got = []
        for r in rows:
            p = int(np.searchsorted(rb, r, side="right") - 1)
            got.extend(device_rows(ctx, panels[p], [r - rb[p]]))
        same, rel = compare_rows(got, *oracle.spgemm(take_rows(T[(i, k)], rows), T[(k, j)]))
        ok &= same
        max_rel = max(max_rel, rel)
        checked += len(rows)
    res["oracle_sample"] = {"tiles": [list(keys[t]) for t in picks], "rows_checked": checked,
                            "structure_equal": bool(ok), "max_rel_err": max_rel}
    res["max_panels_per_tile"] = big
    log(f"[first stage] oracle sample: {res['oracle_sample']}, up to {big} panels per tile")
    b.close()
    # per-kernel HIP-event times of one eager pass, and the dominant kernel's
    # algorithmic bytes from each tile's row classes
    e = Batch(ctx, [dT[(i, k)] for i, j in keys], [dT[(k, j)] for i, j in keys], graph=False)
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    e.run()
    ctx.synchronize()
    ctx.set_timing(False)
    kt = ctx.timings()
    e.close()
    lb = hb = 0
    for t, (i, j) in enumerate(keys):
        At, Bt = T[(i, k)], T[(k, j)]
        blen = np.diff(Bt.ptr.astype(np.int64))
        f = np.zeros(At.nrows, np.int64)
        np.add.at(f, At.row_ids(), blen[At.idx])
        l_, h_ = class_bytes(np.diff(At.ptr.astype(np.int64)), f, per_row_out[t], 12 * Bt.nnz + 4 * (Bt.nrows + 1),
                             tun["light_max_flops"])
        lb += l_
        hb += h_

    def alg(name, launches):
        if name == "num_heavy":
            return (hb / max(launches, 1), "heavy rows of the 16 products: 12 B per C entry written and per A "
                                           "entry read + B read once per tile")
        if name == "num_light":
            return (lb / max(launches, 1), "light rows of the 16 products: 12 B per C entry written and per A "
                                           "entry read + B read once per tile")
        return None
    res["kernels_ms"] = {kk: round(v[1], 3) for kk, v in sorted(kt.items())}
    res["roofline"] = kernel_roofline(kt, 1, alg)
    res.update({"workload": f"R-MAT scale {scale} (ef 16, seed 2) cut {g}x{g}; SUMMA stage k={k} of ranks that "
                            f"start there: {g * g} tile products, no partial C, results as int32 row panels",
                "mults": mults, "nnz_out": nnz_out, "tuning": tun, "tuning_source": tsrc})
    ctx.close()
    return res


def galerkin_section(n, steps, tuning="auto"):
    """AMG Galerkin product C = P^T (A P): 27-point Laplacian on n^3 (26/-1),
    smoothed-aggregation P (3^3 aggregates, damped Jacobi w = 2/3).  Plain:
    two tr_spgemm calls; batched: the two products as captured graphs (P^T
    from the on-device transpose, resident).  Checked against the oracle on
    sampled rows of C."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from trident_amd import workloads
    from trident_amd.device import Batch, Context, spgemm, spgemm_flops
    tun, tsrc = use_tuning(f"galerkin:{n}", tuning)
    t0 = time.perf_counter()
    A, P = workloads.galerkin_operators(n)
    log(f"[galerkin] operators ready in {time.perf_counter() - t0:.1f}s: A {A.nrows} x {A.ncols} nnz {A.nnz}, "
        f"P {P.nrows} x {P.ncols} nnz {P.nnz}")
    ctx = Context(0)
    dA, dP = ctx.upload(A), ctx.upload(P)
    ctx.set_timing(True)
    ctx.reset_timing()
    t = time.perf_counter()
    dPt = dP.transpose()
    ctx.synchronize()
    t_tr = time.perf_counter() - t
    tr_kernels = {k: round(v[1], 3) for k, v in ctx.timings().items()}
    ctx.set_timing(False)
    AP = spgemm(dA, dP)
    mults = spgemm_flops(dA, dP) + spgemm_flops(dPt, AP)
    C = spgemm(dPt, AP)
    ctx.synchronize()
    ref = C.download()
    APh = AP.download()
    AP.free()
    C.free()

    def plain():
        ap = spgemm(dA, dP)
        c = spgemm(dPt, ap)
        ctx.synchronize()
        return ap, c
    ap, c = plain()
    ts = []
    for _ in range(steps):
        ap.free(), c.free()
        t = time.perf_counter()
        ap, c = plain()
        ts.append(time.perf_counter() - t)
    ap.free(), c.free()
    t_plain = float(np.median(ts))
    b1 = Batch(ctx, [dA], [dP])
    ap_b = b1.result(0)
    b2 = Batch(ctx, [dPt], [ap_b])
    ts = []
    for _ in range(steps):
        t = time.perf_counter()
        b1.run()
        b2.run()
        ctx.synchronize()
        ts.append(time.perf_counter() - t)
    b1.check()
    b2.check()
    t_b = float(np.median(ts))
    got = b2.result(0).download()
    ok = bool(np.array_equal(got.ptr, ref.ptr) and np.array_equal(got.idx, ref.idx) and np.array_equal(got.val, ref.val))
    # oracle on sampled rows of C: the AP rows they need, then those rows of P^T (AP)
    rng = np.random.default_rng(5)
    Pt = oracle.transpose(P)
    from trident_amd import gen
    Pt = gen.from_arrays(P.ncols, P.nrows, *Pt)
    rows = np.sort(rng.choice(Pt.nrows, 64, replace=False))
    needed = np.unique(np.concatenate([Pt.idx[Pt.ptr[r]:Pt.ptr[r + 1]] for r in rows]))
    APn = oracle.spgemm(take_rows(A, needed), P)
    APs = scatter_rows(A.nrows, P.ncols, needed, *APn)
    Cr = oracle.spgemm(take_rows(Pt, rows), APs)
    gi = [(got.idx[got.ptr[r]:got.ptr[r + 1]], got.val[got.ptr[r]:got.ptr[r + 1]]) for r in rows]
    same, max_rel = compare_rows(gi, *Cr)
    # per-kernel HIP-event times of one eager pass of both products
    e1 = Batch(ctx, [dA], [dP], graph=False)
    e2 = Batch(ctx, [dPt], [e1.result(0)], graph=False)
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    e1.run()
    e2.run()
    ctx.synchronize()
    ctx.set_timing(False)
    kt = ctx.timings()
    e2.close()
    e1.close()
    cut = tun["light_max_flops"]
    # per-row classes of both products
    alen_a, alen_pt = np.diff(A.ptr.astype(np.int64)), np.diff(Pt.ptr.astype(np.int64))
    plen, aplen = np.diff(P.ptr.astype(np.int64)), np.diff(APh.ptr.astype(np.int64))
    f1 = np.zeros(A.nrows, np.int64)
    np.add.at(f1, A.row_ids(), plen[A.idx])
    f2 = np.zeros(Pt.nrows, np.int64)
    np.add.at(f2, Pt.row_ids(), aplen[Pt.idx])
    l1, h1 = class_bytes(alen_a, f1, aplen, 12 * P.nnz + 4 * (P.nrows + 1), cut)
    l2, h2 = class_bytes(alen_pt, f2, np.diff(got.ptr.astype(np.int64)), 12 * APh.nnz + 4 * (APh.nrows + 1), cut)

    def alg(name, launches):
        if name == "num_light":
            return ((l1 + l2) / max(launches, 1), "light rows of both products: 12 B per C entry written and per "
                                                  "A entry read + B read once per product")
        if name == "num_heavy":
            return ((h1 + h2) / max(launches, 1), "heavy rows of both products: 12 B per C entry written and per "
                                                  "A entry read + B read once per product")
        return None
    step_bytes = (12 * (A.nnz + P.nnz) + 4 * (A.nrows + P.nrows + 2) + 12 * (Pt.nnz + APh.nnz) +
                  4 * (Pt.nrows + APh.nrows + 2) + 12 * (APh.nnz + got.nnz) + 4 * (APh.nrows + got.nrows + 2))
    res = {"workload": f"C = P^T A P, A = 27-point Laplacian {n}^3 (26/-1), P = SA prolongator (3^3 aggregates, "
                       "unnormalised tentative P, w = 2/3)", "mults": mults,
           "transpose_ms": round(1000 * t_tr, 3), "transpose_kernels_ms": tr_kernels,
           "plain": {"ms": round(1000 * t_plain, 3), "GFLOP/s": round(2 * mults / t_plain / 1e9, 2)},
           "batched": {"ms": round(1000 * t_b, 3), "GFLOP/s": round(2 * mults / t_b / 1e9, 2)},
           "speedup_batched_vs_plain": round(t_plain / t_b, 3),
           "batched_matches_plain": ok, "nnz_AP": APh.nnz, "nnz_C": got.nnz,
           "oracle_sample": {"rows_checked": int(len(rows)), "structure_equal": bool(same), "max_rel_err": max_rel},
           "kernels_ms": {kk: round(v[1], 3) for kk, v in sorted(kt.items())},
           "roofline": kernel_roofline(kt, 1, alg),
           "step_alg_GBs": round(step_bytes / t_b / 1e9, 2), "step_frac": round(step_bytes / t_b / 1e9 / HBM_PEAK_GBS, 5),
           "tuning": tun, "tuning_source": tsrc}
    log(f"[galerkin] plain {res['plain']} batched {res['batched']} oracle {res['oracle_sample']}")
    b2.close()
    b1.close()
    ctx.close()
    return res


def serve_section(nreq, rates, budget_gb, workers, tuning="auto"):
    """Served stream of `nreq` mixed requests with Poisson arrivals: 60% A*A on
    R-MAT scales 12-18 (edge factor 8 or 16), 40% P^T A P on 27-point
    Laplacians 16^3-96^3 (seed 3).  For each arrival rate (a multiple of what
    the plain path sustains: every request alone, one after another) it
    reports p50/p99 latency, throughput and the ratio to plain; probe
    requests are checked against the oracle."""
    import threading  # noqa: F401  (the server's workers are native threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from trident_amd import gen
    from trident_amd._lib import TridentError
    from trident_amd.device import (Context, DeviceCsx, Server, TR_REQ_GALERKIN, TR_REQ_SPGEMM, spgemm,
                                    spgemm_flops)
    tun, tsrc = use_tuning("serve", tuning)
    rng = np.random.default_rng(3)
    t0 = time.perf_counter()
    ctx = Context(0)
    host_rm, rm = {}, {}
    for s_ in range(12, 19):
        for ef in (8, 16):
            host_rm[(s_, ef)] = gen.rmat(s_, ef, seed=100 * s_ + ef)
            rm[(s_, ef)] = ctx.upload(host_rm[(s_, ef)])
    host_lp, lp = {}, {}
    for n in (16, 32, 48, 64, 80, 96):
        host_lp[n] = (gen.laplace27(n), gen.sa_prolongator(n))
        lp[n] = (ctx.upload(host_lp[n][0]), ctx.upload(host_lp[n][1]))
    log(f"[serve] operands ready in {time.perf_counter() - t0:.1f}s")
    # plain reference: every request type alone, one after another
    kinds = [("spgemm", k) for k in rm] + [("galerkin", n) for n in lp]
    plain_t, flops = {}, {}
    for kind, key in kinds:
        for rep in range(2):
            if kind == "spgemm":
                t = time.perf_counter()
                c = spgemm(rm[key], rm[key])
                ctx.synchronize()
                dt = time.perf_counter() - t
                if rep == 1:
                    flops[(kind, key)] = spgemm_flops(rm[key], rm[key])
                c.free()
            else:
                A, P = lp[key]
                Pt = P.transpose()
                ctx.synchronize()
                t = time.perf_counter()
                ap = spgemm(A, P)
                c = spgemm(Pt, ap)
                ctx.synchronize()
                dt = time.perf_counter() - t
                if rep == 1:
                    flops[(kind, key)] = spgemm_flops(A, P) + spgemm_flops(Pt, ap)
                ap.free(), c.free(), Pt.free()
        plain_t[(kind, key)] = dt
    ctx.pool_trim()  # the reference runs' outputs stay cached in this context's pool otherwise
    reqs = []
    for _ in range(nreq):
        if rng.random() < 0.6:
            reqs.append(("spgemm", (int(rng.integers(12, 19)), int(rng.choice([8, 16])))))
        else:
            reqs.append(("galerkin", int(rng.choice(list(lp)))))
    tot_flops = sum(flops[r] for r in reqs)
    tot_plain = sum(plain_t[r] for r in reqs)
    plain_gflops = 2 * tot_flops / tot_plain / 1e9
    log(f"[serve] plain: {plain_gflops:.2f} GFLOP/s over the mix ({tot_plain:.1f} s sequential)")
    srv = Server(mem_budget=int(budget_gb * (1 << 30)), workers=workers, max_batch=32)
    ops = {k: srv.register(v) for k, v in rm.items()}
    lops = {n: (srv.register(A), srv.register(P)) for n, (A, P) in lp.items()}
    capacity = nreq / tot_plain  # requests/s the plain path sustains
    results = []
    for factor in rates:
        rate = factor * capacity
        gaps = rng.exponential(1.0 / rate, size=nreq)
        srv.reset_stats()
        ids = []
        t_start = time.perf_counter()
        t_next = t_start
        for i, (kind, key) in enumerate(reqs):
            t_next += gaps[i]
            now = time.perf_counter()
            if t_next > now:
                time.sleep(t_next - now)
            if kind == "spgemm":
                ids.append(srv.submit(TR_REQ_SPGEMM, ops[key], ops[key], discard=True))
            else:
                ids.append(srv.submit(TR_REQ_GALERKIN, lops[key][0], lops[key][1], discard=True))
            if i % 500 == 0:
                log(f"[serve] rate x{factor}: submitted {i}/{nreq}")
        failed = 0
        for rid in ids:
            try:
                srv.wait(rid, want_result=False)
            except TridentError as e:
                failed += 1
                log(f"[serve] request {rid} failed: {e}")
            srv.release(rid)
        wall = time.perf_counter() - t_start
        st = srv.stats()
        st.update({"failed_requests": failed, "rate_x_plain_capacity": factor, "arrival_rate_per_s": round(rate, 2),
                   "wall_s": round(wall, 2), "served_GFLOP/s": round(2 * tot_flops / wall / 1e9, 2),
                   "vs_plain": round((2 * tot_flops / wall / 1e9) / plain_gflops, 3)})
        st = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}
        log(f"[serve] {st}")
        results.append(st)
    # probes: served results of small requests against the oracle
    probes = [("spgemm", (12, 8)), ("spgemm", (13, 16)), ("galerkin", 16)]
    pid = []
    for kind, key in probes:
        if kind == "spgemm":
            pid.append(srv.submit(TR_REQ_SPGEMM, ops[key], ops[key]))
        else:
            pid.append(srv.submit(TR_REQ_GALERKIN, lops[key][0], lops[key][1]))
    probe_ok = True
    for (kind, key), rid in zip(probes, pid):
        d, _ = srv.wait(rid)
        H = DeviceCsx(ctx, d).download()
        if kind == "spgemm":
            ep, ei, ev = oracle.spgemm(host_rm[key], host_rm[key])
        else:
            Ah, Ph = host_lp[key]
            APo = gen.from_arrays(Ah.nrows, Ph.ncols, *oracle.spgemm(Ah, Ph))
            Pto = gen.from_arrays(Ph.ncols, Ph.nrows, *oracle.transpose(Ph))
            ep, ei, ev = oracle.spgemm(Pto, APo)
        probe_ok &= bool(np.array_equal(H.ptr.astype(np.int64), ep) and np.array_equal(H.idx, ei)
                         and np.array_equal(H.val.view(np.uint64), ev.view(np.uint64)))
        srv.release(rid)
    srv.close()
    ctx.close()
    best = max(results, key=lambda r: r["served_GFLOP/s"])
    return {"workload": f"{nreq} requests, Poisson arrivals: 60% A*A on R-MAT s12-18 (ef 8/16), 40% P^T A P on "
                        "27-pt Laplacians 16^3-96^3", "requests": nreq, "plain_GFLOP/s": round(plain_gflops, 2),
            "best_served_GFLOP/s": best["served_GFLOP/s"], "best_vs_plain": best["vs_plain"], "rates": results,
            "oracle_probes": {"requests": [f"{k}:{v}" for k, v in probes], "bit_exact": probe_ok},
            "workers": workers, "tuning": tun, "tuning_source": tsrc}


def section_line(args, res, value_key, ms_key, steps):
    """The one-line JSON of a form run on its own (--workload stage:/galerkin:/serve)."""
    return {"metric": METRIC, "value": value_key, "unit": "GFLOP/s", "n_gpus": 1, "steps": steps,
            "warmup": 1, "ms_per_step": ms_key, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic", "config": {"workload": res["workload"], "forms": res}}


def stage_bench(args):
    _, scale, k = args.workload.split(":")
    res = stage_section(int(scale), int(k), args.steps, args.tuning, with_compress=True)
    print(json.dumps(section_line(args, res, res["batched"]["GFLOP/s"], res["batched"]["ms"], args.steps)), flush=True)
    return 0


def galerkin_bench(args):
    res = galerkin_section(int(args.workload.split(":")[1]), args.steps, args.tuning)
    print(json.dumps(section_line(args, res, res["batched"]["GFLOP/s"], res["batched"]["ms"], args.steps)), flush=True)
    return 0


def serve_bench(args):
    parts = args.workload.split(":")
    nreq = int(parts[1]) if len(parts) > 1 else 2000
    res = serve_section(nreq, args.serve_rates, args.serve_budget_gb, args.serve_workers, args.tuning)
    best = max(res["rates"], key=lambda r: r["served_GFLOP/s"])
    print(json.dumps(section_line(args, res, best["served_GFLOP/s"], round(1000 * best["wall_s"] / nreq, 3), nreq)),
          flush=True)
    return 0


def copy_bench(args):
    """Plain device copy of a known size: the calibration run for the
    FETCH_SIZE / WRITE_SIZE counters, whose readings for the copy kernel are
    compared with the bytes it must move.  copy:G uses torch's elementwise
    y = x * 1 over fp64; hipcopy:G this library's own streaming copy kernel
    (tr_calibration_copy), a second, independent calibration."""
    gib = float(args.workload.split(":")[1])
    if args.workload.startswith("hipcopy:"):
        from trident_amd.device import Context
        nbytes = (int(gib * (1 << 30)) // 16) * 16
        ctx = Context(0)
        src, dst = ctx.malloc(nbytes), ctx.malloc(nbytes)
        ctx.check(ctx.lib.tr_memset(ctx.h, ctypes.c_void_p(src), 1, nbytes), "memset")

        def run():
            ctx.check(ctx.lib.tr_calibration_copy(ctx.h, ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes), "copy")
        sync = ctx.synchronize
        n = nbytes // 8
    else:
        import torch
        n = int(gib * (1 << 30)) // 8
        x = torch.rand(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)

        def run():
            torch.mul(x, 1.0, out=y)
        sync = torch.cuda.synchronize
    for _ in range(max(args.warmup, 1)):
        run()
    sync()
    t = time.perf_counter()
    for _ in range(args.steps):
        run()
    sync()
    dt = (time.perf_counter() - t) / max(args.steps, 1)
    nbytes = 8 * n
    out = {"metric": "device copy GB/s", "value": round(2 * nbytes / dt / 1e9, 2), "unit": "GB/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1000 * dt, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"copy of {nbytes} bytes (read {nbytes}, write {nbytes} per launch)",
                      "read_KiB": nbytes / 1024, "write_KiB": nbytes / 1024}}
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat:20")
    ap.add_argument("--cpu-mults", type=float, default=1.5e9,
                    help="multiplications in the CPU-baseline row sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--serve-budget-gb", type=float, default=160.0)
    ap.add_argument("--serve-workers", type=int, default=1,
                    help="server worker threads (one batch at a time keeps the GPU's caches for one product)")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[1.0, 4.0])
    ap.add_argument("--sections", default="all",
                    help="forms measured after the plain product on its default workload: all, or none, or a "
                         "comma list of batched,first_stage,galerkin,served")
    ap.add_argument("--stage", default="22:3", help="batched form: R-MAT scale and SUMMA stage, S:k")
    ap.add_argument("--first-stage", default="22:2",
                    help="batched first stage (no partial C, results beyond int32 as panels): S:k")
    ap.add_argument("--galerkin", type=int, default=128, help="Galerkin form: Laplacian edge n (n^3 grid)")
    ap.add_argument("--serve-requests", type=int, default=2000, help="served form: requests in the stream")
    ap.add_argument("--tuning", default="auto",
                    help="auto (trident_amd/tuning/<workload>.json if present), none, or a JSON path")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU's share of the path; --gpus {args.gpus} is not supported "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    if args.cpu_child:
        return cpu_child()
    # the plain product's CPU-baseline child starts before this process loads
    # the HIP library or touches the GPU
    plain = not args.workload.startswith(("copy:", "hipcopy:", "stage:", "galerkin:", "serve"))
    child = start_cpu_child() if plain and not args.no_cpu else None
    if args.workload.startswith(("copy:", "hipcopy:")):
        return copy_bench(args)
    args.tuning_used, args.tuning_source = apply_tuning(args)
    if args.workload.startswith("stage:"):
        return stage_bench(args)
    if args.workload.startswith("galerkin:"):
        return galerkin_bench(args)
    if args.workload.startswith("serve"):
        return serve_bench(args)

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
    heavy = f > args.tuning_used["light_max_flops"]
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
    traffic, traffic_src = load_traffic(dname, args.workload)
    achieved = alg / per_launch_s / 1e9 if per_launch_s > 0 else 0.0
    bytes_alg_total = 12 * nnzC + 4 * (A.nrows + 1) + 2 * B_bytes
    roofline = {
        "bound": "hbm",
        "kernel": dname,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": None if traffic is None else int(traffic),
        "traffic_source": traffic_src,
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
        cpu, check = cpu_baseline_and_check(child, A, f, ctx, panels, rb, hptrs, args.cpu_mults)
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
            "tuning": args.tuning_used,
            "tuning_source": args.tuning_source,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_sample": check,
        "kernels": kernels,
    }
    for P in panels:
        P.free()
    ctx.close()
    # the batched, Galerkin and served forms, each with its dominant kernel's
    # roofline and an oracle check on sampled rows
    want = set() if args.sections == "none" else (
        {"batched", "first_stage", "galerkin", "served"} if args.sections == "all" else set(args.sections.split(",")))
    if args.workload != "rmat:20":
        want = set() if args.sections == "all" else want
    if "batched" in want:
        sc, kk = (int(x) for x in args.stage.split(":"))
        out["batched"] = stage_section(sc, kk, args.steps, args.tuning)
    if "first_stage" in want:
        sc, kk = (int(x) for x in args.first_stage.split(":"))
        out["batched_first_stage"] = first_stage_section(sc, kk, args.steps, args.tuning)
    if "galerkin" in want:
        out["galerkin"] = galerkin_section(args.galerkin, args.steps, args.tuning)
    if "served" in want:
        out["served"] = serve_section(args.serve_requests, args.serve_rates, args.serve_budget_gb,
                                      args.serve_workers, args.tuning)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
