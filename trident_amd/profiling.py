"""Summaries of rocprofv3 kernel traces (host-side analysis of the CSV that
`rocprofv3 --kernel-trace --output-format csv` writes).

overlap: how long kernels of one group (e.g. the partial-C add) ran while
kernels of another group (e.g. the multiply) were running on a different
queue -- the measure of the accumulation overlap of a stage loop, where the
add of stage s runs on the accumulation thread's stream while the main
thread multiplies stage s + 1.

    python -m trident_amd.profiling overlap TRACE.csv --a k_add_ --b k_num_,k_sym_,k_row_flops
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sys
from typing import Dict, List, Sequence, Tuple


def read_trace(path: str) -> List[Dict[str, object]]:
    """Kernel dispatches of a rocprofv3 kernel-trace CSV (a file, or a
    directory searched for *kernel_trace.csv): name, queue, start, end (ns)."""
    if os.path.isdir(path):
        files = sorted(glob.glob(os.path.join(path, "**", "*kernel_trace.csv"), recursive=True))
        if not files:
            raise FileNotFoundError(f"no *kernel_trace.csv under {path}")
    else:
        files = [path]
    rows = []
    for fn in files:
        with open(fn, newline="") as f:
            for r in csv.DictReader(f):
                q = r.get("Stream_Id") or r.get("Queue_Id") or "0"
                rows.append({"name": r["Kernel_Name"], "queue": f"{r.get('Agent_Id', '')}:{q}",
                             "start": int(r["Start_Timestamp"]), "end": int(r["End_Timestamp"])})
    rows.sort(key=lambda d: d["start"])
    return rows


def _match(name: str, keys: Sequence[str]) -> bool:
    return any(k in name for k in keys)


def _union(iv: List[Tuple[int, int]]) -> List[Tuple[int, int]]:
    out: List[Tuple[int, int]] = []
    for s, e in sorted(iv):
        if out and s <= out[-1][1]:
            out[-1] = (out[-1][0], max(out[-1][1], e))
        else:
            out.append((s, e))
    return out


def _intersect(a: List[Tuple[int, int]], b: List[Tuple[int, int]]) -> int:
    i = j = 0
    tot = 0
    while i < len(a) and j < len(b):
        s, e = max(a[i][0], b[j][0]), min(a[i][1], b[j][1])
        if s < e:
            tot += e - s
        if a[i][1] < b[j][1]:
            i += 1
        else:
            j += 1
    return tot


def overlap(rows: List[Dict[str, object]], keys_a: Sequence[str], keys_b: Sequence[str]) -> Dict[str, object]:
    """Busy time of group A, of group B, and the time during which a group-A
    kernel ran while a group-B kernel ran on another queue (ns and the share
    of A's busy time), plus the queues each group used."""
    a = [r for r in rows if _match(str(r["name"]), keys_a)]
    b = [r for r in rows if _match(str(r["name"]), keys_b)]
    tot = 0
    qa = sorted({str(r["queue"]) for r in a})
    for q in qa:
        ua = _union([(int(r["start"]), int(r["end"])) for r in a if r["queue"] == q])
        ub = _union([(int(r["start"]), int(r["end"])) for r in b if r["queue"] != q])
        tot += _intersect(ua, ub)
    busy_a = sum(e - s for s, e in _union([(int(r["start"]), int(r["end"])) for r in a]))
    busy_b = sum(e - s for s, e in _union([(int(r["start"]), int(r["end"])) for r in b]))
    span = (max(int(r["end"]) for r in rows) - min(int(r["start"]) for r in rows)) if rows else 0
    return {"kernels_a": len(a), "kernels_b": len(b), "queues_a": qa,
            "queues_b": sorted({str(r["queue"]) for r in b}), "busy_a_ms": busy_a / 1e6,
            "busy_b_ms": busy_b / 1e6, "a_during_b_ms": tot / 1e6,
            "a_during_b_share": (tot / busy_a) if busy_a else 0.0, "trace_span_ms": span / 1e6}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    o = sub.add_parser("overlap", help="time group-A kernels ran alongside group-B kernels on other queues")
    o.add_argument("trace", help="kernel_trace.csv or a rocprofv3 output directory")
    o.add_argument("--a", default="k_add_", help="comma list of name fragments of group A")
    o.add_argument("--b", default="k_num_,k_sym_,k_row_flops", help="comma list of name fragments of group B")
    args = ap.parse_args(argv)
    res = overlap(read_trace(args.trace), args.a.split(","), args.b.split(","))
    for k, v in res.items():
        print(f"{k}: {v:.3f}" if isinstance(v, float) else f"{k}: {v}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
