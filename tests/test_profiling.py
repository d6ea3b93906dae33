"""The kernel-trace overlap summary (trident_amd/profiling.py) on a
synthetic rocprofv3 CSV whose overlap is known by construction."""
import csv

from trident_amd import profiling

FIELDS = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"]


def _write(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for name, q, s, e in rows:
            w.writerow({"Kind": "KERNEL_DISPATCH", "Agent_Id": 1, "Queue_Id": q, "Stream_Id": q,
                        "Kernel_Name": name, "Start_Timestamp": s, "End_Timestamp": e})


def test_overlap_known_answer(tmp_path):
    p = tmp_path / "t_kernel_trace.csv"
    ns = 1_000_000
    _write(p, [
        ("void tr::(anonymous namespace)::k_num_heavy_wg<2048, 256>(int)", 1, 0, 10 * ns),
        ("void tr::(anonymous namespace)::k_add_pass<true>(int)", 2, 4 * ns, 14 * ns),   # 6 ms during the multiply
        ("void tr::(anonymous namespace)::k_sym_heavy<20, 1024>(int)", 1, 12 * ns, 13 * ns),  # +1 ms
        ("void tr::(anonymous namespace)::k_add_pass<false>(int)", 1, 20 * ns, 22 * ns),  # same queue: no overlap
        ("void tr::(anonymous namespace)::k_num_light<64, 4>(int)", 1, 21 * ns, 23 * ns),
    ])
    r = profiling.overlap(profiling.read_trace(str(p)), ["k_add_"], ["k_num_", "k_sym_"])
    assert r["kernels_a"] == 2 and r["kernels_b"] == 3
    assert abs(r["busy_a_ms"] - 12.0) < 1e-9
    assert abs(r["a_during_b_ms"] - 7.0) < 1e-9
    assert abs(r["a_during_b_share"] - 7.0 / 12.0) < 1e-9
    # directory form finds the same file
    r2 = profiling.overlap(profiling.read_trace(str(tmp_path)), ["k_add_"], ["k_num_", "k_sym_"])
    assert r2 == r
