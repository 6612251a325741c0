"""The kernel-trace summary used for the serving profiles: GPU busy time is the
union of the kernel intervals and kernels in flight = total kernel time / busy
time, on a hand-made trace with known answers (CPU only)."""
import csv
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

HDR = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id", "Kernel_Id", "Kernel_Name",
       "Correlation_Id", "Start_Timestamp", "End_Timestamp", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
       "Accum_VGPR_Count", "SGPR_Count", "Workgroup_Size_X", "Workgroup_Size_Y", "Workgroup_Size_Z",
       "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]


def test_busy_union_and_concurrency(tmp_path):
    t = 10 ** 15
    kernels = [("void bielcc_impl::k_panel_lds<double, 64>(int, int)", 1000, 3000, 40),
               ("Cijk_Ailk_Bljk_ZB_MT128x64x8_MI16x16x4x1_SN_1LDSB1", 2000, 5000, 500),
               ("void bielcc_impl::k_trsv<float, 8>(int)", 7000, 8000, 10),
               ("void bielcc_impl::k_panel_lds<float, 64>(int, double*, int*)", 8000, 8500, 40),
               ("void bielcc_impl::k_panel_lds<double, 64>(int, int)", 20000, 26000, 100)]   # second window
    with open(tmp_path / "k.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HDR)
        for n, a, b, g in kernels:
            w.writerow(["KERNEL_DISPATCH", 1, 1, 1, 1, 1, 1, n, 1, t + a, t + b, 0, 0, 0, 0, 0, 256, 1, 1, 256 * g,
                        1, 1])
    c = lambda x: {"mono": t + x, "boot": t + x}
    json.dump({"a": [c(0), c(10000)], "b": [c(20000), c(30000)]}, open(tmp_path / "w.json", "w"))
    out = bench.trace_summary(str(tmp_path / "k.csv"), str(tmp_path / "w.json")).splitlines()
    a = next(s for s in out if s.startswith("a:"))
    b = next(s for s in out if s.startswith("b:"))
    # window a: busy [1000, 5000] + [7000, 8500] = 5500 of 10000 ns; kernel time 6500 ns
    assert "4 kernels" in a and "(55%)" in a and "(1.18 in flight" in a
    assert "1 kernels" in b and "(60%)" in b and "(1.00 in flight" in b
    panels = [s for s in out if s.startswith("k_panel_lds<double>")]
    assert panels[0].endswith("<=64: 2 us (1)") and panels[1].endswith("<=128: 6 us (1)")
    # the FP32 panel (double* among its parameters) is counted as FP32
    assert [s for s in out if s.startswith("k_panel_lds<float>")] == ["k_panel_lds<float> by workgroups: <=64: 0 us (1)"]
