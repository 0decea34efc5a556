"""bench.py's host-side helpers (no GPU): the per-kernel counter table of
rocprofv3 --pmc output and the kernel-name shortening it groups by."""
import csv
import io
import os
import sys
from contextlib import redirect_stdout

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
          "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
          "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def write_run(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, FIELDS)
        w.writeheader()
        for did, name, val, t0, t1 in rows:
            w.writerow({k: 0 for k in FIELDS} | {"Dispatch_Id": did, "Kernel_Name": name, "Counter_Name": counter,
                                                 "Counter_Value": val, "Start_Timestamp": t0, "End_Timestamp": t1})


def test_short_kernel_name():
    assert bench.short_kernel_name("void leaf_kernel<double, 16, 4>(LeafArgs<double>)") == "leaf_kernel<double, 16, 4>"
    assert bench.short_kernel_name("gemm2_lower_kernel(GemmDesc, int, int)") == "gemm2_lower_kernel"
    assert (bench.short_kernel_name("void gemm16_kernel<3>((anonymous namespace)::H16<3>::t const*, int)")
            == "gemm16_kernel<3>")
    assert bench.short_kernel_name("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_pmc_table_scales_on_the_copy(tmp_path):
    """A 1 GiB copy whose FETCH_SIZE reads half its bytes sets the scale to 2;
    a kernel's bytes, rate and dispatch count follow from it."""
    gib = 1 << 30
    fetch, write = tmp_path / "fetch.csv", tmp_path / "write.csv"
    copy = "__amd_rocclr_copyBuffer"
    k = "void gemm2_kernel<1>(GemmDesc, int, int)"
    write_run(fetch, "FETCH_SIZE", [(1, copy, gib / 1024 / 2, 0, 1000), (2, k, 1e6, 2000, 1002000),
                                    (3, k, 1e6, 2000000, 3000000)])
    write_run(write, "WRITE_SIZE", [(1, copy, gib / 1024, 0, 1000), (2, k, 5e5, 2000, 1002000),
                                    (3, k, 5e5, 2000000, 3000000)])
    out = io.StringIO()
    with redirect_stdout(out):
        assert bench.pmc_table([str(fetch), str(write)], float(gib)) == 0
    text = out.getvalue()
    row = next(line for line in text.splitlines() if "gemm2_kernel<1>" in line)
    cells = [c.strip() for c in row.strip("|").split("|")]
    assert cells[1] == "2"                                   # dispatches
    assert abs(float(cells[2]) - 2.0) < 1e-9                 # ms: two 1 ms dispatches
    fetched = 2 * 1e6 * 1024 * 2.0 / 1e9                     # KiB x 1024 x scale
    written = 2 * 5e5 * 1024 * 1.0 / 1e9
    assert abs(float(cells[3]) - fetched) < 1e-3
    assert abs(float(cells[4]) - written) < 1e-3
    assert abs(float(cells[5]) - (fetched + written) / 2e-3) < 1.0
    assert "FETCH_SIZE x 2.000, WRITE_SIZE x 1.000" in text
