"""Benchmark: lu2 LU + solve on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 65536] [--nb 1024]
                    [--pivot gepp|compat] [--no-cpu] [--no-mixed] [--no-lowp]
                    [--no-batch]

A step is one full factor + solve of lu2's augmented n x (n+1) system
(lu2.c:1122 lumcolumn: blocked recursive LU with partial pivoting, then the
back substitution) with the matrix already resident in HBM.  The matrix is
lu2's drand48 matrix (lu2.c:70-85, seed 1), generated on the device once;
before every step the pristine copy is restored device-to-device, outside the
timed region.  Each step is bracketed by synchronisations and timed; the
value is 2n^3/3 flops (the metric's count, lu2.c:1304) over the mean step
time.  The HPL-2 scaled residual of the last solution is checked (< 16).

roofline: the dominant kernel, the trailing-update DGEMM (A22 -= L21*U12,
lu2.c:580): its algorithmic flops over the summed duration of its launches,
from HIP events recorded by the library on the stream the kernel runs on,
against the MI355X FP64 matrix peak.

mixed_precision: the same system solved by an LU of an FP32 copy whose
trailing updates run on FP16 operands split hi + lo (FP16x2) or on FP32
operands, then FP64 refinement (classical, GMRES-IR when it stalls) to the
HPL-2 residual target; time-to-solution, iterations, error against the FP64
solution, speed-up over the FP64 step.

lowp_16bit: the 16-bit configuration -- lu2's matrix made diagonally
dominant (a_ii += sum_j |a_ij|), n=98304, trailing updates on FP16 operands
with per-block power-of-two scaling, FP64 refinement; error against an FP64
GPU solve of the same system.

batched: 4096 independent systems, n ~ U{32..512}, 1-4 right-hand sides,
per-system drand48 streams, in one graph-captured persistent launch with the
autotuned shape; FP64 and FP32-factor + FP64-refinement forms.

cpu_baseline: the CPU oracle (a C restatement of lu2's algorithm, OpenMP) on
the host cores: lu2's own n=2048 nb=256 and a bounded sample, n=16384 (the
value); --cpu-full adds n=32768 nb=512.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK_TFLOPS = 78.6   # MI355X dense FP64 matrix, published
TRAIL_KERNEL_NAME = {1: "gemm2_kernel<1> (trailing update A22 -= L21*U12, operands loaded straight into LDS)",
                     0: "gemm_kernel<double, true, 1, 1> (trailing update A22 -= L21*U12, register-staged)"}
HBM_PEAK_GBS = 8000.0
# HBM bytes of one M=N=32768 trailing update per (tile, K), measured with
# rocprofv3 FETCH_SIZE (x 2.000) + WRITE_SIZE (x 1.000) on bench.py --gemm-probe
TRAIL_TRAFFIC = {(1, 2048): (68.70e9, "trailing_dgemm_m32768_k2048_pmc.md"),
                 (1, 1024): (42.91e9, "trailing_dgemm_m32768_k1024_pmc.md")}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n: int, nb: int, pivot: int, full: bool):
    """The CPU oracle (lu2o_lumcolumn: lu2's blocked recursive LU, OpenMP,
    AVX2) on the host cores: lu2's own configuration n=2048 nb=256, then the
    bounded sample n x nb that gives `value` (and, with full, the drop-in
    configuration's size n=32768 nb=512, about 1.5-3 minutes).  Timed region
    as lu2.c:1287-1296: factorization and back substitution."""
    from oracle import lu2_oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    runs = []
    sizes = [(2048, 256), (n, nb)] + ([(32768, 512)] if full else [])
    for nn, bb in sizes:
        a0 = O.randomsetmat(nn, 1)
        t = time.perf_counter()
        x, _, _ = O.lumcolumn(a0, bb, pivot)
        dt = time.perf_counter() - t
        res = O.residuals(a0, x)
        runs.append({"n": nn, "nb": bb, "s": dt, "tflops": 2.0 * nn ** 3 / 3.0 / dt / 1e12,
                     "hpl2_residual": float(res[6])})
        log(f"cpu oracle n={nn} nb={bb}: {dt:.2f} s {runs[-1]['tflops'] * 1e3:.1f} GFLOP/s")
        del a0, x
    main = runs[1]
    return {"value": main["tflops"], "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"oracle lumcolumn (C restatement of lu2.c:1122, OpenMP, {threads} threads on "
                      f"{cpu_model()}) n={main['n']} nb={main['nb']} seed 1, {main['s']:.2f} s wall, "
                      f"HPL-2 residual {main['hpl2_residual']:.3g}",
            "runs": runs}


def lowp_config(ctx, n: int, nb: int) -> dict:
    """16-bit trailing updates on lu2's matrix made diagonally dominant."""
    import numpy as np

    import lu2_amd as L

    lda = n + 16
    dA0 = ctx.alloc(8 * n * lda)
    dx = ctx.alloc(8 * n)
    ctx.randomsetmat(n, 1, dA0, lda)
    ctx.make_diag_dominant(n, dA0, lda)
    ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
    res = {}
    for name, prec in (("fp16", L.PREC_FP16), ("bf16", L.PREC_BF16)):
        L._lib.lu2mi_synchronize(ctx.handle)
        t0 = time.perf_counter()
        info = ctx.solve_mixed(n, dA0, lda, nb, dx, prec, maxiter=30, target=1.0)
        dt = time.perf_counter() - t0
        res[name] = {"ms_per_solve": dt * 1e3, "value": 2.0 * n ** 3 / 3 / dt / 1e12, "unit": "TFLOP/s",
                     "factor_ms": info["t_factor_ms"], "refine_ms": info["t_refine_ms"],
                     "iterations": info["iterations"], "gmres_used": bool(info["gmres_used"]),
                     "hpl2_residual_initial": info["resid_hpl2_initial"],
                     "hpl2_residual": float(ctx.residual_dev(n, dA0, lda, dx)[6]),
                     "x": dx.download(np.float64, n)}
        log(f"lowp {name} n={n} nb={nb}: {dt:.3f} s iterations {info['iterations']} "
            f"residual {res[name]['hpl2_residual']:.3g}")
    # the FP64 GPU solve of the same system, for the error figures
    dA = ctx.alloc(8 * n * lda)
    dA.copy_from(dA0)
    L._lib.lu2mi_synchronize(ctx.handle)
    t0 = time.perf_counter()
    ctx.lumcolumn_dev(n, dA, lda, 512, None, dx)
    t64 = time.perf_counter() - t0
    x64 = dx.download(np.float64, n)
    for b in (dA, dA0, dx):
        b.free()
    for v in res.values():
        v["err_vs_fp64"] = float(np.abs(v.pop("x") - x64).max() / np.abs(x64).max())
        v["speedup_vs_fp64"] = t64 * 1e3 / v["ms_per_solve"]
    log(f"lowp reference FP64 solve n={n}: {t64:.3f} s")
    return {"workload": f"lu2 matrix + diag(row abs-sum), n={n}, nb={nb}, pivot=gepp, seed 1",
            "fp64_ms": t64 * 1e3, **res}


def batched_config(ctx, seed: int) -> dict:
    """4096 systems, n ~ U{32..512}, 1-4 right-hand sides, tuned launch shape."""
    import numpy as np

    import lu2_amd as L

    rng = np.random.default_rng(2024)
    nsys = 4096
    n = rng.integers(32, 513, nsys)
    nrhs = rng.integers(1, 5, nsys)
    flops = float((2.0 * n.astype(float) ** 3 / 3 + 2.0 * n.astype(float) ** 2 * nrhs).sum())
    out = {"workload": f"{nsys} systems, n ~ U{{32..512}}, nrhs ~ U{{1..4}}, per-system drand48 "
                       f"(seed base {seed} + system id), one graph-captured persistent launch",
           "flops_per_batch": flops}
    dA = dX = None
    for name, prec in (("fp64", L.PREC_FP64), ("fp32_ir", L.PREC_FP32)):
        b = L.Batch(ctx, n, nrhs, prec=prec, maxiter=10, target=1.0)
        if dA is None:
            dA = ctx.alloc(8 * b.a_elems)
            dX = ctx.alloc(8 * b.x_elems)
            b.generate(dA, seed)
        cfg = b.tune(dA, dX, reps=2)
        infos = [b.solve(dA, dX) for _ in range(4)]       # first replay captures the graph
        ms = sum(i["t_ms"] for i in infos[1:]) / 3
        info = infos[-1]
        _, iters, _ = b.results()
        out[name] = {"ms_per_batch": ms, "systems_per_s": nsys / ms * 1e3, "value": flops / ms / 1e9,
                     "unit": "TFLOP/s", "frac_of_fp64_peak": flops / ms / 1e9 / FP64_MFMA_PEAK_TFLOPS,
                     "max_hpl2_residual": info["max_resid_hpl2"], "failed": info["n_failed"],
                     "max_iterations": int(iters.max()), "mean_iterations": float(iters.mean()),
                     "tuned": {"threads_per_system": cfg[0], "panel_width_small": cfg[1],
                               "panel_width_large": cfg[2], "split_n": cfg[3]}}
        log(f"batched {name}: {ms:.2f} ms  {nsys / ms * 1e3:.0f} systems/s  max residual "
            f"{info['max_resid_hpl2']:.3g}  tuned {cfg}")
        b.close()
    dA.free()
    dX.free()
    return out


def gemm_probe(m: int, n: int, k: int, reps: int, trail_kernel: int, lower: bool = False) -> int:
    """The trailing-update product on its own: C[m x n] -= A[m x k] B[k x n]
    inside one row-major matrix (A = rows k.., columns 0..k; B = rows 0..k,
    columns k..; C = rows k.., columns k..), as the factorization lays it
    out, with the n=65536 factorization's leading dimension 65552 (or
    m + k + 16 if larger), after a 1 GiB device-to-device copy that
    calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters.  With lower, the
    U12 product's form instead: C = A B with A lower triangular (entries
    above its diagonal are not read), whose flops are 2 sum_r min(r+1, k) n.
    Prints one JSON line with the median time and rate over `reps` launches."""
    import lu2_amd as L

    lda = max(m + k + 16, 65552)
    ctx = L.Context(0)
    ctx.set_option(L.OPT_TRAIL_KERNEL, trail_kernel)
    ctx.set_option(L.OPT_DGEMM_TRAILING, 2 if lower else 1)
    dX = ctx.alloc(8 * (m + k) * lda)
    ctx.randomsetmat(m + k, 1, dX, lda)
    cal = ctx.alloc(1 << 30)
    tmp = ctx.alloc(1 << 30)
    tmp.copy_from(cal)                       # the calibration copy
    offA, offB, offC = 8 * k * lda, 8 * k, 8 * (k * lda + k)
    times = []
    for _ in range(reps):
        L._lib.lu2mi_synchronize(ctx.handle)
        t0 = time.perf_counter()
        if lower:
            ctx.dgemm_dev(m, n, k, 1.0, dX, lda, dX, lda, 0.0, dX, lda, offs=(offA, offB, offC))
        else:
            ctx.dgemm_dev(m, n, k, -1.0, dX, lda, dX, lda, 1.0, dX, lda, offs=(offA, offB, offC))
        times.append(time.perf_counter() - t0)
    for b in (dX, cal, tmp):
        b.free()
    ctx.close()
    t = sorted(times)[len(times) // 2]
    fl = 2.0 * n * sum(min(r + 1, k) for r in range(m)) if lower else 2.0 * m * n * k
    kname = ("gemm2_lower_kernel (U12 form, row tiles paired)" if trail_kernel else
             "gemm_lower_kernel<double> (U12 form, register-staged)") if lower else TRAIL_KERNEL_NAME[trail_kernel]
    print(json.dumps({"probe": "U12 product" if lower else "trailing-update DGEMM", "kernel": kname,
                      "M": m, "N": n, "K": k, "lda": lda, "reps": reps,
                      "median_ms": t * 1e3, "tflops": fl / t / 1e12,
                      "frac_of_fp64_peak": fl / t / 1e12 / FP64_MFMA_PEAK_TFLOPS,
                      "algorithmic_bytes": 8.0 * ((1 if lower else 2) * m * n + m * k + k * n)}), flush=True)
    return 0


def leaf_probe(rows: int, reps: int) -> int:
    """One 16-column panel leaf on its own: lu2's column-major panel seam
    (lu2tlib.c:1130 cm_column_decomposition_recursive, m = 16, the width
    lu2.c:1082 uses) on a random rows x 16 panel, which the library factors
    with one leaf launch (plus its host<->device copies).  For counter runs
    of the leaf kernel; prints the wall time of each call."""
    import ctypes

    import numpy as np

    import lu2_amd as L

    lib = ctypes.CDLL(L.LIB_PATH)
    lib.cm_column_decomposition_recursive.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                      ctypes.c_void_p, ctypes.c_int]
    rng = np.random.default_rng(rows)
    a0 = rng.uniform(-0.5, 0.5, (16, rows))
    pv = np.zeros(16, np.int32)
    for r in range(reps):
        a = a0.copy()
        t0 = time.perf_counter()
        lib.cm_column_decomposition_recursive(rows, a.ctypes.data, 16, pv.ctypes.data, 0)
        print(json.dumps({"probe": "panel leaf (16 columns)", "rows": rows, "call": r,
                          "wall_ms": (time.perf_counter() - t0) * 1e3,
                          "leaf_bytes": 2.0 * rows * 16 * 8}), flush=True)
    return 0


def host_gemm_probe(m: int, n: int, k: int) -> int:
    """A host-memory product through the reference's mydgemm seam
    (lu2lib.c:2079, C -= A B on malloc'd row-major matrices): operands cross
    PCIe through the library's pinned bounce buffers in overlapped row chunks.
    Prints the wall time and the PCIe bytes per second of the second call
    (the first allocates the staging)."""
    import ctypes

    import numpy as np

    import lu2_amd as L

    lib = ctypes.CDLL(L.LIB_PATH)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.mydgemm.argtypes = [ci, ci, ci, cd, vp, ci, vp, ci, cd, vp, ci]
    lib.lu2mi_host_gemm_stats.argtypes = [vp, ci]
    rng = np.random.default_rng(0)
    a = rng.uniform(-0.5, 0.5, (m, k))
    b = rng.uniform(-0.5, 0.5, (k, n))
    c = rng.uniform(-0.5, 0.5, (m, n))
    for _ in range(2):
        t0 = time.perf_counter()
        lib.mydgemm(m, n, k, -1.0, a.ctypes.data, k, b.ctypes.data, n, 1.0, c.ctypes.data, n)
        dt = time.perf_counter() - t0
    st = np.zeros(4)
    lib.lu2mi_host_gemm_stats(st.ctypes.data, 4)
    print(json.dumps({"probe": "host-memory mydgemm seam", "M": m, "N": n, "K": k, "wall_ms": dt * 1e3,
                      "bytes_up": st[0], "bytes_down": st[1], "chunks": int(st[3]),
                      "pcie_gbs": (st[0] + st[1]) / dt / 1e9,
                      "pcie_peak_gbs_per_direction": 64.0,
                      "tflops": 2.0 * m * n * k / dt / 1e12}), flush=True)
    return 0


def short_kernel_name(name: str) -> str:
    """`void leaf_kernel<double, 16, 4>(double*, ...)` -> `leaf_kernel<double, 16, 4>`."""
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    name = name.replace("(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and i > 0:
            return name[:i]
    return name


def pmc_table(paths: list, copy_bytes: float | None) -> int:
    """Per-kernel totals of rocprofv3 --pmc counter files (run_counter_collection.csv,
    one per run of the same program, dispatches in the same order): dispatches,
    summed duration, and for the counters present FETCH_SIZE / WRITE_SIZE in GB
    (scaled so that the largest device-to-device copy of the run, copy_bytes
    long, reads and writes copy_bytes), the HBM rate over the summed duration,
    and the matrix-core busy fraction SQ_VALU_MFMA_BUSY_CYCLES over
    GRBM_GUI_ACTIVE / 8 XCDs x 1024 SIMDs.  Prints a markdown table."""
    import csv
    from collections import defaultdict

    ctr = defaultdict(lambda: defaultdict(float))     # kernel -> counter -> sum
    disp = defaultdict(int)
    ns = defaultdict(int)
    copy_val = {}
    for k, p in enumerate(paths):
        seen = set()
        with open(p, newline="") as f:
            for r in csv.DictReader(f):
                name = short_kernel_name(r["Kernel_Name"])
                cname, val = r["Counter_Name"], float(r["Counter_Value"])
                ctr[name][cname] += val
                if name == "__amd_rocclr_copyBuffer":
                    copy_val[cname] = max(copy_val.get(cname, 0.0), val)
                did = r["Dispatch_Id"]
                if k == 0 and did not in seen:
                    seen.add(did)
                    disp[name] += 1
                    ns[name] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    scale = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        scale[c] = (copy_bytes / 1024.0 / copy_val[c]) if (copy_bytes and copy_val.get(c)) else 1.0
    have = {c for d in ctr.values() for c in d}
    hdr = ["kernel", "dispatches", "ms"]
    if "FETCH_SIZE" in have:
        hdr.append("fetched GB")
    if "WRITE_SIZE" in have:
        hdr.append("written GB")
    if have & {"FETCH_SIZE", "WRITE_SIZE"}:
        hdr.append("HBM GB/s")
    mfma = {"SQ_VALU_MFMA_BUSY_CYCLES", "GRBM_GUI_ACTIVE"} <= have
    if mfma:
        hdr.append("MFMA busy %")
    print("| " + " | ".join(hdr) + " |")
    print("|" + "---|" + "---:|" * (len(hdr) - 1))
    for name in sorted(ctr, key=lambda x: -ns[x]):
        d, ms = ctr[name], ns[name] / 1e6
        row = [f"`{name}`", str(disp[name]), f"{ms:.2f}"]
        gb = 0.0
        for c, col in (("FETCH_SIZE", "fetched GB"), ("WRITE_SIZE", "written GB")):
            if col in hdr:
                v = d.get(c, 0.0) * 1024.0 * scale[c] / 1e9
                gb += v
                row.append(f"{v:.3f}")
        if "HBM GB/s" in hdr:
            row.append(f"{gb / (ms / 1e3):.0f}" if ms > 0 else "-")
        if mfma:
            den = d.get("GRBM_GUI_ACTIVE", 0.0) / 8.0 * 1024.0
            row.append(f"{100.0 * d.get('SQ_VALU_MFMA_BUSY_CYCLES', 0.0) / den:.1f}" if den > 0 else "-")
        print("| " + " | ".join(row) + " |")
    print(f"\nscale: FETCH_SIZE x {scale['FETCH_SIZE']:.3f}, WRITE_SIZE x {scale['WRITE_SIZE']:.3f}")
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--nb", type=int, default=1024)
    ap.add_argument("--pivot", choices=["gepp", "compat"], default="gepp")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=16384)
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the CPU oracle at n=32768 nb=512 (minutes)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-steps", type=int, default=1)
    ap.add_argument("--mixed-nb", type=int, default=None,
                    help="block size of both mixed solves (default: 2048 FP16x2, 1024 FP32)")
    ap.add_argument("--no-lowp", action="store_true")
    ap.add_argument("--lowp-n", type=int, default=98304)
    ap.add_argument("--lowp-nb", type=int, default=2048)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--leaf-w", type=int, default=16, choices=(16, 32),
                    help="panel leaf width (columns factored per leaf launch)")
    ap.add_argument("--lookahead", type=int, default=1, choices=(0, 1),
                    help="factor the next panel on a second stream during the trailing update")
    ap.add_argument("--subpanel", type=int, default=None,
                    help="widest panel part factored by one persistent launch (0 = one launch per "
                         "step; library default if omitted)")
    ap.add_argument("--wide16", type=int, default=0, choices=(0, 1),
                    help="16-bit trailing updates in 256 x 256 tiles (1) or 128 x 128 (0, default)")
    ap.add_argument("--backsub128", type=int, default=0, choices=(0, 1),
                    help="back substitutions in 128-row blocks (1) or 64 (0, default)")
    ap.add_argument("--strict-sync", type=int, default=0, choices=(0, 1),
                    help="leaf and back substitution publish with agent-scope release/acquire")
    ap.add_argument("--gemm-probe", type=int, nargs=3, metavar=("M", "N", "K"), default=None,
                    help="time the trailing-update DGEMM alone at this shape (counter runs)")
    ap.add_argument("--probe-reps", type=int, default=1)
    ap.add_argument("--probe-lower", action="store_true",
                    help="--gemm-probe of the U12 product's form (A lower triangular, C = A B)")
    ap.add_argument("--leaf-probe", type=int, default=None, metavar="ROWS",
                    help="one 16-column panel leaf of ROWS rows through the column-major seam")
    ap.add_argument("--host-gemm", type=int, nargs=3, metavar=("M", "N", "K"), default=None,
                    help="time a host-memory product through the mydgemm seam (PCIe staging)")
    ap.add_argument("--trail-kernel", type=int, default=1, choices=(0, 1),
                    help="trailing-update DGEMM tile: 1 direct-to-LDS (default), 0 register-staged")
    ap.add_argument("--pmc-table", nargs="+", metavar="CSV", default=None,
                    help="summarise rocprofv3 --pmc counter files per kernel (no GPU use)")
    ap.add_argument("--pmc-copy-bytes", type=float, default=None,
                    help="bytes of the run's largest device-to-device copy (counter calibration)")
    args = ap.parse_args()

    if args.pmc_table:
        return pmc_table(args.pmc_table, args.pmc_copy_bytes)
    if args.gemm_probe:
        return gemm_probe(*args.gemm_probe, args.probe_reps, args.trail_kernel, args.probe_lower)
    if args.host_gemm:
        return host_gemm_probe(*args.host_gemm)
    if args.leaf_probe:
        return leaf_probe(args.leaf_probe, args.probe_reps)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the LU on a single MI355X "
              "(no multi-GPU path); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np

    import lu2_amd as L

    n, nb = args.n, args.nb
    lda = n + 16
    piv_mode = L.PIV_GEPP if args.pivot == "gepp" else L.PIV_COMPAT
    ctx = L.Context(0)
    ncu, arch = ctx.device_info()
    log(f"device {arch} CUs={ncu}; n={n} nb={nb} pivot={args.pivot}")
    ctx.set_option(L.OPT_PIVMODE, piv_mode)
    ctx.set_option(L.OPT_LOOKAHEAD, args.lookahead)
    ctx.set_option(L.OPT_LEAF_W, args.leaf_w)
    ctx.set_option(L.OPT_TRAIL_KERNEL, args.trail_kernel)
    ctx.set_option(L.OPT_STRICT_SYNC, args.strict_sync)
    ctx.set_option(L.OPT_WIDE16, args.wide16)
    ctx.set_option(L.OPT_BACKSUB128, args.backsub128)
    if args.subpanel is not None:
        ctx.set_option(L.OPT_SUBPANEL, args.subpanel)
    nbytes = 8 * n * lda
    dA0 = ctx.alloc(nbytes)
    dA = ctx.alloc(nbytes)
    dx = ctx.alloc(8 * n)
    dp = ctx.alloc(4 * n)
    ctx.randomsetmat(n, args.seed, dA0, lda)
    log("matrix generated on device")

    def one_step(timed: bool) -> float:
        dA.copy_from(dA0)                      # restore, outside the timed region
        ctx.set_option(L.OPT_TIMING, 1 if timed else 0)
        L._lib.lu2mi_synchronize(ctx.handle)
        t0 = time.perf_counter()
        ctx.lumcolumn_dev(n, dA, lda, nb, dp, dx)   # returns after the device is idle
        return time.perf_counter() - t0

    flops = 2.0 * n ** 3 / 3.0
    for w in range(args.warmup):
        dt = one_step(False)
        log(f"warmup {w}: {dt:.3f} s")
    times, gemm_ms, gemm_fl, gemm_launch = [], 0.0, 0.0, 0
    subs = {}
    for s in range(args.steps):
        dt = one_step(True)
        tm = ctx.timing()
        for name, v in ctx.subsystem_timing().items():
            if isinstance(v, dict):
                acc = subs.setdefault(name, {"ms": 0.0, "flops": 0.0, "bytes": 0.0})
                for k in acc:
                    acc[k] += v[k]
        times.append(dt)
        gemm_ms += tm["gemm_ms"]
        gemm_fl += tm["gemm_flops"]
        gemm_launch += tm["gemm_launches"]
        log(f"step {s}: {dt:.3f} s  {2 * n ** 3 / 3 / dt / 1e12:.2f} TFLOP/s  trailing DGEMM "
            f"{tm['gemm_ms']:.1f} ms over {tm['gemm_launches']} launches")
    res = ctx.residual_dev(n, dA0, lda, dx)
    log(f"HPL-2 scaled residual {res[6]:.4g}")
    mean = sum(times) / len(times)

    mixed = None
    if not args.no_mixed:
        # FP32-copy LU with FP16x2 / FP32 trailing updates + FP64 refinement on
        # the same matrix (dA0 is read only)
        x64 = dx.download(np.float64, n)
        dxm = ctx.alloc(8 * n)
        mixed = {}
        try:
            # block sizes per form (sweeps at n=65536): the FP16x2 updates round
            # their operands once per block update, so fewer, wider blocks keep
            # the factors accurate enough for 3 refinement steps (nb=1024: 8);
            # the FP32 form converges in 4 either way and nb=1024 shortens its
            # panels
            nb16 = args.mixed_nb or 2048
            nb32 = args.mixed_nb or 1024
            for name, prec, mnb in (("fp16x2", L.PREC_FP16X2, nb16), ("fp32", L.PREC_FP32, nb32)):
                tms = []
                for w in range(args.warmup + args.mixed_steps):
                    L._lib.lu2mi_synchronize(ctx.handle)
                    t0 = time.perf_counter()
                    info = ctx.solve_mixed(n, dA0, lda, mnb, dxm, prec, maxiter=30, target=1.0)
                    dt = time.perf_counter() - t0
                    if w >= args.warmup:
                        tms.append(dt)
                    log(f"mixed {name} {'warmup' if w < args.warmup else 'step'}: {dt:.3f} s  iterations "
                        f"{info['iterations']} (gmres {info['gmres_used']}) residual {info['resid_hpl2']:.3g}")
                xm = dxm.download(np.float64, n)
                rm = ctx.residual_dev(n, dA0, lda, dxm)
                tm = sum(tms) / len(tms)
                mixed[name] = {
                    "precision": ("FP32 factors; trailing updates on FP16 operands split hi+lo "
                                  "(hi*hi+hi*lo+lo*hi), per-block power-of-two scaling, FP32 accumulation"
                                  if prec == L.PREC_FP16X2 else
                                  "FP32 factors; FP32 MFMA trailing updates, FP32 accumulation"),
                    "nb": mnb, "value": flops / tm / 1e12, "unit": "TFLOP/s", "ms_per_solve": tm * 1e3,
                    "factor_ms": info["t_factor_ms"], "refine_ms": info["t_refine_ms"],
                    "iterations": info["iterations"], "gmres_used": bool(info["gmres_used"]),
                    "gmres_inner": info["gmres_inner"],
                    "hpl2_residual_initial": info["resid_hpl2_initial"], "hpl2_residual": float(rm[6]),
                    "err_vs_fp64": float(np.abs(xm - x64).max() / np.abs(x64).max()),
                    "speedup_vs_fp64": mean / tm}
        finally:
            dxm.free()

    achieved = gemm_fl / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else None
    out = {
        "metric": "LU solve TFLOP/s (2n^3/3), n=65536: FP64 vs mixed-prec+IR at HPL residual<16",
        "value": flops / mean / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mean * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: lu2 drand48 matrix (lu2.c:70-85) generated on the device",
        "config": {"workload": f"lu2 FP64 LU with partial pivoting + solve, n={n}, nb={nb}, "
                               f"pivot={args.pivot}, 1 RHS, seed {args.seed}",
                   "n": n, "nb": nb, "pivot": args.pivot, "lookahead": args.lookahead, "leaf_w": args.leaf_w,
                   "subpanel": args.subpanel if args.subpanel is not None else 0,
                   "trail_kernel": args.trail_kernel, "strict_sync": args.strict_sync, "wide16": args.wide16,
                   "backsub128": args.backsub128},
        "tflops_2n3_3_plus_2n2": (flops + 2.0 * n * n) / mean / 1e12,
        "frac_of_fp64_peak": flops / mean / 1e12 / FP64_MFMA_PEAK_TFLOPS,
        "hpl2_residual": float(res[6]),
        "roofline": {
            "bound": "mfma",
            "kernel": TRAIL_KERNEL_NAME[args.trail_kernel],
            "achieved": achieved,
            "peak": FP64_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": (achieved / FP64_MFMA_PEAK_TFLOPS) if achieved else None,
            # bytes of one trailing update of the factorization's M=N=32768, K=nb shape,
            # from FETCH_SIZE and WRITE_SIZE counter runs of --gemm-probe scaled on
            # its 1 GiB copy (profiles/trailing_dgemm_m32768_k{nb}_pmc.md); None for
            # a block size without such a run
            "traffic": TRAIL_TRAFFIC.get((args.trail_kernel, nb), (None, None))[0],
            "traffic_launch": f"M=N=32768 K={nb}",
            "traffic_algorithmic": 8.0 * (2 * 32768 * 32768 + 2 * 32768 * nb),
            "launches_per_step": gemm_launch // max(1, args.steps),
            "avg_launch_ms": gemm_ms / max(1, gemm_launch),
        },
    }
    # lu2's print_timers groups (timerlib.c:47-70) from the library's HIP
    # events, per step; with the lookahead schedule the panel's stream runs
    # beside the main one, so its time overlaps the others
    out["subsystems"] = {}
    for name, acc in subs.items():
        ms = acc["ms"] / args.steps
        e = {"ms_per_step": ms}
        if acc["flops"] > 0 and ms > 0:
            e["tflops"] = acc["flops"] / args.steps / ms / 1e9
            e["frac_of_fp64_peak"] = e["tflops"] / FP64_MFMA_PEAK_TFLOPS
        if acc["bytes"] > 0 and ms > 0:
            e["gbs"] = acc["bytes"] / args.steps / ms / 1e6
            e["frac_of_hbm"] = e["gbs"] / HBM_PEAK_GBS
        out["subsystems"][name] = e
    if mixed is not None:
        out["mixed_precision"] = mixed
    for b in (dA0, dA, dx, dp):
        b.free()
    if not args.no_lowp:
        out["lowp_16bit"] = lowp_config(ctx, args.lowp_n, args.lowp_nb)
    if not args.no_batch:
        out["batched"] = batched_config(ctx, args.seed)
    ctx.close()
    if not args.no_cpu:
        log(f"cpu baseline: oracle n={args.cpu_n}")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, 256, L.PIV_GEPP, args.cpu_full)
    print(json.dumps(out), flush=True)
    return 0 if res[6] < 16 else 1


if __name__ == "__main__":
    sys.exit(main())
