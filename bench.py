"""Benchmark: lu2 LU + solve on one MI355X (FP64 path).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 65536] [--nb 512]
                    [--pivot gepp|compat] [--no-cpu]

A step is one full factor + solve of lu2's augmented n x (n+1) system
(lu2.c:1122 lumcolumn: blocked recursive LU with partial pivoting, then the
back substitution) with the matrix already resident in HBM.  The matrix is
lu2's drand48 matrix (lu2.c:70-85, seed 1), generated on the device once;
before every step the pristine copy is restored device-to-device, outside the
timed region.  Each step is bracketed by synchronisations and timed; the
value is 2n^3/3 flops (the metric's count, lu2.c:1304) over the mean step
time.  The HPL-2 scaled residual of the last solution is checked (< 16).

The roofline object describes the dominant kernel, the trailing-update
DGEMM (A22 -= L21*U12, lu2.c:580): its algorithmic flops over the summed
duration of its launches, from HIP events recorded by the library on the
stream the kernel runs on, against the MI355X FP64 matrix peak.

mixed_precision: the same system solved by an FP32 factorization (FP32 MFMA
trailing updates) and FP64 iterative refinement to the same residual bar,
with its time-to-solution, iterations and error against the FP64 solution.

cpu_baseline: the CPU oracle (a C restatement of lu2's algorithm, OpenMP) on
a bounded sample, n=16384, on the host cores.
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
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n: int, nb: int, pivot: int):
    import numpy as np

    from oracle import lu2_oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    a0 = O.randomsetmat(n, 1)
    t = time.perf_counter()
    x, _, _ = O.lumcolumn(a0, nb, pivot)
    dt = time.perf_counter() - t
    res = O.residuals(a0, x)
    flops = 2.0 * n ** 3 / 3.0
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"oracle lumcolumn (C restatement of lu2.c:1122, OpenMP) n={n} nb={nb} "
                      f"seed 1, {dt:.2f} s wall, HPL-2 residual {res[6]:.3g}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--nb", type=int, default=512)
    ap.add_argument("--pivot", choices=["gepp", "compat"], default="gepp")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=16384)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-steps", type=int, default=1)
    args = ap.parse_args()

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
    for s in range(args.steps):
        dt = one_step(True)
        tm = ctx.timing()
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
        # FP32 factorization + FP64 refinement on the same matrix (dA0 is read only)
        x64 = dx.download(np.float64, n)
        dxm = ctx.alloc(8 * n)
        try:
            tms = []
            for w in range(args.warmup + args.mixed_steps):
                L._lib.lu2mi_synchronize(ctx.handle)
                t0 = time.perf_counter()
                info = ctx.solve_mixed(n, dA0, lda, nb, dxm, L.PREC_FP32, maxiter=30, target=1.0)
                dt = time.perf_counter() - t0
                if w >= args.warmup:
                    tms.append(dt)
                log(f"mixed {'warmup' if w < args.warmup else 'step'}: {dt:.3f} s  iterations "
                    f"{info['iterations']} (gmres {info['gmres_used']}) residual {info['resid_hpl2']:.3g}")
            xm = dxm.download(np.float64, n)
            rm = ctx.residual_dev(n, dA0, lda, dxm)
            tm = sum(tms) / len(tms)
            mixed = {"precision": "fp32 factors, fp32 MFMA updates, fp64 refinement",
                     "value": flops / tm / 1e12, "unit": "TFLOP/s", "ms_per_solve": tm * 1e3,
                     "factor_ms": info["t_factor_ms"], "refine_ms": info["t_refine_ms"],
                     "iterations": info["iterations"], "gmres_used": bool(info["gmres_used"]),
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
                   "n": n, "nb": nb, "pivot": args.pivot},
        "tflops_2n3_3_plus_2n2": (flops + 2.0 * n * n) / mean / 1e12,
        "frac_of_fp64_peak": flops / mean / 1e12 / FP64_MFMA_PEAK_TFLOPS,
        "hpl2_residual": float(res[6]),
        "roofline": {
            "bound": "mfma",
            "kernel": "gemm_f64_kernel (trailing update A22 -= L21*U12)",
            "achieved": achieved,
            "peak": FP64_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": (achieved / FP64_MFMA_PEAK_TFLOPS) if achieved else None,
            "traffic": None,
            "launches_per_step": gemm_launch // max(1, args.steps),
            "avg_launch_ms": gemm_ms / max(1, gemm_launch),
        },
    }
    if mixed is not None:
        out["mixed_precision"] = mixed
    for b in (dA0, dA, dx, dp):
        b.free()
    ctx.close()
    if not args.no_cpu:
        log(f"cpu baseline: oracle n={args.cpu_n}")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, 256, L.PIV_GEPP)
    print(json.dumps(out), flush=True)
    return 0 if res[6] < 16 else 1


if __name__ == "__main__":
    sys.exit(main())
