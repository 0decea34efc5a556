"""Build liblu2mi.so (HIP, gfx950), the lu2mi driver, and the CPU oracle.

    python -m lu2_amd.build          # or __graft_entry__.build()

Everything is built in-tree so the shared objects travel with the source.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["lu2mi_gemm.hip", "lu2mi_lu.hip", "lu2mi_mixed.hip", "lu2mi_batched.hip", "lu2mi_lowp.hip",
           "lu2mi_lowp_wide.hip", "lu2mi_api.hip"]
LIB = os.path.join(HERE, "liblu2mi.so")
DRIVER = os.path.join(HERE, "lu2mi")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed-FP32 instructions are kept out of the generated code on purpose
# MFMA accumulators in ordinary VGPRs: with AGPR accumulators the compiler
# copies every accumulator across the K-loop back-edge.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# MFMA accumulators in ordinary VGPRs, except where a wave holds more
# accumulators than the 256 architectural VGPRs (the 16-bit wide tile keeps
# its 256 in AGPRs)
VGPR_FORM = ["-mllvm", "-amdgpu-mfma-vgpr-form=true"]
AGPR_ACC_SOURCES = {"lu2mi_lowp_wide.hip"}
# Kernels that must not spill registers to scratch: the direct-to-LDS DGEMM
# tile counts on its loads being the only vector-memory operations in flight
# at its barriers (a spilling variant gave wrong products).  The build fails
# if the compiler reports scratch use for any kernel in these files.
NOSPILL_SOURCES = {"lu2mi_gemm.hip"}
RESOURCE_REMARKS = ["-Rpass-analysis=kernel-resource-usage"]


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def _check_no_scratch(src: str, remarks: str) -> None:
    """Fail when the compiler's resource-usage remarks give any kernel of src
    a nonzero ScratchSize (register spills)."""
    name, seen = None, 0
    for line in remarks.splitlines():
        if "Function Name:" in line:
            name = line.split("Function Name:", 1)[1].split()[0]
        elif "ScratchSize [bytes/lane]:" in line:
            seen += 1
            size = int(line.split("ScratchSize [bytes/lane]:", 1)[1].split()[0])
            if size:
                raise RuntimeError(f"{src}: kernel {name} uses {size} bytes/lane of scratch (register spills)")
    if not seen:
        raise RuntimeError(f"{src}: no resource-usage remarks from the compiler")


def build_library() -> str:
    from concurrent.futures import ThreadPoolExecutor

    objs = [os.path.join(CSRC, "_build", src.replace(".hip", ".o")) for src in SOURCES]
    os.makedirs(os.path.join(CSRC, "_build"), exist_ok=True)
    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    def compile_one(so):
        src, obj = so
        r = _run([HIPCC, *HIPFLAGS, *([] if src in AGPR_ACC_SOURCES else VGPR_FORM),
                  *(RESOURCE_REMARKS if src in NOSPILL_SOURCES else []),
                  "-I", os.path.join(ROOT, "include"), "-I", CSRC,
                  "-c", os.path.join(CSRC, src), "-o", obj])
        if src in NOSPILL_SOURCES:
            _check_no_scratch(src, r.stderr)

    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(compile_one, zip(SOURCES, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> str:
    _run(["gcc", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(HERE, "lu2mi_main.c"), "-o", DRIVER, "-L", HERE, "-llu2mi",
          "-Wl,-rpath,$ORIGIN"])
    return DRIVER


REFERENCE = os.environ.get("LU2_REFERENCE", "/root/reference")


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    # lu2's own driver from the reference sources, linked against the library
    # (only where the reference is present; the built binary travels)
    if os.path.exists(os.path.join(REFERENCE, "lu2.c")):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={REFERENCE}"])


def build_all() -> None:
    build_library()
    build_driver()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB, DRIVER)
