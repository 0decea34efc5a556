"""C ABI checks that need no GPU: the library loads and exports every symbol
that include/*.h declares; the driver and the benchmark refuse what they do
not support without touching the device."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "lu2_amd", "liblu2mi.so")
DRIVER = os.path.join(ROOT, "lu2_amd", "lu2mi")


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:int|void|double|char|long)\s*\*?\s*(\w+)\s*\(", text, re.M):
            names.append(m.group(1))
    return names


def test_headers_declare_the_seams():
    names = set(declared_functions())
    for sym in ["mygdrdgemm", "gdrblas_dgemm", "gdr_check_and_restart", "gdr_dgemm_", "gdr_init",
                "gdr_free", "gdrdgemm_set_procname", "gdrsetboardid", "gdrsetnboards",
                "gdrsetforceswapab", "mydgemm", "matmul_for_small_nk", "reset_gdr", "lumcolumn",
                "lu2mi_init", "lu2mi_lumcolumn_dev", "lu2mi_randomsetmat"]:
        assert sym in names, sym


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(LIB)          # loading does not initialise the GPU
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        getattr(lib, n)


def test_python_binding_raises_without_library(tmp_path):
    import lu2_amd

    saved = lu2_amd._lib
    lu2_amd._lib = None
    try:
        with pytest.raises(lu2_amd.Lu2miError):
            lu2_amd.load_library(str(tmp_path / "missing.so"))
    finally:
        lu2_amd._lib = saved


def test_driver_usage():
    assert os.path.exists(DRIVER)
    r = subprocess.run([DRIVER, "-h"], capture_output=True, text=True)
    assert r.returncode == 1
    assert "-n: size of matrix" in r.stderr and "-b: block size" in r.stderr


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_hip_sources_disable_packed_fp32_and_use_gfx950():
    src = open(os.path.join(ROOT, "lu2_amd", "build.py")).read()
    assert "-packed-fp32-ops" in src and "gfx950" in src
