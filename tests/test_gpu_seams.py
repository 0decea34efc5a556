"""The reference's link seams exported by liblu2mi.so, called with host
memory exactly as lu2's objects call them (gdrdgemm.h, the GRAPE-DR device
entry, lu2lib.h, lu2.c:1122 lumcolumn)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import lu2_amd

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = ctypes.POINTER
c_int, c_double = ctypes.c_int, ctypes.c_double


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(lu2_amd.LIB_PATH)
    dp = ctypes.c_void_p
    L.mygdrdgemm.argtypes = [c_int, c_int, c_int, c_double, dp, c_int, dp, c_int, c_double, dp, c_int]
    L.mydgemm.argtypes = L.mygdrdgemm.argtypes
    L.gdrblas_dgemm.argtypes = [c_int, c_int, c_int, c_int, c_int, c_int, c_double, dp, c_int, dp, c_int,
                                c_double, dp, c_int]
    L.gdr_dgemm_.argtypes = [P(c_int)] * 5 + [P(c_double)] * 2 + [P(c_int)] * 3 + [dp, dp, dp, P(c_int)]
    L.matmul_for_small_nk.argtypes = [c_int, dp, c_int, dp, c_int, dp, c_int, c_int, c_int]
    L.gdr_check_and_restart.argtypes = [dp, dp, dp]
    L.reset_gdr.argtypes = [c_int, dp, c_int, dp, c_int]
    L.lumcolumn.argtypes = [c_int, dp, dp, c_int, dp, dp, c_int]
    L.cm_column_decomposition_recursive.argtypes = [c_int, dp, c_int, dp, c_int]
    return L


def ptr(a):
    return a.ctypes.data


def bound(k, alpha, a, b, beta, c):
    u = 2.0 ** -53
    g = (k + 2) * u / (1 - (k + 2) * u)
    return 2 * g * (abs(alpha) * (np.abs(a) @ np.abs(b)) + abs(beta) * np.abs(c)) + 1e-300


def test_mygdrdgemm_strided(lib):
    rng = np.random.default_rng(0)
    m, n, k = 200, 150, 96
    A = rng.standard_normal((m, k + 7))
    B = rng.standard_normal((k, n + 3))
    C = rng.standard_normal((m, n + 5))
    want = -1.0 * A[:, :k] @ B[:, :n] + 1.0 * C[:, :n]
    lib.mygdrdgemm(m, n, k, -1.0, ptr(A), k + 7, ptr(B), n + 3, 1.0, ptr(C), n + 5)
    assert np.all(np.abs(C[:, :n] - want) <= bound(k, 1, A[:, :k], B[:, :n], 1, want))
    D = rng.standard_normal((m, n))
    lib.mydgemm(m, n, k, 2.0, ptr(A), k + 7, ptr(B), n + 3, 0.0, ptr(D), n)
    assert np.allclose(D, 2.0 * A[:, :k] @ B[:, :n], rtol=1e-13, atol=1e-12)


def test_gdr_dgemm_column_major(lib):
    """gdrdgemm.c:185-186 passes row-major operands as column-major with A<->B
    swapped; the device entry is column-major Fortran."""
    rng = np.random.default_rng(1)
    m, n, k = 70, 90, 40
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((k, n)))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    want = 0.5 * A @ B - 2.0 * C
    args = [ctypes.byref(c_int(v)) for v in (1, 1, m, n, k)]
    done = c_int(0)
    lib.gdr_dgemm_(*args, ctypes.byref(c_double(0.5)), ctypes.byref(c_double(-2.0)),
                   ctypes.byref(c_int(m)), ctypes.byref(c_int(k)), ctypes.byref(c_int(m)),
                   ptr(A), ptr(B), ptr(C), ctypes.byref(done))
    assert done.value == 1
    assert np.allclose(C, want, rtol=1e-13, atol=1e-12)
    # a transposed operand is not handled by the device: doneflag stays 0
    done2 = c_int(7)
    lib.gdr_dgemm_(ctypes.byref(c_int(0)), *args[1:], ctypes.byref(c_double(1.0)), ctypes.byref(c_double(0.0)),
                   ctypes.byref(c_int(m)), ctypes.byref(c_int(k)), ctypes.byref(c_int(m)),
                   ptr(A), ptr(B), ptr(C), ctypes.byref(done2))
    assert done2.value == 0


@pytest.mark.parametrize("order,ta,tb", [(101, 111, 111), (102, 111, 111), (101, 112, 111), (102, 111, 112),
                                         (101, 112, 112)])
def test_gdrblas_dgemm(lib, order, ta, tb):
    rng = np.random.default_rng(order + ta + tb)
    M, N, K = 33, 47, 29
    opA = rng.standard_normal((M, K))
    opB = rng.standard_normal((K, N))
    C = rng.standard_normal((M, N))
    want = 1.5 * opA @ opB + 0.25 * C
    # storage of A, B, C for the requested order / transposes
    def store(op, trans):
        mat = op.T if trans == 112 else op
        return np.ascontiguousarray(mat) if order == 101 else np.asfortranarray(mat)
    As, Bs = store(opA, ta), store(opB, tb)
    Cs = np.ascontiguousarray(C) if order == 101 else np.asfortranarray(C)
    lda = As.shape[1] if order == 101 else As.shape[0]
    ldb = Bs.shape[1] if order == 101 else Bs.shape[0]
    ldc = N if order == 101 else M
    lib.gdrblas_dgemm(order, ta, tb, M, N, K, 1.5, ptr(As), lda, ptr(Bs), ldb, 0.25, ptr(Cs), ldc)
    assert np.allclose(Cs, want, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("m,n,k,beta", [(5000, 2000, 600, 1.0), (3001, 777, 64, 0.0), (40, 30, 0, 0.5)])
def test_host_gemm_chunked(lib, m, n, k, beta):
    """Host-memory products pass through pinned bounce buffers in row chunks
    (32 MiB each; 5000 x (600 + 2000) doubles is four chunks) whose packing,
    transfers and products overlap: the result is the product, every chunk
    lands in its rows, and the bytes moved are those of the operands."""
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((m, k + 3))
    B = rng.standard_normal((k, n + 5))
    C = rng.standard_normal((m, n + 7))
    want = -1.0 * A[:, :k] @ B[:, :n] + beta * C[:, :n]
    lib.mydgemm(m, n, k, -1.0, ptr(A), k + 3, ptr(B), n + 5, beta, ptr(C), n + 7)
    assert np.allclose(C[:, :n], want, rtol=1e-12, atol=1e-11)
    st = np.zeros(4)
    lib.lu2mi_host_gemm_stats.argtypes = [ctypes.c_void_p, c_int]
    assert lib.lu2mi_host_gemm_stats(st.ctypes.data, 4) == 0
    up = 8.0 * (k * n + m * k + (m * n if beta != 0 else 0))
    assert st[0] == up and st[1] == 8.0 * m * n
    rows = (32 << 20) // (8 * max(k + (n if beta != 0 else 0), n))
    assert st[3] == -(-m // min(rows, m))


def test_matmul_for_small_nk(lib):
    rng = np.random.default_rng(2)
    m, kk, n = 300, 16, 8
    a = rng.standard_normal((m, kk + 4))
    b = rng.standard_normal((kk, n + 2))
    c = rng.standard_normal((m, n + 1))
    want = c[:, :n] - a[:, :kk] @ b[:, :n]
    lib.matmul_for_small_nk(kk + 4, ptr(a), n + 2, ptr(b), n + 1, ptr(c), m, kk, n)
    assert np.allclose(c[:, :n], want, rtol=1e-13, atol=1e-12)


def test_gdr_check_and_restart_and_reset_gdr(lib, capfd):
    N = 2048
    a, b, c = (np.zeros((N, N)) for _ in range(3))
    lib.gdr_check_and_restart(ptr(a), ptr(b), ptr(c))
    assert np.array_equal(c, b) and b[3, 5] == 3 * N + 5
    lib.reset_gdr(16, None, 64, None, 16)   # scratch is the library's own
    err = capfd.readouterr().err
    assert "gdr_check_and_restart passed" in err


@pytest.mark.parametrize("n,nb,seed", [(512, 64, 1), (1024, 128, 12345)])
def test_lumcolumn_host_dropin(lib, oracle, n, nb, seed):
    """lu2.c:1122 lumcolumn(n, a, b, nb, awork, pv, 1) on a host a[n][n+16]:
    the GPU factors with lu2's COMPAT rule; a holds the stored factors, b the
    solution, pv the last block's pivots."""
    a = oracle.randomsetmat(n, seed)
    a0 = a.copy()
    b = np.zeros(n)
    pv = np.zeros(n, np.int32)
    awork = np.zeros((nb, n))
    lib.lumcolumn(n, ptr(a), ptr(b), nb, ptr(awork), ptr(pv), 1)
    xo, po, _ = oracle.lumcolumn(a0, nb, oracle.PIV_COMPAT)
    assert np.array_equal(pv[:nb], po[n - nb:])
    assert np.abs(b - xo).max() <= 1e-5 * np.abs(xo).max()
    # the seam is copy-in, the device factorization, copy-out: bit-identical
    # to the device-pointer entry point on the same input
    xd, pd, fd = lu2_amd.solve(a0, nb, lu2_amd.PIV_COMPAT)
    assert np.array_equal(pd[n - nb:], pv[:nb])
    assert np.array_equal(b, xd)
    assert np.array_equal(a[:, : n + 1], fd[:, : n + 1])


# more than 32768 rows: the leaf holds four rows per thread; more than 65536:
# more than 64 workgroups, so fewer threads poll each exchange slot (three,
# two, and one at 128 workgroups)
@pytest.mark.parametrize("n,m,i", [(300, 16, 0), (1000, 16, 64), (777, 32, 100), (40000, 16, 0), (41000, 16, 1000),
                                   (70000, 16, 0), (100000, 16, 0), (130100, 16, 0)])
def test_cm_column_decomposition_recursive(lib, oracle, n, m, i):
    """lu2tlib.c:1130 on the column-major awork of lu2.c:1072-1085: same
    pivots as the oracle's restatement, values to rounding (lu2's rule)."""
    rng = np.random.default_rng(n + m + i)
    rm = np.zeros((n, i + m + 3))
    rm[:, i:i + m] = rng.uniform(-0.5, 0.5, (n, m))
    cm = np.zeros((m, n))
    cm[:, :] = rm[:, i:i + m].T
    cm_in = cm.copy()
    pv = np.zeros(m, np.int32)
    lib.cm_column_decomposition_recursive(n, ptr(cm), m, ptr(pv), i)
    po, fo = oracle.cm_panel(rm, m, i, oracle.PIV_COMPAT)
    assert np.array_equal(pv, po)
    want = fo[i:, i:i + m]
    assert np.abs(cm[:, i:].T - want).max() <= 1e-12 * np.abs(want).max() * m
    assert np.array_equal(cm[:, :i], cm_in[:, :i])     # rows above i untouched


@pytest.mark.parametrize("pivot", ["gepp", "compat"])
def test_driver_lu2mi(oracle, tmp_path, pivot):
    """The lu2-style driver: CLI, the Emax / Nswap...Gflops / HPL lines and
    lu2's per-subsystem timer report.  Its Emax (lu2.c:45-60, evaluated on
    the device against the regenerated matrix) is checked against the
    oracle's Emax of the same x: two evaluations of max|Ax - b| differ by at
    most 2 gamma_n (|A||x| + |b|), i.e. 2 units of the HPL-2 scale."""
    import re

    n, nb = 2048, 256
    exe = os.path.join(ROOT, "lu2_amd", "lu2mi")
    xfile = tmp_path / "x.bin"
    r = subprocess.run([exe, "-n", str(n), "-b", str(nb), "-s", "1", "-p", pivot, "-x", str(xfile)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert f"N={n} Seed=1 NB={nb}" in out
    assert re.search(r"Nswap=\d+ cpsec =\s+\S+ wsec=\S+ \S+ Gflops", out), out
    if pivot == "gepp":
        assert "PASSED" in out
    for name in ("swaprows", "col dec t", "solve tri", "DGEMMrest", "back sub", "Total"):
        m = re.search(name + r"\s+time=\s*(\S+) ops/cycle=(\S+)", out)
        assert m and float(m.group(1)) > 0 and float(m.group(2)) > 0, (name, out)
    emax = float(re.search(r"Emax=\s*(\S+)", out).group(1))
    x = np.fromfile(xfile, dtype=np.float64)
    a0 = oracle.randomsetmat(n, 1)
    res = oracle.residuals(a0, x)
    unit = 2.0 ** -53 * (res[1] * res[3] + res[5]) * n
    assert abs(emax - oracle.emax(a0, x)) <= 2 * unit + 1e-3 * emax, (emax, oracle.emax(a0, x), unit)


DROPIN = os.path.join(ROOT, "oracle", "_ref", "lu2_dropin")


@pytest.mark.skipif(not os.path.exists(DROPIN),
                    reason="lu2's own driver is compiled from the reference sources (make -C oracle ref), "
                           "which were not available where this tree was built")
@pytest.mark.parametrize("pivot", ["gepp", "compat"])
def test_reference_driver_dropin(oracle, pivot):
    """lu2.c's own main (lu2.c:1220-1309) with lu2_mp_nompi.c, timerlib.c and
    second.c, compiled unmodified and linked against liblu2mi.so (no CBLAS):
    its panel leaves (cm_column_decomposition_recursive), TRSM and trailing
    products (mydgemm, matmul_for_small_nk) and reset_gdr all run through the
    library's seams on host memory.  With the GEPP leaf rule its Emax
    (lu2.c:45-60 on the regenerated matrix) is that of a backward-stable
    solve; with lu2's own rule it is within the spread that rule shows
    between two correct implementations."""
    import re

    n, nb = 2048, 256
    env = dict(os.environ, LU2MI_PIVOT=pivot, OMP_NUM_THREADS="8")
    r = subprocess.run([DROPIN, "-n", str(n), "-b", str(nb), "-s", "1"], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    out = r.stdout
    assert f"N={n} Seed=1 NB={nb}" in out
    m = re.search(r"Emax=\s*(\S+)", out)
    assert m, out
    emax = float(m.group(1))
    assert re.search(r"Nswap=\d+ cpsec =\s+\S+ wsec=\S+ \S+ Gflops", out), out
    for name in ("swaprows", "DGEMM2K", "Total"):
        assert name in out
    a0 = oracle.randomsetmat(n, 1)
    xo, _, _ = oracle.lumcolumn(a0, nb, oracle.PIV_GEPP if pivot == "gepp" else oracle.PIV_COMPAT)
    eo = oracle.emax(a0, xo)
    if pivot == "gepp":
        assert emax <= 1e-9 and emax <= 100 * eo, (emax, eo)
    else:
        assert np.isfinite(emax) and emax <= 1e4 * eo, (emax, eo)
