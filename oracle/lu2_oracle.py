"""ctypes bindings to the CPU oracle (oracle/_build/liblu2oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the lu2_amd package.  See
lu2_oracle.h for the functions and the reference lines they restate.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LU2O_LIB selects another build of the same source, e.g. the sanitizer
# build (make -C oracle asan: _build/liblu2oracle_asan.so, run with libasan
# preloaded)
LIB = os.environ.get("LU2O_LIB") or os.path.join(HERE, "_build", "liblu2oracle.so")
PIV_COMPAT = 0
PIV_GEPP = 1

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        c_int, c_long, c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
        L.lu2o_randomsetmat.argtypes = [c_int, c_int, _dp, c_long]
        L.lu2o_randomrows.argtypes = [c_int, c_int, c_long, c_int, _dp, c_long]
        L.lu2o_dgemm.argtypes = [c_int, c_int, c_int, c_double, _dp, c_long, _dp, c_long, c_double,
                                 _dp, c_long]
        L.lu2o_lumcolumn.argtypes = [c_int, _dp, c_long, c_int, c_int, _ip, _dp]
        L.lu2o_lumcolumn.restype = c_int
        L.lu2o_gepp_unblocked.argtypes = L.lu2o_lumcolumn.argtypes
        L.lu2o_gepp_unblocked.restype = c_int
        L.lu2o_backward_sub.argtypes = [c_int, _dp, c_long, _dp]
        L.lu2o_emax.argtypes = [c_int, _dp, c_long, _dp]
        L.lu2o_emax.restype = c_double
        L.lu2o_residuals.argtypes = [c_int, _dp, c_long, _dp, _dp]
        L.lu2o_factor_solve.argtypes = [c_int, _dp, c_long, c_int, _ip, _dp]
        L.lu2o_cm_panel.argtypes = [c_int, _dp, c_long, c_int, _ip, c_int, c_int]
        L.lu2o_cm_panel.restype = c_int
        L.lu2o_diag_dominant.argtypes = [c_int, _dp, c_long]
        L.lu2o_random_aug.argtypes = [c_int, c_int, c_long, _dp, c_long]
        L.lu2o_gepp_solve.argtypes = [c_int, c_int, _dp, c_long, _ip, _dp]
        L.lu2o_gepp_solve.restype = c_int
        L.lu2o_kat_identity.argtypes = [c_int]
        L.lu2o_kat_identity.restype = ctypes.c_long
        _lib = L
    return _lib


def randomsetmat(n: int, seed: int, lda: int | None = None) -> np.ndarray:
    """lu2.c:70-85: n x lda augmented matrix, b in column n, pad columns 0."""
    lda = n + 16 if lda is None else lda
    a = np.zeros((n, lda))
    lib().lu2o_randomsetmat(n, seed, a, lda)
    return a


def randomrows(n: int, seed: int, r0: int, nrows: int, lda: int) -> np.ndarray:
    """Rows [r0, r0+nrows) of the lu2.c:70-85 matrix."""
    a = np.zeros((nrows, lda))
    lib().lu2o_randomrows(n, seed, r0, nrows, a, lda)
    return a


def lumcolumn(a: np.ndarray, nb: int, pivmode: int):
    """Blocked recursive LU (lu2.c:1122).  Returns (x, piv, factored copy)."""
    f = np.array(a, dtype=np.float64, copy=True, order="C")
    n, lda = f.shape
    piv = np.zeros(n, np.int32)
    x = np.zeros(n)
    rc = lib().lu2o_lumcolumn(n, f, lda, nb, pivmode, piv, x)
    if rc:
        raise ValueError(f"lu2o_lumcolumn rc={rc}")
    return x, piv, f


def gepp_unblocked(a: np.ndarray, nb: int, pivmode: int):
    f = np.array(a, dtype=np.float64, copy=True, order="C")
    n, lda = f.shape
    piv = np.zeros(n, np.int32)
    x = np.zeros(n)
    rc = lib().lu2o_gepp_unblocked(n, f, lda, nb, pivmode, piv, x)
    if rc:
        raise ValueError(f"lu2o_gepp_unblocked rc={rc}")
    return x, piv, f


def dgemm(alpha, a, b, beta, c):
    c = np.array(c, dtype=np.float64, copy=True, order="C")
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m, k = a.shape
    n = b.shape[1]
    lib().lu2o_dgemm(m, n, k, alpha, a, max(k, 1), b, max(n, 1), beta, c, max(n, 1))
    return c


def backward_sub(f: np.ndarray) -> np.ndarray:
    n = f.shape[0]
    x = np.zeros(n)
    lib().lu2o_backward_sub(n, np.ascontiguousarray(f), f.shape[1], x)
    return x


def residuals(a0: np.ndarray, x: np.ndarray) -> np.ndarray:
    n, lda = a0.shape
    out = np.zeros(10)
    lib().lu2o_residuals(n, np.ascontiguousarray(a0), lda, np.ascontiguousarray(x), out)
    return out


def emax(a0: np.ndarray, x: np.ndarray) -> float:
    n, lda = a0.shape
    return lib().lu2o_emax(n, np.ascontiguousarray(a0), lda, np.ascontiguousarray(x))


def factor_solve(f: np.ndarray, nb: int, piv: np.ndarray, r: np.ndarray) -> np.ndarray:
    n, lda = f.shape
    out = np.array(r, dtype=np.float64, copy=True)
    lib().lu2o_factor_solve(n, np.ascontiguousarray(f), lda, nb, np.ascontiguousarray(piv, np.int32), out)
    return out


def cm_panel(a_rm: np.ndarray, m: int, i: int, pivmode: int):
    """lu2tlib.c:1130 on columns [i, i+m) of a row-major n x lda matrix."""
    f = np.array(a_rm, dtype=np.float64, copy=True, order="C")
    n, lda = f.shape
    pv = np.zeros(m, np.int32)
    rc = lib().lu2o_cm_panel(n, f, lda, m, pv, i, pivmode)
    if rc:
        raise ValueError(f"lu2o_cm_panel rc={rc}")
    return pv, f


def kat_identity(nmat: int) -> int:
    return lib().lu2o_kat_identity(nmat)


def random_aug(n: int, nrhs: int, seed: int, lda: int | None = None) -> np.ndarray:
    """n x lda augmented [A | B] with nrhs right-hand sides, drawn row by row
    after srand48(seed) (lu2.c:70-85's rule); padding columns 0."""
    lda = n + nrhs if lda is None else lda
    a = np.empty((n, lda))
    lib().lu2o_random_aug(n, n + nrhs, seed, a, lda)
    return a


def gepp_solve(a_aug: np.ndarray, nrhs: int):
    """Textbook GEPP + back substitution on [A | B]; returns (piv, x[n][nrhs],
    factored copy)."""
    f = np.ascontiguousarray(a_aug, dtype=np.float64).copy()
    n = f.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    x = np.zeros(n * nrhs)
    rc = lib().lu2o_gepp_solve(n, nrhs, f, f.shape[1], piv, x)
    if rc:
        raise ValueError(f"lu2o_gepp_solve: {rc}")
    return piv, x.reshape(n, nrhs), f


def hpl2_multi(a_aug: np.ndarray, x: np.ndarray) -> float:
    """Max over right-hand sides of the HPL-2 scaled residual
    ||Ax-b||_oo / (eps (||A||_oo ||x||_oo + ||b||_oo) n), eps = 2^-53."""
    n = a_aug.shape[0]
    nrhs = x.shape[1]
    A = a_aug[:, :n]
    B = a_aug[:, n:n + nrhs]
    r = B - A @ x
    an = np.abs(A).sum(axis=1).max()
    eps = 2.0 ** -53
    res = np.abs(r).max(axis=0) / (eps * (an * np.abs(x).max(axis=0) + np.abs(B).max(axis=0)) * n)
    return float(res.max())


def diag_dominant(a: np.ndarray) -> np.ndarray:
    """Copy of the augmented matrix with a[i][i] += sum_{j<n} |a[i][j]|."""
    f = np.array(a, dtype=np.float64, copy=True, order="C")
    lib().lu2o_diag_dominant(f.shape[0], f, f.shape[1])
    return f
