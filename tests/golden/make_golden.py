"""Write the golden fixtures in tests/golden/ from sources independent of the
CPU oracle:

  drand48.npz    glibc drand48 after srand48(seed) (ctypes into libc), the
                 generator of lu2.c:70-85; first 64 draws and the draws at
                 a few far indices, seeds 1 and 12345.
  lapack_*.npz   the lu2 matrix built from glibc drand48, factored by LAPACK
                 dgetrf (scipy.linalg.lu_factor): standard partial-pivoting
                 pivot sequence and the solution x (lu_solve).
  compat_*.npz   pivots of a plain numpy elimination that uses lu2's leaf
                 rule (odd column j pivots on column j-1, lu2tlib.c:42-52 via
                 lu2tlib.c:1050); the rule alone, no blocking.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
libc = ctypes.CDLL(None)
libc.drand48.restype = ctypes.c_double
libc.srand48.argtypes = [ctypes.c_long]


def glibc_matrix(n: int, seed: int) -> np.ndarray:
    """n x (n+1) augmented matrix in lu2's draw order (b last in each row)."""
    libc.srand48(seed)
    a = np.empty((n, n + 1))
    for r in range(n):
        for c in range(n + 1):
            a[r, c] = libc.drand48() - 0.5
    return a


def compat_pivots(a_aug: np.ndarray) -> np.ndarray:
    """Unblocked elimination with lu2's leaf pivot rule (pivots only)."""
    a = np.array(a_aug[:, :-1], dtype=np.float64)
    n = a.shape[0]
    piv = np.zeros(n, np.int32)
    for j in range(n):
        col = j - 1 if (j & 1) else j
        p = j + int(np.argmax(np.abs(a[j:, col])))   # first maximum
        piv[j] = p
        if p != j:
            a[[j, p], :] = a[[p, j], :]
        d = a[j, j]
        u = a[j, j + 1:] / d
        a[j + 1:, j + 1:] -= np.outer(a[j + 1:, j], u)
    return piv


def main() -> None:
    draws = {}
    for seed in (1, 12345):
        libc.srand48(seed)
        seq = np.array([libc.drand48() for _ in range(200_001)])
        draws[f"seed{seed}_first"] = seq[:64]
        draws[f"seed{seed}_idx"] = np.array([0, 1, 1000, 65536, 200_000])
        draws[f"seed{seed}_val"] = seq[[0, 1, 1000, 65536, 200_000]]
    np.savez(os.path.join(HERE, "drand48.npz"), **draws)

    for n, seed in ((256, 1), (512, 12345), (1024, 1)):
        a = glibc_matrix(n, seed)
        lu, ipiv = sl.lu_factor(a[:, :n])
        x = sl.lu_solve((lu, ipiv), a[:, n])
        np.savez(os.path.join(HERE, f"lapack_n{n}_s{seed}.npz"), piv=ipiv.astype(np.int32), x=x)
    for n, seed in ((64, 1), (256, 12345)):
        a = glibc_matrix(n, seed)
        np.savez(os.path.join(HERE, f"compat_n{n}_s{seed}.npz"), piv=compat_pivots(a))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
