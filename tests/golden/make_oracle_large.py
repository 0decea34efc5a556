"""Write the large-size oracle fixtures in tests/golden/: the CPU oracle's
blocked recursive LU (lu2.c:1122 lumcolumn restated in oracle/lu2_oracle.c)
on lu2's drand48 matrix at sizes where the oracle takes minutes, so that the
GPU tests can hold the device path to it without re-running it.

  oracle_n{n}_nb{nb}_{gepp|compat}.npz   piv (global pivot rows, int32) and x

For the GEPP cases the pivot sequence is also checked here against LAPACK
dgetrf (scipy.linalg.lu_factor) on the same matrix: standard partial
pivoting chooses the same rows whatever the blocking, so the fixture does
not rest on the oracle alone.

    python tests/golden/make_oracle_large.py      # about 15 minutes on 8 cores
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import lu2_oracle as O  # noqa: E402

CASES = [(32768, 512, "gepp"), (34816, 512, "gepp"), (34816, 512, "compat")]


def main() -> None:
    for n, nb, mode in CASES:
        out = os.path.join(HERE, f"oracle_n{n}_nb{nb}_{mode}.npz")
        a = O.randomsetmat(n, 1)
        t = time.time()
        x, piv, f = O.lumcolumn(a, nb, O.PIV_GEPP if mode == "gepp" else O.PIV_COMPAT)
        del f
        print(f"oracle n={n} nb={nb} {mode}: {time.time() - t:.1f} s, HPL-2 {O.residuals(a, x)[6]:.3g}",
              flush=True)
        if mode == "gepp":
            lu, ipiv = sl.lu_factor(a[:, :n], overwrite_a=True, check_finite=False)
            del lu
            mism = int((ipiv.astype(np.int32) != piv).sum())
            print(f"  LAPACK dgetrf pivot mismatches: {mism}", flush=True)
            assert mism == 0, "oracle GEPP pivots differ from LAPACK"
        del a
        np.savez_compressed(out, piv=piv, x=x)
        print("  wrote", out, os.path.getsize(out), "bytes", flush=True)


if __name__ == "__main__":
    main()
