"""Pins of the CPU oracle (oracle/lu2_oracle.c) -- CPU only.

The reference lu2 needs an external CBLAS that this image lacks, so it is not
built; its only known-answer test is gdrdgemm.c:17-58 (I*B == B).  The
oracle is pinned against glibc's drand48, LAPACK dgetrf (SciPy), a plain
numpy elimination with lu2's leaf rule, the textbook unblocked elimination,
and closed forms.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sl

from conftest import golden


def test_drand48_matches_glibc(oracle):
    libc = ctypes.CDLL(None)
    libc.drand48.restype = ctypes.c_double
    libc.srand48.argtypes = [ctypes.c_long]
    for seed in (1, 12345, 7):
        n = 37
        a = oracle.randomsetmat(n, seed, n + 5)
        libc.srand48(seed)
        want = np.array([[libc.drand48() - 0.5 for _ in range(n + 1)] for _ in range(n)])
        assert np.array_equal(a[:, : n + 1], want)
        assert np.all(a[:, n + 1:] == 0.0)


def test_drand48_golden_far_indices(oracle):
    g = golden("drand48.npz")
    for seed in (1, 12345):
        st = ctypes.c_ulonglong()
        L = oracle.lib()
        L.lu2o_srand48.argtypes = [ctypes.c_long, ctypes.POINTER(ctypes.c_ulonglong)]
        L.lu2o_drand48.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
        L.lu2o_drand48.restype = ctypes.c_double
        L.lu2o_srand48(seed, ctypes.byref(st))
        seq = np.array([L.lu2o_drand48(ctypes.byref(st)) for _ in range(200_001)])
        assert np.array_equal(seq[:64], g[f"seed{seed}_first"])
        assert np.array_equal(seq[g[f"seed{seed}_idx"]], g[f"seed{seed}_val"])


@pytest.mark.parametrize("n,seed", [(256, 1), (512, 12345), (1024, 1)])
@pytest.mark.parametrize("nb", [64, 128])
def test_gepp_matches_lapack_fixture(oracle, n, seed, nb):
    g = golden(f"lapack_n{n}_s{seed}.npz")
    a0 = oracle.randomsetmat(n, seed)
    x, piv, _ = oracle.lumcolumn(a0, nb, oracle.PIV_GEPP)
    assert np.array_equal(piv, g["piv"])
    assert np.abs(x - g["x"]).max() <= 1e-10 * np.abs(g["x"]).max()


@pytest.mark.parametrize("n,seed", [(64, 1), (256, 12345)])
def test_compat_matches_numpy_rule(oracle, n, seed):
    g = golden(f"compat_n{n}_s{seed}.npz")
    a0 = oracle.randomsetmat(n, seed)
    for nb in (16, 32, 64):
        _, piv, _ = oracle.lumcolumn(a0, nb, oracle.PIV_COMPAT)
        assert np.array_equal(piv, g["piv"])


def test_pins_detect_wrong_rule(oracle):
    """A planted mistake -- the other pivot rule -- is caught by the pins."""
    a0 = oracle.randomsetmat(256, 1)
    _, pc, _ = oracle.lumcolumn(a0, 64, oracle.PIV_COMPAT)
    assert not np.array_equal(pc, golden("lapack_n256_s1.npz")["piv"])
    a1 = oracle.randomsetmat(256, 12345)
    _, pg, _ = oracle.lumcolumn(a1, 64, oracle.PIV_GEPP)
    assert not np.array_equal(pg, golden("compat_n256_s12345.npz")["piv"])


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n,nb", [(128, 32), (256, 64), (512, 128), (1024, 256)])
def test_blocked_equals_unblocked(oracle, mode, n, nb):
    """lu2's blocked recursion and the textbook elimination agree: same pivots
    and (gepp) the same stored factors up to rounding."""
    a0 = oracle.randomsetmat(n, 3)
    x1, p1, f1 = oracle.lumcolumn(a0, nb, mode)
    x2, p2, f2 = oracle.gepp_unblocked(a0, nb, mode)
    assert np.array_equal(p1, p2)
    assert np.all(p1 >= np.arange(n)) and np.all(p1 < n)
    if mode == oracle.PIV_GEPP:
        scale = np.abs(f2[:, : n + 1]).max()
        assert np.abs(f1[:, : n + 1] - f2[:, : n + 1]).max() <= 1e-12 * n * scale
        assert np.abs(x1 - x2).max() <= 1e-11 * np.abs(x2).max()


def test_hpl_residual_gepp(oracle):
    n = 1024
    a0 = oracle.randomsetmat(n, 1)
    x, _, _ = oracle.lumcolumn(a0, 128, oracle.PIV_GEPP)
    r = oracle.residuals(a0, x)
    assert r[6] < 16.0
    # the residual functional against a direct numpy evaluation
    A, b = a0[:, :n], a0[:, n]
    assert np.isclose(r[1], np.abs(A).sum(1).max(), rtol=1e-12)
    assert np.isclose(r[2], np.abs(A).sum(0).max(), rtol=1e-12)
    assert r[3] == np.abs(x).max() and r[5] == np.abs(b).max()
    assert np.isclose(r[6], r[0] / (2.0 ** -53 * (r[1] * r[3] + r[5]) * n), rtol=1e-14)
    # ||Ax-b|| is itself a rounding-level quantity: another summation order
    # agrees only in magnitude
    rinf = np.abs(A @ x - b).max()
    assert 0.5 < r[0] / rinf < 2.0
    assert oracle.emax(a0, x) == r[0]


def test_kat_identity_2048(oracle):
    """gdrdgemm.c:17-58: I(2048^2) * B == B exactly, B[i][j] = 2048 i + j."""
    assert oracle.kat_identity(2048) == 0


def test_dgemm_bound(oracle):
    rng = np.random.default_rng(0)
    for m, n, k in [(7, 5, 3), (130, 257, 129), (64, 1, 512)]:
        a, b, c = rng.standard_normal((m, k)), rng.standard_normal((k, n)), rng.standard_normal((m, n))
        got = oracle.dgemm(-1.0, a, b, 1.0, c)
        ref = c - a @ b
        gamma = k * 2.0 ** -53 / (1 - k * 2.0 ** -53)
        assert np.all(np.abs(got - ref) <= 2 * gamma * (np.abs(a) @ np.abs(b) + np.abs(c)) + 1e-300)
        got2 = oracle.dgemm(0.5, a, b, 0.0, np.full((m, n), np.nan))
        assert np.allclose(got2, 0.5 * a @ b)


def test_backward_sub_closed_form(oracle):
    rng = np.random.default_rng(1)
    n = 200
    f = np.zeros((n, n + 16))
    U = np.triu(rng.uniform(-0.5, 0.5, (n, n)), 1)
    f[:, :n] = U + np.diag(rng.uniform(1, 2, n))  # diagonal holds d, unused
    x_true = rng.standard_normal(n)
    f[:, n] = (U + np.eye(n)) @ x_true
    x = oracle.backward_sub(f)
    assert np.abs(x - x_true).max() < 1e-12 * np.abs(x_true).max() * n


@pytest.mark.parametrize("mode", [0, 1])
def test_factor_solve_new_rhs(oracle, mode):
    """Re-solving with lu2's stored factors (block pivots, D^-1 L11, L21,
    unit-upper U) reproduces the solve for a new right-hand side."""
    n, nb = 512, 64
    a0 = oracle.randomsetmat(n, 5)
    x, piv, f = oracle.lumcolumn(a0, nb, mode)
    # the original b through the stored factors gives back x
    xb = oracle.factor_solve(f, nb, piv, a0[:, n])
    assert np.abs(xb - x).max() <= 1e-9 * np.abs(x).max()
    rhs = np.random.default_rng(2).standard_normal(n)
    y = oracle.factor_solve(f, nb, piv, rhs)
    ref = sl.solve(a0[:, :n], rhs)
    tol = 1e-9 if mode == oracle.PIV_GEPP else 1e-4
    assert np.abs(y - ref).max() <= tol * np.abs(ref).max()


def test_reference_limits_rejected(oracle):
    a0 = oracle.randomsetmat(96, 1)
    with pytest.raises(ValueError):
        oracle.lumcolumn(a0, 48, 1)      # nb not a power of two
    with pytest.raises(ValueError):
        oracle.lumcolumn(a0, 64, 1)      # n % nb != 0


# ---- batched systems: the per-system generator and GEPP solve -------------

def test_random_aug_is_randomsetmat_rule(oracle):
    """With one right-hand side the general generator is lu2's randomsetmat
    (itself pinned against glibc's drand48 draws)."""
    for n, seed in [(1, 5), (7, 1), (64, 12345)]:
        a = oracle.random_aug(n, 1, seed, lda=n + 4)
        assert np.array_equal(a[:, :n + 1], oracle.randomsetmat(n, seed)[:, :n + 1])
        assert not a[:, n + 1:].any()
    # more right-hand sides continue the same row-order stream
    a3 = oracle.random_aug(4, 3, 9)           # 4 x 7 draws
    flat = oracle.random_aug(1, 27, 9)[0]     # 28 draws in one row
    assert np.array_equal(a3.ravel(), flat)


@pytest.mark.parametrize("n,nrhs,seed", [(1, 1, 1), (2, 3, 2), (37, 4, 3), (200, 2, 4), (512, 8, 5)])
def test_gepp_solve_matches_lapack(oracle, n, nrhs, seed):
    import scipy.linalg as sl

    a = oracle.random_aug(n, nrhs, seed)
    piv, x, _ = oracle.gepp_solve(a, nrhs)
    lu, ip = sl.lu_factor(a[:, :n])
    assert np.array_equal(piv, ip)
    xr = sl.lu_solve((lu, ip), a[:, n:])
    assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-11
    assert oracle.hpl2_multi(a, x) < 1.0


def test_gepp_solve_agrees_with_lu2_form(oracle):
    """Standard GEPP pivots do not depend on lu2's stored form or blocking:
    with one right-hand side the pivot sequence equals the lu2-form GEPP."""
    n = 256
    a = oracle.randomsetmat(n, 77)
    piv, x, _ = oracle.gepp_solve(a[:, :n + 1], 1)
    x2, p2, _ = oracle.gepp_unblocked(a, 64, oracle.PIV_GEPP)
    assert np.array_equal(piv, p2)
    assert np.abs(x[:, 0] - x2).max() / np.abs(x2).max() < 1e-12


def test_gepp_solve_closed_forms(oracle):
    # 2x2 with a pivot swap: [[1, 2 | 5], [3, 4 | 6]] -> x = (-4, 4.5)
    a = np.array([[1.0, 2.0, 5.0], [3.0, 4.0, 6.0]])
    piv, x, f = oracle.gepp_solve(a, 1)
    assert list(piv) == [1, 1]
    assert np.allclose(x[:, 0], [-4.0, 4.5], rtol=0, atol=1e-14)   # 1/3 rounds
    assert f[1, 0] == 1.0 / 3.0   # multiplier a*(1/d)
    # ties go to the first row; a NaN never wins
    a = np.array([[1.0, 0.0, 1.0], [-1.0, 1.0, 0.0]])
    piv, _, _ = oracle.gepp_solve(a, 1)
    assert piv[0] == 0
    a = np.hstack([np.array([[np.nan, 1.0, 0.0], [0.5, 0.0, 1.0], [0.25, 1.0, 1.0]]), np.ones((3, 1))])
    piv, _, _ = oracle.gepp_solve(a, 1)
    assert piv[0] == 1


def test_hpl2_multi_detects_planted_error(oracle):
    a = oracle.random_aug(100, 2, 8)
    _, x, _ = oracle.gepp_solve(a, 2)
    assert oracle.hpl2_multi(a, x) < 1.0
    x[3, 1] += 1e-9
    assert oracle.hpl2_multi(a, x) > 16
