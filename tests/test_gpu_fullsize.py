"""The device path at the configured sizes.

* n = 32768 and n = 34816 (nb = 512) against the CPU oracle's own results
  (tests/golden/oracle_*.npz, written by make_oracle_large.py; the GEPP pivot
  sequences there were also matched against LAPACK dgetrf): at n = 32768 this
  is the drop-in configuration's bar (pivots identical, x within 1e-10 of
  ||x||); at n = 34816 the first two panels have more than 32768 rows, where
  the leaf holds four rows per thread.
* The device residual (lu2.c:45-60 Emax, lu2_mpi.c:2170-2229 resid1..3 and
  the HPL-2 form) against the oracle's residuals on the same x.
* n = 65536 (FP64 at nb = 1024 and both mixed-precision forms) and n = 98304 (FP16 and
  BF16 updates on the diagonally dominant matrix) through properties that do
  not depend on size: HPL-2 scaled residual, a valid pivot sequence, the
  mixed solutions against the FP64 one.
"""
import numpy as np
import pytest

import lu2_amd as L
from conftest import golden

pytestmark = pytest.mark.gpu


def device_solve(ctx, n, nb, mode, seed=1, want_residual=True):
    """lu2's matrix generated on the device, factored and solved there."""
    lda = n + 16
    dA0 = ctx.alloc(8 * n * lda) if want_residual else None
    dA = ctx.alloc(8 * n * lda)
    dx = ctx.alloc(8 * n)
    dp = ctx.alloc(4 * n)
    try:
        ctx.randomsetmat(n, seed, dA, lda)
        if dA0 is not None:
            dA0.copy_from(dA)
        ctx.set_option(L.OPT_PIVMODE, mode)
        ctx.lumcolumn_dev(n, dA, lda, nb, dp, dx)
        x = dx.download(np.float64, n)
        piv = dp.download(np.int32, n)
        res = ctx.residual_dev(n, dA0, lda, dx) if dA0 is not None else None
    finally:
        for d in (dA0, dA, dx, dp):
            if d is not None:
                d.free()
    return x, piv, res


def forward_bound(n, hpl2_a, hpl2_b):
    """How far apart two backward-stable solutions of lu2's matrix may be.
    x1 - x2 = A^-1 (r2 - r1), so ||x1 - x2|| / ||x|| <= kappa (eta1 + eta2)
    with eta = ||r|| / (||A|| ||x||) ~ HPL-2 * n * eps the normwise backward
    errors.  For an n x n matrix of i.i.d. entries kappa_2 is about
    sigma_max / sigma_min ~ (2 sqrt(n) s) / (s / sqrt(n)) = 2n; the bound
    takes 10n.  At n = 65536 with HPL-2 ~ 2e-3 this is ~2e-8 (two correct
    FP64 factorizations of this matrix measured ~2e-9 apart)."""
    eps = 2.0 ** -53
    return 10.0 * n * (hpl2_a + hpl2_b) * n * eps


def valid_pivots(piv):
    n = len(piv)
    return bool(np.all(piv >= np.arange(n)) and np.all(piv < n))


# ------------------------------------------------------ against the oracle
@pytest.mark.parametrize("n,nb,mode", [(32768, 512, "gepp"), (34816, 512, "gepp"), (34816, 512, "compat")])
def test_lu_large_against_oracle(ctx, n, nb, mode):
    """GEPP: the whole pivot sequence identical, x within 1e-10.  lu2's own
    rule (COMPAT) grows the entries of this matrix (the oracle's own HPL-2
    residual is 3.5e3 here), so the last-bit differences between two correct
    evaluation orders grow with it until two candidates of some column tie to
    within them; from there the sequences may part.  COMPAT is held to
    identical pivots through column 16384 -- every panel with more than
    32768 rows (the four-rows-per-thread leaf) and the next 15 blocks -- and
    to at most one column in a thousand differing after that."""
    g = golden(f"oracle_n{n}_nb{nb}_{mode}.npz")
    pm = L.PIV_GEPP if mode == "gepp" else L.PIV_COMPAT
    x, piv, res = device_solve(ctx, n, nb, pm, want_residual=(mode == "gepp"))
    mism = np.flatnonzero(piv != g["piv"])
    if mode == "gepp":
        assert mism.size == 0, f"{mism.size} pivots differ, first at column {mism[:5]}"
        assert res[6] < 16.0
        xo = g["x"]
        err = np.abs(x - xo).max() / np.abs(xo).max()
        if n <= 32768:
            assert err <= 1e-10      # the drop-in configuration's bar (n = 32768)
        else:
            # the oracle's own HPL-2 residual for this case (make_oracle_large.py)
            assert err <= forward_bound(n, res[6], 4.46e-3), err
    else:
        assert mism.size == 0 or mism[0] >= 16384, f"pivots differ from column {mism[:5]}"
        assert mism.size <= n // 1000, f"{mism.size} pivots differ"


@pytest.mark.parametrize("n,seed,perturb", [(2048, 1, 0.0), (3001, 7, 0.0), (2048, 12345, 1e-6)])
def test_residual_dev_matches_oracle(ctx, oracle, n, seed, perturb):
    """lu2mi_residual_dev against the oracle's residuals on the same x.  The
    norms of A, x and b are sums and maxima of the same numbers (agree to
    the rounding of an n-term sum); max|Ax-b| is a difference of two
    evaluations of A x, each within gamma_n (|A||x| + |b|) of the exact
    value, i.e. within 1 in HPL-2 units, so the two agree within 2 units.
    With a perturbed x the residual is large and must agree to 1e-6 relative
    (the evaluation error is ~1e-10 of it); the other forms follow from the
    same components."""
    lda = n + 16
    a0 = oracle.randomsetmat(n, seed)
    x = np.linalg.solve(a0[:, :n], a0[:, n])
    if perturb:
        x = x + perturb * np.random.default_rng(seed).standard_normal(n)
    dA, dx = ctx.alloc(8 * n * lda), ctx.alloc(8 * n)
    try:
        dA.upload(a0)
        dx.upload(x)
        got = ctx.residual_dev(n, dA, lda, dx)
    finally:
        dA.free()
        dx.free()
    want = oracle.residuals(a0, x)
    np.testing.assert_allclose(got[1:6], want[1:6], rtol=n * 2.0 ** -52)
    assert abs(got[6] - want[6]) <= 2.0, (got[6], want[6])
    if perturb:
        # the residual (~4e-5) dwarfs the evaluation error (<= ~1e-10), so a
        # wrong index or a dropped term would show at O(1)
        np.testing.assert_allclose(got[0], want[0], rtol=1e-6)
        np.testing.assert_allclose(got[6:], want[6:], rtol=1e-6)
    # each scaled form is its definition on the device's own components
    eps, eps_lu2 = 2.0 ** -53, 2e-16
    np.testing.assert_allclose(got[6], got[0] / (eps * (got[1] * got[3] + got[5]) * n), rtol=1e-14)
    np.testing.assert_allclose(got[7], got[0] / (eps_lu2 * got[2] * n), rtol=1e-14)
    np.testing.assert_allclose(got[8], got[0] / (eps_lu2 * got[2] * got[4]), rtol=1e-14)
    np.testing.assert_allclose(got[9], got[0] / (eps_lu2 * got[1] * got[3] * n), rtol=1e-14)


def test_residual_dev_keeps_nan(ctx, oracle):
    """A solution with a NaN must not pass as a small residual: the maxima
    keep NaN (fmax would drop it)."""
    n = 1000
    lda = n + 16
    a0 = oracle.randomsetmat(n, 2)
    x = np.linalg.solve(a0[:, :n], a0[:, n])
    x[17] = np.nan
    dA, dx = ctx.alloc(8 * n * lda), ctx.alloc(8 * n)
    try:
        dA.upload(a0)
        dx.upload(x)
        got = ctx.residual_dev(n, dA, lda, dx)
    finally:
        dA.free()
        dx.free()
    assert np.isnan(got[0]) and np.isnan(got[3]) and np.isnan(got[6])


# ------------------------------------------------- full-size properties
N3, NB3 = 65536, 1024
# block size of each mixed form at n=65536 (as bench.py runs them)
NB_MIXED = {L.PREC_FP16X2: 2048, L.PREC_FP32: 1024}


@pytest.fixture(scope="module")
def fp64_n65536(ctx):
    """The metric's configuration: FP64 LU + solve of lu2's matrix, n=65536,
    nb=1024, GEPP; the pristine matrix stays resident for the mixed solves."""
    n, lda = N3, N3 + 16
    dA0 = ctx.alloc(8 * n * lda)
    dA = ctx.alloc(8 * n * lda)
    dx = ctx.alloc(8 * n)
    dp = ctx.alloc(4 * n)
    ctx.randomsetmat(n, 1, dA0, lda)
    dA.copy_from(dA0)
    ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
    ctx.lumcolumn_dev(n, dA, lda, NB3, dp, dx)
    out = {"dA0": dA0, "x": dx.download(np.float64, n), "piv": dp.download(np.int32, n),
           "res": ctx.residual_dev(n, dA0, lda, dx)}
    for d in (dA, dx, dp):
        d.free()
    yield out
    dA0.free()


def test_fp64_n65536_properties(fp64_n65536):
    r = fp64_n65536
    assert r["res"][6] < 16.0
    assert valid_pivots(r["piv"])
    assert np.all(np.isfinite(r["x"]))


@pytest.mark.parametrize("prec", [L.PREC_FP16X2, L.PREC_FP32])
def test_mixed_n65536(ctx, fp64_n65536, prec):
    n, lda = N3, N3 + 16
    dx = ctx.alloc(8 * n)
    try:
        ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
        info = ctx.solve_mixed(n, fp64_n65536["dA0"], lda, NB_MIXED[prec], dx, prec, maxiter=30, target=1.0)
        x = dx.download(np.float64, n)
        res = ctx.residual_dev(n, fp64_n65536["dA0"], lda, dx)
    finally:
        dx.free()
    assert info["converged"] and info["iterations"] >= 1
    assert res[6] <= 1.0
    x64 = fp64_n65536["x"]
    err = np.abs(x - x64).max() / np.abs(x64).max()
    assert err <= forward_bound(N3, res[6], fp64_n65536["res"][6]), err


N4, NB4 = 98304, 2048


@pytest.fixture(scope="module")
def dd_n98304(ctx):
    """The 16-bit configuration's matrix: lu2's matrix with a_ii += sum_j |a_ij|."""
    n, lda = N4, N4 + 16
    dA0 = ctx.alloc(8 * n * lda)
    ctx.randomsetmat(n, 1, dA0, lda)
    ctx.make_diag_dominant(n, dA0, lda)
    yield dA0
    dA0.free()


@pytest.mark.parametrize("prec", [L.PREC_FP16, L.PREC_BF16])
def test_lowp_n98304_diag_dominant(ctx, dd_n98304, prec):
    n, lda = N4, N4 + 16
    dx = ctx.alloc(8 * n)
    try:
        ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
        info = ctx.solve_mixed(n, dd_n98304, lda, NB4, dx, prec, maxiter=30, target=1.0)
        res = ctx.residual_dev(n, dd_n98304, lda, dx)
        x = dx.download(np.float64, n)
    finally:
        dx.free()
    assert info["converged"]
    assert res[6] <= 1.0
    assert np.all(np.isfinite(x))
