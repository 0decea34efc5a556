"""Mixed precision: FP32 MFMA GEMM and the FP32-factor + FP64 refinement solve.

The FP32 factorization may choose different pivots than the FP64 one (its
rounding differs), so the bar is the solution: after refinement the HPL-2
scaled residual meets the target and x agrees with the FP64 oracle's x within
1e-10 relative to ||x|| (gepp)."""
import numpy as np
import pytest

import lu2_amd as L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,k,alpha,beta", [(128, 128, 16, -1.0, 1.0), (300, 257, 100, -1.0, 1.0),
                                              (1000, 999, 512, 1.0, 0.0), (513, 130, 33, 0.75, -0.5)])
def test_sgemm_error_bound(ctx, m, n, k, alpha, beta):
    rng = np.random.default_rng(m + n + k)
    a = rng.standard_normal((m, k)).astype(np.float32)
    b = rng.standard_normal((k, n)).astype(np.float32)
    c = rng.standard_normal((m, n)).astype(np.float32)
    dA, dB, dC = ctx.alloc(a.nbytes), ctx.alloc(b.nbytes), ctx.alloc(c.nbytes)
    dA.upload(a)
    dB.upload(b)
    dC.upload(c)
    ctx.sgemm_dev(m, n, k, alpha, dA, k, dB, n, beta, dC, n)
    got = dC.download(np.float32, m * n).reshape(m, n).astype(np.float64)
    for d in (dA, dB, dC):
        d.free()
    a64, b64, c64 = a.astype(np.float64), b.astype(np.float64), c.astype(np.float64)
    ref = alpha * (a64 @ b64) + beta * c64
    u = 2.0 ** -24
    gamma = (k + 2) * u / (1 - (k + 2) * u)
    bound = 2 * gamma * (abs(alpha) * (np.abs(a64) @ np.abs(b64)) + abs(beta) * np.abs(c64))
    assert np.all(np.abs(got - ref) <= bound + 1e-30)


@pytest.mark.parametrize("n,nb,seed", [(512, 64, 1), (2048, 256, 12345), (4096, 512, 1)])
def test_mixed_fp32_refinement(ctx, oracle, n, nb, seed):
    lda = n + 16
    a0 = oracle.randomsetmat(n, seed)
    dA = ctx.alloc(8 * n * lda)
    dx = ctx.alloc(8 * n)
    try:
        dA.upload(a0)
        ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
        info = ctx.solve_mixed(n, dA, lda, nb, dx, L.PREC_FP32, maxiter=30, target=1.0)
        x = dx.download(np.float64, n)
        after = dA.download(np.float64, n * lda).reshape(n, lda)
    finally:
        dA.free()
        dx.free()
    assert np.array_equal(after, a0)                      # the FP64 matrix is untouched
    assert info["iterations"] >= 1
    assert info["resid_hpl2"] <= 1.0
    assert oracle.residuals(a0, x)[6] <= 2.0
    xo, _, _ = oracle.lumcolumn(a0, nb, oracle.PIV_GEPP)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()


def strailing_product(ctx, a, b, c):
    """C - A B in FP32 through the kernel of the FP32 factorization's trailing
    update (leading dimensions a multiple of 4 floats, 16-byte rows, as the
    factorization's)."""
    m, k = a.shape
    n = b.shape[1]
    lda, ldb = (k + 3) // 4 * 4, (n + 3) // 4 * 4
    ap = np.zeros((m, lda), np.float32); ap[:, :k] = a
    bp = np.zeros((k, ldb), np.float32); bp[:, :n] = b
    cp = np.zeros((m, ldb), np.float32); cp[:, :n] = c
    dA, dB, dC = ctx.alloc(ap.nbytes), ctx.alloc(bp.nbytes), ctx.alloc(cp.nbytes)
    dA.upload(ap)
    dB.upload(bp)
    dC.upload(cp)
    ctx.set_option(L.OPT_DGEMM_TRAILING, 1)
    try:
        ctx.sgemm_dev(m, n, k, -1.0, dA, lda, dB, ldb, 1.0, dC, ldb)
    finally:
        ctx.set_option(L.OPT_DGEMM_TRAILING, 0)
    out = dC.download(np.float32, m * ldb).reshape(m, ldb)
    for d in (dA, dB, dC):
        d.free()
    assert np.array_equal(out[:, n:], cp[:, n:])      # nothing written past column n
    return out[:, :n].astype(np.float64)


@pytest.mark.parametrize("m,n,k", [(128, 128, 32), (300, 257, 512), (1000, 999, 1024), (256, 384, 64),
                                   (129, 130, 33)])
def test_strailing_gemm_error_bound(ctx, m, n, k):
    """The FP32 trailing-update kernel against the componentwise bound
    |C - fl(C - AB)| <= gamma_{K+2} (|A||B| + |C|) in FP32 (u = 2^-24), edge
    tiles (M, N not multiples of 128, N not a multiple of 4, K not a multiple
    of 32) included."""
    rng = np.random.default_rng(m * 31 + n * 7 + k)
    a = rng.standard_normal((m, k)).astype(np.float32)
    b = rng.standard_normal((k, n)).astype(np.float32)
    c = rng.standard_normal((m, n)).astype(np.float32)
    got = strailing_product(ctx, a, b, c)
    a64, b64, c64 = a.astype(np.float64), b.astype(np.float64), c.astype(np.float64)
    ref = c64 - a64 @ b64
    u = 2.0 ** -24
    gamma = (k + 2) * u / (1 - (k + 2) * u)
    bound = 2 * gamma * (np.abs(a64) @ np.abs(b64) + np.abs(c64))
    assert np.all(np.abs(got - ref) <= bound + 1e-30)


def test_strailing_gemm_edge_tiles_bit_identical(ctx):
    """FP32 edge tiles (bounds-checked loads) perform exactly the arithmetic
    of interior tiles (unchecked 16-byte loads): a 300 x 257 product equals
    the corner of a 512 x 512 one, bit for bit."""
    rng = np.random.default_rng(13)
    k = 96
    a = rng.standard_normal((512, k)).astype(np.float32)
    b = rng.standard_normal((k, 512)).astype(np.float32)
    c = rng.standard_normal((512, 512)).astype(np.float32)
    big = strailing_product(ctx, a, b, c)
    small = strailing_product(ctx, a[:300], b[:, :257], c[:300, :257])
    assert np.array_equal(big[:300, :257], small)
