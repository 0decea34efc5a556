"""GPU parity: the HIP path against the CPU oracle and the golden fixtures.

Bars: pivot sequence identical to the reference LU; x within
1e-10 relative to ||x|| (gepp -- lu2's own COMPAT rule is numerically unstable,
two correct implementations of it differ by up to ~1e-6 at n=2048, so COMPAT
is held to identical pivots and 1e-5); HPL-2 scaled residual < 16 (gepp).
"""
import numpy as np
import pytest

import lu2_amd as L
from conftest import golden

pytestmark = pytest.mark.gpu


def gpu_solve(ctx, a0, nb, mode):
    return L.solve(a0, nb, mode, ctx)


# ---------------------------------------------------------------- generator
@pytest.mark.parametrize("n,seed,lda", [(300, 1, 316), (1000, 12345, 1017), (64, 7, 80)])
def test_randomsetmat_bit_exact(ctx, oracle, n, seed, lda):
    dA = ctx.alloc(8 * n * lda)
    dA.upload(np.full(n * lda, np.nan))
    ctx.randomsetmat(n, seed, dA, lda)
    got = dA.download(np.float64, n * lda).reshape(n, lda)
    dA.free()
    assert np.array_equal(got, oracle.randomsetmat(n, seed, lda))


def test_randomsetmat_far_rows(ctx, oracle):
    """Jump-ahead at full size: rows near the start and deep into an n=32768
    matrix against the sequence run sequentially on the CPU."""
    n, lda = 32768, 32784
    dA = ctx.alloc(8 * n * lda)
    try:
        ctx.randomsetmat(n, 1, dA, lda)
        head = dA.download(np.float64, 4 * lda).reshape(4, lda)
        deep = dA.download(np.float64, 4 * lda, offset=8 * 20000 * lda).reshape(4, lda)
    finally:
        dA.free()
    assert np.array_equal(head, oracle.randomrows(n, 1, 0, 4, lda))
    assert np.array_equal(deep, oracle.randomrows(n, 1, 20000, 4, lda))


# --------------------------------------------------------------------- GEMM
def test_gemm_known_answer_identity(ctx):
    """gdrdgemm.c:17-58: I(2048^2) * B == B exactly."""
    N = 2048
    ia = np.eye(N)
    b = (np.arange(N)[:, None] * N + np.arange(N)[None, :]).astype(np.float64)
    dA, dB, dC = ctx.alloc(ia.nbytes), ctx.alloc(b.nbytes), ctx.alloc(b.nbytes)
    dA.upload(ia)
    dB.upload(b)
    dC.upload(np.zeros_like(b))
    ctx.dgemm_dev(N, N, N, 1.0, dA, N, dB, N, 0.0, dC, N)
    c = dC.download(np.float64, N * N).reshape(N, N)
    for d in (dA, dB, dC):
        d.free()
    assert np.array_equal(c, b)


@pytest.mark.parametrize("m,n,k,alpha,beta", [
    (128, 128, 16, -1.0, 1.0), (300, 257, 100, -1.0, 1.0), (1000, 999, 512, -1.0, 1.0),
    (64, 1, 512, 1.0, 0.0), (513, 130, 33, 0.75, -0.5), (1, 700, 3, 2.0, 1.0), (256, 256, 512, 1.0, 0.0)])
def test_gemm_error_bound(ctx, m, n, k, alpha, beta):
    rng = np.random.default_rng(m * 7 + n)
    a, b, c = rng.standard_normal((m, k)), rng.standard_normal((k, n)), rng.standard_normal((m, n))
    dA, dB, dC = ctx.alloc(a.nbytes), ctx.alloc(b.nbytes), ctx.alloc(c.nbytes)
    dA.upload(a)
    dB.upload(b)
    dC.upload(c)
    ctx.dgemm_dev(m, n, k, alpha, dA, k, dB, n, beta, dC, n)
    got = dC.download(np.float64, m * n).reshape(m, n)
    for d in (dA, dB, dC):
        d.free()
    ref = alpha * (a @ b) + beta * c
    u = 2.0 ** -53
    gamma = (k + 2) * u / (1 - (k + 2) * u)
    bound = 2 * gamma * (abs(alpha) * (np.abs(a) @ np.abs(b)) + abs(beta) * np.abs(c))
    assert np.all(np.abs(got - ref) <= bound + 1e-300)


def test_gemm_strided_submatrix(ctx):
    """Leading dimensions larger than the operand widths (lu2's RDIM = n+16)."""
    rng = np.random.default_rng(3)
    big = rng.standard_normal((600, 616))
    m, n, k = 300, 280, 256
    dX = ctx.alloc(big.nbytes)
    dX.upload(big)
    # C = rows 300.., cols 300..; A = rows 300.., cols 0..; B = rows 0.., cols 300..
    ld = 616
    offA, offB, offC = 8 * (300 * ld), 8 * (300), 8 * (300 * ld + 300)
    ctx.dgemm_dev(m, n, k, -1.0, dX, ld, dX, ld, 1.0, dX, ld, offs=(offA, offB, offC))
    got = dX.download(np.float64, 600 * ld).reshape(600, ld)
    dX.free()
    want = big.copy()
    want[300:600, 300:580] -= big[300:600, 0:256] @ big[0:256, 300:580]
    assert np.abs(got - want).max() < 1e-12 * 256 * 4
    assert np.array_equal(got[:300], big[:300])


@pytest.mark.parametrize("m,n,k", [(1000, 64, 300), (1000, 16, 48), (64, 900, 200), (40, 700, 130),
                                   (3000, 128, 128), (600, 256, 96)])
def test_gemm_narrow_tiles_bit_identical(ctx, m, n, k):
    """Products with M or N <= 64, or too small to fill the device in 128x128
    tiles, run in 64x64 tiles; each element gets the same MFMA sequence as in
    the 128x128 tile of a product wide enough for that kernel."""
    rng = np.random.default_rng(m + n + k)
    # the wide product has at least 2048 tiles of 128x128 (128 x 17 and more)
    # and keeps 16-byte rows (the vector path)
    wide_m, wide_n = max(m, 16384), max(n, 17 * 128)
    a = rng.standard_normal((wide_m, k))
    b = rng.standard_normal((k, wide_n))
    c = rng.standard_normal((wide_m, wide_n))
    out = {}
    for tag, (mm, nn) in (("wide", (wide_m, wide_n)), ("narrow", (m, n))):
        dA, dB, dC = ctx.alloc(8 * mm * k), ctx.alloc(8 * k * nn), ctx.alloc(8 * mm * nn)
        dA.upload(np.ascontiguousarray(a[:mm]))
        dB.upload(np.ascontiguousarray(b[:, :nn]))
        dC.upload(np.ascontiguousarray(c[:mm, :nn]))
        ctx.dgemm_dev(mm, nn, k, -1.0, dA, k, dB, nn, 1.0, dC, nn)
        out[tag] = dC.download(np.float64, mm * nn).reshape(mm, nn)
        for d in (dA, dB, dC):
            d.free()
    assert np.array_equal(out["wide"][:m, :n], out["narrow"])
    assert np.abs(out["narrow"] - (c[:m, :n] - a[:m] @ b[:, :n])).max() < 1e-12 * k


def trailing_product(ctx, a, b, c, kernel):
    """C - A B through the trailing-update kernel (even leading dimensions,
    16-byte aligned rows: the shape the factorization hands it)."""
    m, k = a.shape
    n = b.shape[1]
    lda, ldb = k + (k & 1), n + (n & 1)
    ap = np.zeros((m, lda)); ap[:, :k] = a
    bp = np.zeros((k, ldb)); bp[:, :n] = b
    cp = np.zeros((m, ldb)); cp[:, :n] = c
    dA, dB, dC = ctx.alloc(ap.nbytes), ctx.alloc(bp.nbytes), ctx.alloc(cp.nbytes)
    dA.upload(ap)
    dB.upload(bp)
    dC.upload(cp)
    ctx.set_option(L.OPT_TRAIL_KERNEL, kernel)
    ctx.set_option(L.OPT_DGEMM_TRAILING, 1)
    try:
        ctx.dgemm_dev(m, n, k, -1.0, dA, lda, dB, ldb, 1.0, dC, ldb)
    finally:
        ctx.set_option(L.OPT_DGEMM_TRAILING, 0)
        ctx.set_option(L.OPT_TRAIL_KERNEL, 1)   # the library default
    out = dC.download(np.float64, m * ldb).reshape(m, ldb)
    for d in (dA, dB, dC):
        d.free()
    assert np.array_equal(out[:, n:], cp[:, n:])      # nothing written past column n
    return out[:, :n]


@pytest.mark.parametrize("kernel", [1, 0])
@pytest.mark.parametrize("m,n,k", [(128, 128, 16), (300, 257, 512), (1000, 999, 2048), (256, 384, 32),
                                   (640, 129, 48), (129, 1, 64), (1, 300, 16), (384, 256, 40)])
def test_trailing_gemm_error_bound(ctx, kernel, m, n, k):
    """The trailing-update kernels (operands loaded straight into LDS, or
    staged through registers) against the componentwise bound
    |C - fl(C - AB)| <= gamma_{K+2} (|A||B| + |C|), edge tiles (M, N not
    multiples of 128, odd N, K not a multiple of 16) included."""
    rng = np.random.default_rng(m * 31 + n * 7 + k)
    a, b, c = rng.standard_normal((m, k)), rng.standard_normal((k, n)), rng.standard_normal((m, n))
    got = trailing_product(ctx, a, b, c, kernel)
    ref = c - a @ b
    u = 2.0 ** -53
    gamma = (k + 2) * u / (1 - (k + 2) * u)
    bound = 2 * gamma * (np.abs(a) @ np.abs(b) + np.abs(c))
    assert np.all(np.abs(got - ref) <= bound + 1e-300)


def test_trailing_gemm_edge_tiles_bit_identical(ctx):
    """Edge tiles (bounds-checked loads through registers) perform exactly
    the arithmetic of interior tiles (loads straight into LDS): the 300 x 257
    product equals the corresponding corner of a 512 x 512 product, whose
    tiles are all interior, bit for bit."""
    rng = np.random.default_rng(11)
    k = 96
    a = rng.standard_normal((512, k))
    b = rng.standard_normal((k, 512))
    c = rng.standard_normal((512, 512))
    big = trailing_product(ctx, a, b, c, 1)
    small = trailing_product(ctx, a[:300], b[:, :257], c[:300, :257], 1)
    assert np.array_equal(big[:300, :257], small)


# ----------------------------------------------------------------------- LU
# (256, 256): one block; (512, 256): two blocks, so the split schedule's
# remaining columns are b alone; (64, 32): the smallest block
CASES = [(256, 64, 1), (512, 64, 12345), (1024, 128, 1), (1024, 256, 12345),
         (2048, 256, 1), (2048, 512, 12345), (2048, 128, 7), (512, 32, 3),
         (256, 256, 5), (512, 256, 9), (64, 32, 2)]


@pytest.mark.parametrize("n,nb,seed", CASES)
def test_lu_gepp_parity(ctx, oracle, n, nb, seed):
    a0 = oracle.randomsetmat(n, seed)
    x, piv, fac = gpu_solve(ctx, a0, nb, L.PIV_GEPP)
    xo, po, fo = oracle.lumcolumn(a0, nb, oracle.PIV_GEPP)
    assert np.array_equal(piv, po)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()
    assert oracle.residuals(a0, x)[6] < 16.0
    # lu2's whole stored factor form (L/d in the diagonal blocks, unscaled
    # L21, unit-upper U, transformed b in column n)
    scale = np.abs(fo[:, : n + 1]).max()
    assert np.abs(fac[:, : n + 1] - fo[:, : n + 1]).max() <= 1e-11 * scale


@pytest.mark.parametrize("n,nb,seed", CASES)
def test_lu_compat_parity(ctx, oracle, n, nb, seed):
    a0 = oracle.randomsetmat(n, seed)
    x, piv, _ = gpu_solve(ctx, a0, nb, L.PIV_COMPAT)
    xo, po, _ = oracle.lumcolumn(a0, nb, oracle.PIV_COMPAT)
    assert np.array_equal(piv, po)
    assert np.abs(x - xo).max() <= 1e-5 * np.abs(xo).max()


@pytest.mark.parametrize("n,seed", [(256, 1), (512, 12345), (1024, 1)])
def test_lu_gepp_matches_lapack_fixture(ctx, oracle, n, seed):
    g = golden(f"lapack_n{n}_s{seed}.npz")
    a0 = oracle.randomsetmat(n, seed)
    x, piv, _ = gpu_solve(ctx, a0, 128 if n >= 256 else 64, L.PIV_GEPP)
    assert np.array_equal(piv, g["piv"])
    assert np.abs(x - g["x"]).max() <= 1e-10 * np.abs(g["x"]).max()


@pytest.mark.parametrize("n,seed", [(64, 1), (256, 12345)])
def test_lu_compat_matches_rule_fixture(ctx, oracle, n, seed):
    g = golden(f"compat_n{n}_s{seed}.npz")
    a0 = oracle.randomsetmat(n, seed)
    _, piv, _ = gpu_solve(ctx, a0, 32, L.PIV_COMPAT)
    assert np.array_equal(piv, g["piv"])


@pytest.mark.parametrize("leaf_w", [16, 32])
def test_lu_leaf_widths_and_wide_lda(ctx, oracle, leaf_w):
    n, nb = 1024, 128
    a0 = oracle.randomsetmat(n, 11, n + 40)
    ctx.set_option(L.OPT_LEAF_W, leaf_w)
    try:
        x, piv, _ = gpu_solve(ctx, a0, nb, L.PIV_GEPP)
    finally:
        ctx.set_option(L.OPT_LEAF_W, 16)   # the library default
    xo, po, _ = oracle.lumcolumn(a0, nb, oracle.PIV_GEPP)
    assert np.array_equal(piv, po)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()


def test_lu_single_block(ctx, oracle):
    n = 256
    a0 = oracle.randomsetmat(n, 2)
    x, piv, _ = gpu_solve(ctx, a0, n, L.PIV_GEPP)
    xo, po, _ = oracle.lumcolumn(a0, n, oracle.PIV_GEPP)
    assert np.array_equal(piv, po)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()


def test_lu_rejects_reference_limits(ctx, oracle):
    a0 = oracle.randomsetmat(96, 1)
    with pytest.raises(L.Lu2miError):
        gpu_solve(ctx, a0, 48, L.PIV_GEPP)   # nb not a power of two
    with pytest.raises(L.Lu2miError):
        gpu_solve(ctx, a0, 64, L.PIV_GEPP)   # n % nb != 0


def test_lu_singular_matrix_completes(ctx, oracle):
    """A zero column gives a zero pivot: the call completes (no hang) and,
    as in lu2, the solution is not finite."""
    n = 256
    a0 = oracle.randomsetmat(n, 4)
    a0[:, 10] = 0.0
    x, _, _ = gpu_solve(ctx, a0, 64, L.PIV_GEPP)
    assert not np.all(np.isfinite(x))


def test_lu_n4096_full_oracle(ctx, oracle):
    n, nb = 4096, 512
    a0 = oracle.randomsetmat(n, 1)
    x, piv, _ = gpu_solve(ctx, a0, nb, L.PIV_GEPP)
    xo, po, _ = oracle.lumcolumn(a0, nb, oracle.PIV_GEPP)
    assert np.array_equal(piv, po)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()


def test_lu_n16384_full_oracle(ctx, oracle):
    """Half the config-2 size against the whole CPU oracle: the matrix is
    generated on the device (bit-identical to the oracle's), factored with
    nb=512, and the pivot sequence must be identical and x within 1e-10."""
    n, nb = 16384, 512
    lda = n + 16
    dA = ctx.alloc(8 * n * lda)
    dx = ctx.alloc(8 * n)
    dp = ctx.alloc(4 * n)
    ctx.randomsetmat(n, 1, dA, lda)
    ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
    ctx.lumcolumn_dev(n, dA, lda, nb, dp, dx)
    x = dx.download(np.float64, n)
    piv = dp.download(np.int32, n)
    for d in (dA, dx, dp):
        d.free()
    a0 = oracle.randomsetmat(n, 1)
    xo, po, _ = oracle.lumcolumn(a0, nb, oracle.PIV_GEPP)
    assert np.array_equal(piv, po)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()


def test_lu_n32768_residual(ctx):
    """Config 2 size (n=32768, nb=512): the HPL-2 scaled residual on the
    device against the regenerated matrix, and the pivot sequence shape."""
    n, nb = 32768, 512
    lda = n + 16
    dA0 = ctx.alloc(8 * n * lda)
    dA = ctx.alloc(8 * n * lda)
    dx = ctx.alloc(8 * n)
    dp = ctx.alloc(4 * n)
    try:
        ctx.randomsetmat(n, 1, dA0, lda)
        dA.copy_from(dA0)
        ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
        ctx.lumcolumn_dev(n, dA, lda, nb, dp, dx)
        res = ctx.residual_dev(n, dA0, lda, dx)
        piv = dp.download(np.int32, n)
    finally:
        for d in (dA0, dA, dx, dp):
            d.free()
    assert res[6] < 16.0
    assert np.all(piv >= np.arange(n)) and np.all(piv < n)


@pytest.mark.parametrize("n,nb,mode", [(4096, 256, L.PIV_GEPP), (2048, 512, L.PIV_COMPAT), (3072, 128, L.PIV_GEPP)])
def test_lookahead_schedule_bit_identical(ctx, n, nb, mode):
    """Factoring panel i+1 on the second stream during block i's split
    trailing update performs the same arithmetic per element: factors,
    pivots and x are bit-identical to the serial schedule."""
    lda = n + 16
    outs = []
    for la in (0, 1):
        dA = ctx.alloc(8 * n * lda)
        dx = ctx.alloc(8 * n)
        dp = ctx.alloc(4 * n)
        ctx.randomsetmat(n, 3, dA, lda)
        ctx.set_option(L.OPT_PIVMODE, mode)
        ctx.set_option(L.OPT_LOOKAHEAD, la)
        ctx.lumcolumn_dev(n, dA, lda, nb, dp, dx)
        outs.append((dA.download(np.float64, n * lda), dx.download(np.float64, n), dp.download(np.int32, n)))
        for d in (dA, dx, dp):
            d.free()
    ctx.set_option(L.OPT_LOOKAHEAD, 1)   # the library default
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,nb,mode", [(4096, 256, L.PIV_GEPP), (2048, 512, L.PIV_COMPAT), (34816, 2048, L.PIV_GEPP)])
def test_strict_sync_bit_identical(ctx, n, nb, mode):
    """Publishing the leaf's exchange slots and the back substitution's
    frontier with agent-scope release/acquire changes only cache maintenance:
    factors, pivots and x are bit-identical to the default publication."""
    lda = n + 16
    outs = []
    for strict in (0, 1):
        dA = ctx.alloc(8 * n * lda)
        dx = ctx.alloc(8 * n)
        dp = ctx.alloc(4 * n)
        ctx.randomsetmat(n, 9, dA, lda)
        ctx.set_option(L.OPT_PIVMODE, mode)
        ctx.set_option(L.OPT_STRICT_SYNC, strict)
        try:
            ctx.lumcolumn_dev(n, dA, lda, nb, dp, dx)
        finally:
            ctx.set_option(L.OPT_STRICT_SYNC, 0)   # the library default
        outs.append((dA.download(np.float64, n * lda), dx.download(np.float64, n), dp.download(np.int32, n)))
        for d in (dA, dx, dp):
            d.free()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,nb,mode,sub", [
    (1024, 256, L.PIV_COMPAT, 256),
    (2048, 512, L.PIV_GEPP, 256),
    (2048, 128, L.PIV_GEPP, 64),
    (4096, 1024, L.PIV_COMPAT, 128),
    (1536, 512, L.PIV_GEPP, 32),
    (20480, 2048, L.PIV_GEPP, 256),
])
def test_subpanel_kernel_bit_identical(ctx, n, nb, mode, sub):
    """The persistent sub-panel launch (leaves, their swaps and the recursion
    nodes of up to 256 columns between grid barriers) performs the same
    arithmetic as one launch per step: factors, pivots and x bit-identical,
    with and without the lookahead schedule."""
    lda = n + 16
    outs = []
    for sw, la in ((0, 0), (sub, 0), (sub, 1)):
        dA = ctx.alloc(8 * n * lda)
        dx = ctx.alloc(8 * n)
        dp = ctx.alloc(4 * n)
        ctx.randomsetmat(n, 5, dA, lda)
        ctx.set_option(L.OPT_PIVMODE, mode)
        ctx.set_option(L.OPT_SUBPANEL, sw)
        ctx.set_option(L.OPT_LOOKAHEAD, la)
        ctx.lumcolumn_dev(n, dA, lda, nb, dp, dx)
        outs.append((dA.download(np.float64, n * lda), dx.download(np.float64, n), dp.download(np.int32, n)))
        for d in (dA, dx, dp):
            d.free()
    ctx.set_option(L.OPT_SUBPANEL, 0)   # the library defaults
    ctx.set_option(L.OPT_LOOKAHEAD, 1)
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)


def test_subpanel_option_rejects(ctx):
    for bad in (8, 48, 512, -1):
        with pytest.raises(L.Lu2miError):
            ctx.set_option(L.OPT_SUBPANEL, bad)


@pytest.mark.parametrize("n,nb,mode", [(4096, 256, L.PIV_GEPP), (3040, 32, L.PIV_GEPP), (2048, 512, L.PIV_COMPAT)])
def test_backsub128_solution(ctx, oracle, n, nb, mode):
    """Back substitution in 128-row blocks (LU2MI_OPT_BACKSUB128): the
    factors and pivots are the 64-row run's bits (only the solve differs);
    x is the same unit-upper solve summed in another order, so it agrees
    with the 64-row x to the solve's rounding, and with the oracle's x as the
    64-row solve does.  n = 3040 leaves a ragged last 128-row block."""
    lda = n + 16
    outs = []
    for b128 in (0, 1):
        dA = ctx.alloc(8 * n * lda)
        dx = ctx.alloc(8 * n)
        dp = ctx.alloc(4 * n)
        ctx.randomsetmat(n, 5, dA, lda)
        ctx.set_option(L.OPT_PIVMODE, mode)
        ctx.set_option(L.OPT_BACKSUB128, b128)
        try:
            ctx.lumcolumn_dev(n, dA, lda, nb, dp, dx)
        finally:
            ctx.set_option(L.OPT_BACKSUB128, 0)
        a = dA.download(np.float64, n * lda).reshape(n, lda)
        outs.append((a[:, :n].copy(), dx.download(np.float64, n), dp.download(np.int32, n), a))
        for d in (dA, dx, dp):
            d.free()
    (f0, x0, p0, _), (f1, x1, p1, a1) = outs
    assert np.array_equal(f0, f1) and np.array_equal(p0, p1)
    if mode == L.PIV_GEPP:
        # two backward-stable triangular solves of a well-conditioned U
        assert np.abs(x1 - x0).max() <= 1e-12 * np.abs(x0).max()
        xo, po, _ = oracle.lumcolumn(oracle.randomsetmat(n, 5), nb, oracle.PIV_GEPP)
        assert np.array_equal(p1, po)
        assert np.abs(x1 - xo).max() <= 1e-10 * np.abs(xo).max()
    # the 128-row solve against the factors themselves (both pivot rules):
    # U x1 = y with U the unit upper triangle and y column n after the
    # elimination, to the backward error of a triangular solve (a wrong
    # index or a dropped term shows at O(1))
    U = np.triu(f1, 1) + np.eye(n)
    y = a1[:, n]
    r = np.abs(U @ x1 - y).max() / (np.abs(U).sum(1).max() * np.abs(x1).max() * n * 2.0 ** -53)
    assert r < 16.0


def test_backsub128_refinement(ctx, oracle):
    """The mixed solve's re-solves with 128-row blocks (FP32 factors, FP64
    vector): same target met, x within 1e-10 of the oracle's."""
    n, nb = 4096, 512
    lda = n + 16
    a0 = oracle.randomsetmat(n, 3)
    dA = ctx.alloc(8 * n * lda)
    dx = ctx.alloc(8 * n)
    try:
        dA.upload(a0)
        ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
        ctx.set_option(L.OPT_BACKSUB128, 1)
        info = ctx.solve_mixed(n, dA, lda, nb, dx, L.PREC_FP32, maxiter=30, target=1.0)
        x = dx.download(np.float64, n)
    finally:
        ctx.set_option(L.OPT_BACKSUB128, 0)
        dA.free()
        dx.free()
    assert info["resid_hpl2"] <= 1.0
    xo, _, _ = oracle.lumcolumn(a0, nb, oracle.PIV_GEPP)
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()
