"""16-bit trailing updates (FP16 / BF16 operands, FP32 accumulation,
per-block power-of-two scaling) and the solves built on them.

The update is compared with an exact model of what it is meant to compute:
every 128-row block of L and 128-column block of U scaled by the power of two
that puts its largest |entry| in [2^13, 2^14), rounded to nearest-even in the
16-bit format, multiplied exactly (FP64) and unscaled.  The only freedom left
is the FP32 accumulation, bounded by K u32 sum|l||u| plus the final rounding.
The solves are checked against the FP64 oracle on lu2's random matrix made
diagonally dominant (the 16-bit configuration's matrix)."""
import numpy as np
import pytest

import lu2_amd as L

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


def round_bf16(x32: np.ndarray) -> np.ndarray:
    b = x32.astype(np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def to16(x: np.ndarray, fmt: int) -> np.ndarray:
    """The operand the matrix cores see: the FP32 value rounded to FP16 or
    BF16, or for FP16x2 the sum hi + lo of hi = fp16(v), lo = fp16(v - hi)."""
    x32 = x.astype(np.float32)
    if fmt == L.LOWP_BF16:
        return round_bf16(x32).astype(np.float64)
    hi = x32.astype(np.float16)
    if fmt == L.LOWP_FP16:
        return hi.astype(np.float64)
    lo = (x32 - hi.astype(np.float32)).astype(np.float16)
    return hi.astype(np.float64), lo.astype(np.float64)


def shifts(blockmax: np.ndarray) -> np.ndarray:
    _, e = np.frexp(blockmax.astype(np.float32))
    return np.where(blockmax > 0, 14 - e, 0)


def model_update(l, u, c, fmt):
    m, k = l.shape
    n = u.shape[1]
    sa = shifts(np.array([np.abs(l[r:r + 128]).max() for r in range(0, m, 128)]))
    sb = shifts(np.array([np.abs(u[:, q:q + 128]).max() for q in range(0, n, 128)]))
    ra = np.repeat(sa, 128)[:m].astype(np.float64)
    rb = np.repeat(sb, 128)[:n].astype(np.float64)
    l16 = to16(l * 2.0 ** ra[:, None], fmt)
    u16 = to16(u * 2.0 ** rb[None, :], fmt)
    unscale = 2.0 ** (-(ra[:, None] + rb[None, :]))
    if fmt == L.LOWP_FP16X2:   # hi*hi + hi*lo + lo*hi (lo*lo is not formed)
        (lh, ll), (uh, ul) = l16, u16
        prod = lh @ uh + lh @ ul + ll @ uh
        mag = np.abs(lh) @ np.abs(uh) + np.abs(lh) @ np.abs(ul) + np.abs(ll) @ np.abs(uh)
        return c - prod * unscale, 3 * k * U32 * mag * unscale
    prod = (l16 @ u16) * unscale
    bound = k * U32 * (np.abs(l16) @ np.abs(u16)) * unscale
    return c - prod, bound


@pytest.mark.parametrize("fmt", [L.LOWP_FP16, L.LOWP_BF16, L.LOWP_FP16X2])
@pytest.mark.parametrize("m,n,k", [(128, 128, 64), (300, 257, 128), (1000, 999, 512), (129, 1, 64)])
def test_update16_matches_rounding_model(ctx, fmt, m, n, k):
    rng = np.random.default_rng(m * 7 + n + k + fmt)
    l = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    u = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    # blocks far outside FP16's range: only the per-block scales keep them
    l[:128] *= np.float32(2.0 ** 20)
    u[:, :min(n, 128)] *= np.float32(2.0 ** -18)
    if m > 256:
        l[256:] *= np.float32(2.0 ** -30)
    c = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    dL, dU, dC = ctx.alloc(l.nbytes), ctx.alloc(u.nbytes), ctx.alloc(c.nbytes)
    dL.upload(l)
    dU.upload(u)
    dC.upload(c)
    ctx.gemm16_dev(fmt, m, n, k, dL, k, dU, n, dC, n)
    got = dC.download(np.float32, m * n).reshape(m, n).astype(np.float64)
    for d in (dL, dU, dC):
        d.free()
    ref, bound = model_update(l.astype(np.float64), u.astype(np.float64), c.astype(np.float64), fmt)
    err = np.abs(got - ref)
    tol = 2 * bound + 2 * U32 * np.abs(ref) + 1e-30
    assert (err <= tol).all(), f"max err/tol {np.max(err / tol):.3g}"


def test_update16_rejects_bad_k(ctx):
    d = ctx.alloc(4 * 64 * 64)
    with pytest.raises(L.Lu2miError):
        ctx.gemm16_dev(L.LOWP_FP16, 64, 64, 32, d, 64, d, 64, d, 64)
    with pytest.raises(L.Lu2miError):
        ctx.gemm16_dev(7, 64, 64, 64, d, 64, d, 64, d, 64)
    d.free()


def test_diag_dominant_generator(ctx, oracle):
    n, lda = 1000, 1016
    dA = ctx.alloc(8 * n * lda)
    ctx.randomsetmat(n, 5, dA, lda)
    ctx.make_diag_dominant(n, dA, lda)
    got = dA.download(np.float64, n * lda).reshape(n, lda)
    dA.free()
    ref = oracle.diag_dominant(oracle.randomsetmat(n, 5, lda))
    off = ~np.eye(n, lda, dtype=bool)
    assert np.array_equal(got[off], ref[off])
    d, dr = np.diag(got[:, :n]), np.diag(ref[:, :n])
    assert np.abs(d - dr).max() <= 4 * n * 2.0 ** -53 * np.abs(dr).max()
    # |a_ii| >= sum_{j != i} |a_ij| (equality for rows whose drawn a_ii was negative)
    offsum = np.abs(got[:, :n]).sum(axis=1) - np.abs(d)
    assert bool((np.abs(d) >= offsum * (1 - 1e-12)).all())


@pytest.mark.parametrize("prec", [L.PREC_FP16, L.PREC_BF16, L.PREC_FP16X2])
@pytest.mark.parametrize("n,nb", [(1024, 128), (4096, 512)])
def test_mixed_16bit_refinement(ctx, oracle, prec, n, nb):
    lda = n + 16
    dA = ctx.alloc(8 * n * lda)
    dx = ctx.alloc(8 * n)
    ctx.randomsetmat(n, 3, dA, lda)
    ctx.make_diag_dominant(n, dA, lda)
    ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
    info = ctx.solve_mixed(n, dA, lda, nb, dx, prec=prec, maxiter=30, target=1.0)
    x = dx.download(np.float64, n)
    res = ctx.residual_dev(n, dA, lda, dx)
    a0 = dA.download(np.float64, n * lda).reshape(n, lda)
    for d in (dA, dx):
        d.free()
    assert info["resid_hpl2"] <= 1.0 and res[6] <= 1.0
    assert info["iterations"] >= 1 and info["resid_hpl2_initial"] > info["resid_hpl2"]
    xo, _, _ = oracle.gepp_unblocked(a0, nb, oracle.PIV_GEPP)
    assert np.abs(x - xo).max() / np.abs(xo).max() < 1e-10


def test_update16_split_is_near_fp32(ctx):
    """FP16x2 operands carry ~22 bits: the update agrees with the exact FP64
    product to FP32-like accuracy, far inside what plain FP16 reaches."""
    rng = np.random.default_rng(99)
    m, n, k = 512, 384, 256
    l = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    u = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    exact = -(l.astype(np.float64) @ u.astype(np.float64))
    errs = {}
    for fmt in (L.LOWP_FP16, L.LOWP_FP16X2):
        dL, dU, dC = ctx.alloc(l.nbytes), ctx.alloc(u.nbytes), ctx.alloc(4 * m * n)
        dL.upload(l)
        dU.upload(u)
        dC.upload(np.zeros(m * n, np.float32))
        ctx.gemm16_dev(fmt, m, n, k, dL, k, dU, n, dC, n)
        got = dC.download(np.float32, m * n).reshape(m, n).astype(np.float64)
        errs[fmt] = np.abs(got - exact).max() / (np.abs(l).astype(np.float64) @ np.abs(u)).max()
        for d in (dL, dU, dC):
            d.free()
    assert errs[L.LOWP_FP16X2] < 4 * U32 * 8
    assert errs[L.LOWP_FP16] > 50 * errs[L.LOWP_FP16X2]


@pytest.mark.parametrize("n,nb", [(2048, 256), (4096, 512)])
def test_mixed_fp16x2_on_lu2_matrix(ctx, oracle, n, nb):
    """lu2's own random matrix (not diagonally dominant): plain FP16 factors
    are too coarse for it, the split operands converge like FP32."""
    lda = n + 16
    dA = ctx.alloc(8 * n * lda)
    dx = ctx.alloc(8 * n)
    ctx.randomsetmat(n, 1, dA, lda)
    ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
    # kappa ~ 1e6 here: agreement with the FP64 solution to 1e-10 needs the
    # residual refined to the FP64 LU's own level (~0.005-0.05), not just < 1
    info = ctx.solve_mixed(n, dA, lda, nb, dx, prec=L.PREC_FP16X2, maxiter=30, target=0.05)
    x = dx.download(np.float64, n)
    a0 = dA.download(np.float64, n * lda).reshape(n, lda)
    for d in (dA, dx):
        d.free()
    assert info["resid_hpl2"] <= 0.05
    xo, _, _ = oracle.gepp_unblocked(a0, nb, oracle.PIV_GEPP)
    assert np.abs(x - xo).max() / np.abs(xo).max() < 1e-10


def test_mixed_16bit_rejects_small_nb(ctx):
    n, lda = 256, 272
    dA = ctx.alloc(8 * n * lda)
    dx = ctx.alloc(8 * n)
    ctx.randomsetmat(n, 1, dA, lda)
    with pytest.raises(L.Lu2miError):
        ctx.solve_mixed(n, dA, lda, 32, dx, prec=L.PREC_FP16)
    for d in (dA, dx):
        d.free()


@pytest.mark.parametrize("prec", [L.PREC_FP32, L.PREC_FP16X2])
def test_mixed_lookahead_bit_identical(ctx, prec):
    """The mixed solve's factorization with and without the lookahead
    schedule performs the same arithmetic: identical iterates and x."""
    n, nb = 2048, 256
    lda = n + 16
    dA = ctx.alloc(8 * n * lda)
    ctx.randomsetmat(n, 9, dA, lda)
    ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
    xs, infos = [], []
    for la in (0, 1):
        dx = ctx.alloc(8 * n)
        ctx.set_option(L.OPT_LOOKAHEAD, la)
        infos.append(ctx.solve_mixed(n, dA, lda, nb, dx, prec=prec, maxiter=30, target=1.0))
        xs.append(dx.download(np.float64, n))
        dx.free()
    ctx.set_option(L.OPT_LOOKAHEAD, 1)   # the library default
    dA.free()
    assert np.array_equal(xs[0], xs[1])
    assert infos[0]["iterations"] == infos[1]["iterations"]
    assert infos[0]["resid_hpl2_initial"] == infos[1]["resid_hpl2_initial"]


@pytest.mark.parametrize("prec", [L.PREC_FP32, L.PREC_FP16X2])
def test_mixed_subpanel_bit_identical(ctx, prec):
    """The FP32 factorization inside the mixed solve with the persistent
    sub-panel launch and with one launch per step: identical iterates and x."""
    n, nb = 2048, 512
    lda = n + 16
    dA = ctx.alloc(8 * n * lda)
    ctx.randomsetmat(n, 11, dA, lda)
    ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
    xs, infos = [], []
    for sw in (0, 256):
        dx = ctx.alloc(8 * n)
        ctx.set_option(L.OPT_SUBPANEL, sw)
        infos.append(ctx.solve_mixed(n, dA, lda, nb, dx, prec=prec, maxiter=30, target=1.0))
        xs.append(dx.download(np.float64, n))
        dx.free()
    ctx.set_option(L.OPT_SUBPANEL, 0)   # the library default
    dA.free()
    assert np.array_equal(xs[0], xs[1])
    assert infos[0]["iterations"] == infos[1]["iterations"]
    assert infos[0]["resid_hpl2_initial"] == infos[1]["resid_hpl2_initial"]


@pytest.mark.parametrize("fmt", [L.LOWP_FP16, L.LOWP_BF16, L.LOWP_FP16X2])
def test_update16_wide_tile_bit_identical(ctx, fmt):
    """The 256 x 256 tile (128 x 64 per wave, inside one scaling-block pair)
    performs each accumulator's matrix-core ops in the 128 x 128 tile's
    order: identical bits, edge tiles included (M, N not multiples of 256,
    the last wide tile reaching past the last scaling block)."""
    m, n, k = 4096 + 100, 4096 + 37, 192
    rng = np.random.default_rng(fmt)
    l = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    u = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    l[:128] *= np.float32(2.0 ** 12)
    u[:, 256:384] *= np.float32(2.0 ** -9)
    c = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    out = []
    for wide in (0, 1):
        dL, dU, dC = ctx.alloc(l.nbytes), ctx.alloc(u.nbytes), ctx.alloc(c.nbytes)
        dL.upload(l)
        dU.upload(u)
        dC.upload(c)
        ctx.set_option(L.OPT_WIDE16, wide)
        try:
            ctx.gemm16_dev(fmt, m, n, k, dL, k, dU, n, dC, n)
        finally:
            ctx.set_option(L.OPT_WIDE16, 0)   # the library default
        out.append(dC.download(np.float32, m * n))
        for d in (dL, dU, dC):
            d.free()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("prec", [L.PREC_FP16X2, L.PREC_BF16])
def test_mixed_wide_tile_bit_identical(ctx, prec):
    """A mixed solve whose trailing updates use the wide tile gives the same
    factors and solution as with the 128 x 128 tile."""
    n, nb = 8192, 512
    lda = n + 16
    dA = ctx.alloc(8 * n * lda)
    ctx.randomsetmat(n, 4, dA, lda)
    ctx.make_diag_dominant(n, dA, lda)
    ctx.set_option(L.OPT_PIVMODE, L.PIV_GEPP)
    xs, infos = [], []
    for wide in (0, 1):
        dx = ctx.alloc(8 * n)
        ctx.set_option(L.OPT_WIDE16, wide)
        try:
            infos.append(ctx.solve_mixed(n, dA, lda, nb, dx, prec=prec, maxiter=30, target=1.0))
        finally:
            ctx.set_option(L.OPT_WIDE16, 0)
        xs.append(dx.download(np.float64, n))
        dx.free()
    dA.free()
    assert np.array_equal(xs[0], xs[1])
    assert infos[0]["iterations"] == infos[1]["iterations"]
