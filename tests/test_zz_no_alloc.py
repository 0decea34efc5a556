"""lu2mi_residual_dev_noalloc: the residual figures with the partial results
in the library's fixed per-device scratch array instead of an allocation.
Compared with lu2mi_residual_dev on the same inputs (same kernels, same
order: bit-identical), on two calls in a row, the second smaller, and the
size limit rejected on the host."""
import numpy as np
import pytest

import lu2_amd as L

pytestmark = pytest.mark.gpu


def test_noalloc_residual_matches_allocating_entry(ctx, oracle):
    dout = ctx.alloc(8 * 10)
    try:
        for n, seed in ((3000, 1), (1000, 12345)):
            lda = n + 16
            a0 = oracle.randomsetmat(n, seed)
            x = np.linalg.solve(a0[:, :n], a0[:, n])
            x += 1e-9 * np.random.default_rng(seed).standard_normal(n)
            dA, dx = ctx.alloc(8 * n * lda), ctx.alloc(8 * n)
            try:
                dA.upload(a0)
                dx.upload(x)
                dout.upload(np.full(10, np.nan))
                ctx.residual_dev_noalloc(n, dA, lda, dx, dout)
                ctx.synchronize()
                got = dout.download(np.float64, 10)
                want = ctx.residual_dev(n, dA, lda, dx)
            finally:
                dA.free()
                dx.free()
            assert np.array_equal(got, want), (n, got, want)
            assert abs(got[6] - oracle.residuals(a0, x)[6]) <= 2.0
    finally:
        dout.free()


def test_noalloc_residual_rejects_size_above_limit(ctx):
    n = L.RESID_NMAX + 1
    buf = ctx.alloc(8 * 16)
    try:
        with pytest.raises(L.Lu2miError):
            # rejected on the host before anything is launched
            ctx.residual_dev_noalloc(n, buf, n + 16, buf, buf)
    finally:
        buf.free()
