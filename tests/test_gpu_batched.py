"""Batched uneven solve: many independent systems in one graph-captured launch.

Every system is checked against the oracle's textbook GEPP on the same
[A | B] (generated on the device from per-system drand48 streams and checked
bit-exact against the oracle's generator first):
  FP64 form: pivot sequence identical, x within max(1e-10, 100 cond_2(A) u)
    relative to ||x|| (per right-hand side; the random systems include
    ill-conditioned ones), HPL-2 scaled residual < 16 and equal to the one
    recomputed here from x within the rounding of two FP64 evaluations;
  FP32 form (FP32 factors + FP64 refinement): refinement reaches the
    residual target and x agrees with the oracle's FP64 x to the same bar.
Graph replays are compared with an eager launch bit for bit, with eager work
and synchronisations between replays.
"""
import numpy as np
import pytest

import lu2_amd as L

pytestmark = pytest.mark.gpu

BASE = 1000


def make_batch(ctx, n, nrhs, prec=L.PREC_FP64, lda=None, maxiter=10, target=1.0):
    b = L.Batch(ctx, n, nrhs, prec=prec, maxiter=maxiter, target=target, lda=lda)
    dA = ctx.alloc(8 * max(1, b.a_elems))
    dX = ctx.alloc(8 * max(1, b.x_elems))
    b.generate(dA, BASE)
    return b, dA, dX


def host_system(b, a_all, s):
    n, r, lda, off = int(b.n[s]), int(b.nrhs[s]), int(b.lda[s]), int(b.a_off[s])
    return a_all[off:off + n * lda].reshape(n, lda)


def check_against_oracle(oracle, b, dA, dX, fp32=False, systems=None):
    a_all = dA.download(np.float64, b.a_elems)
    x_all = dX.download(np.float64, b.x_elems)
    resid, iters, piv = b.results()
    poff = np.concatenate([[0], np.cumsum(b.n)[:-1]])
    for s in (range(len(b.n)) if systems is None else systems):
        n, r = int(b.n[s]), int(b.nrhs[s])
        a = host_system(b, a_all, s)
        ref = oracle.random_aug(n, r, BASE + s, lda=a.shape[1])
        assert np.array_equal(a, ref), f"system {s}: generator differs from drand48"
        opiv, ox, _ = oracle.gepp_solve(a, r)
        x = x_all[b.x_off[s]:b.x_off[s] + n * r].reshape(n, r)
        if not fp32:
            assert np.array_equal(piv[poff[s]:poff[s] + n], opiv), f"system {s} (n={n}): pivots differ"
        err = np.abs(x - ox).max(axis=0) / np.abs(ox).max(axis=0)
        # two backward-stable solutions differ by about cond(A) u: random
        # systems include ill-conditioned ones, so the bar is 1e-10 or
        # 100 cond_2(A) u, whichever is larger
        tol = max(1e-10, 100.0 * np.linalg.cond(a[:, :n]) * 2.0 ** -53)
        assert err.max() < tol, f"system {s} (n={n}, nrhs={r}): rel err {err.max():.3g} tol {tol:.3g}"
        res = oracle.hpl2_multi(a, x)
        assert resid[s] < 16 and res < 16, (s, resid[s], res)
        # the device's residual is the same quantity: each FP64 evaluation of
        # b - Ax carries an error of at most gamma_n (|b| + |A||x|), i.e. about
        # 1 in these scaled units, so two correct evaluations differ by <= 2
        assert abs(resid[s] - res) <= 2.0, (s, resid[s], res)
        if fp32:
            assert resid[s] <= 1.0


SIZES = [1, 2, 3, 16, 17, 31, 32, 33, 47, 64, 100, 129, 255, 256, 300, 511, 512]


def test_batch_generator_bit_exact(ctx, oracle):
    n = [5, 64, 33]
    nrhs = [1, 4, 8]
    lda = [9, 70, 41]   # padded rows
    b, dA, dX = make_batch(ctx, n, nrhs, lda=lda)
    a_all = dA.download(np.float64, b.a_elems)
    for s in range(3):
        got = host_system(b, a_all, s)
        assert np.array_equal(got, oracle.random_aug(n[s], nrhs[s], BASE + s, lda=lda[s]))
    b.close()
    for d in (dA, dX):
        d.free()


@pytest.mark.parametrize("cfg", [(256, 16, 16, 512), (256, 32, 32, 0), (512, 16, 32, 128)])
def test_batch_fp64_parity(ctx, oracle, cfg):
    rng = np.random.default_rng(7)
    n = SIZES + list(rng.integers(1, 513, 15))
    nrhs = list(rng.integers(1, 9, len(n)))
    b, dA, dX = make_batch(ctx, n, nrhs)
    b.set_config(*cfg)
    info = b.solve(dA, dX)
    assert info["n_failed"] == 0 and info["max_resid_hpl2"] < 16
    assert info["max_iterations"] == 0
    assert (info["nt"], info["wp_small"], info["wp_large"], info["split"]) == cfg
    check_against_oracle(oracle, b, dA, dX)
    b.close()
    for d in (dA, dX):
        d.free()


def test_batch_fp32_refinement(ctx, oracle):
    rng = np.random.default_rng(11)
    n = SIZES + list(rng.integers(32, 513, 10))
    nrhs = list(rng.integers(1, 5, len(n)))
    b, dA, dX = make_batch(ctx, n, nrhs, prec=L.PREC_FP32, maxiter=10, target=1.0)
    info = b.solve(dA, dX)
    assert info["n_failed"] == 0 and info["max_resid_hpl2"] <= 1.0
    _, iters, _ = b.results()
    assert info["max_iterations"] == iters.max() >= 1   # FP32 factors alone do not reach 1.0 at n=512
    check_against_oracle(oracle, b, dA, dX, fp32=True)
    b.close()
    for d in (dA, dX):
        d.free()


def test_batch_graph_replays_match_eager(ctx):
    rng = np.random.default_rng(3)
    n = list(rng.integers(32, 300, 200))
    nrhs = list(rng.integers(1, 5, 200))
    b, dA, dX = make_batch(ctx, n, nrhs)
    b.solve(dA, dX, eager=True)
    x0 = dX.download(np.float64, b.x_elems)
    r0 = b.results()
    scratch = ctx.alloc(8 * 4096)
    for rep in range(3):
        scratch.upload(np.full(4096, float(rep)))       # eager work between replays
        dX.upload(np.zeros(b.x_elems))
        b.solve(dA, dX)
        ctx.set_option(L.OPT_TIMING, 0)
        assert np.array_equal(dX.download(np.float64, b.x_elems), x0), f"replay {rep}"
        r = b.results()
        assert np.array_equal(r[0], r0[0]) and np.array_equal(r[2], r0[2])
    scratch.free()
    b.close()
    for d in (dA, dX):
        d.free()


def test_batch_tuner_and_config5_mix(ctx, oracle):
    """The config-5 mix (4096 systems, n ~ U{32..512}, 1-4 right-hand sides),
    tuned, both forms; a sample of systems against the oracle."""
    rng = np.random.default_rng(2024)
    n = rng.integers(32, 513, 4096)
    nrhs = rng.integers(1, 5, 4096)
    b, dA, dX = make_batch(ctx, n, nrhs)
    cfg = b.tune(dA, dX, reps=1)
    assert cfg[0] in (256, 512) and cfg[1] in (16, 32) and cfg[2] in (16, 32)
    info = b.solve(dA, dX)
    assert info["n_failed"] == 0 and info["max_resid_hpl2"] < 16, info
    sample = list(rng.choice(4096, 6, replace=False)) + [int(np.argmax(n))]
    check_against_oracle(oracle, b, dA, dX, systems=sample)
    b.close()
    b32 = L.Batch(ctx, n, nrhs, prec=L.PREC_FP32, maxiter=10, target=1.0)
    b32.tune(dA, dX, reps=1)
    info = b32.solve(dA, dX)
    assert info["n_failed"] == 0 and info["max_resid_hpl2"] <= 1.0, info
    check_against_oracle(oracle, b32, dA, dX, fp32=True, systems=sample)
    b32.close()
    for d in (dA, dX):
        d.free()


def test_batch_empty_and_rejects(ctx):
    b = L.Batch(ctx, [], [])
    info = b.solve(None, None)
    assert info["n_failed"] == 0
    b.close()
    for n, r, lda in [([513], [1], None), ([0], [1], None), ([8], [9], None), ([8], [2], [9])]:
        with pytest.raises(L.Lu2miError):
            L.Batch(ctx, n, r, lda=lda)


def test_batch_singular_system_reports(ctx):
    """A system with a zero column completes; its residual is not finite or
    above 16 and it is counted as failed; the others are unaffected."""
    b, dA, dX = make_batch(ctx, [40, 40], [1, 1])
    a = dA.download(np.float64, b.a_elems)
    a[b.a_off[1]:b.a_off[1] + 40 * 41].reshape(40, 41)[:, 5] = 0.0
    dA.upload(a)
    info = b.solve(dA, dX)
    resid, _, _ = b.results()
    assert resid[0] < 16
    assert not (resid[1] <= 16)
    assert info["n_failed"] == 1
    b.close()
    for d in (dA, dX):
        d.free()
