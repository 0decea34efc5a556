"""lu2_amd -- MI355X-native lu2 LU + solve.

Python bindings (ctypes) to liblu2mi.so, the C ABI declared in
include/lu2mi.h and include/lu2mi_seams.h.  The compute runs in the HIP
kernels of lu2_amd/csrc; this module only moves pointers and sizes.  If the
library has not been built, every call raises: there is no CPU fallback.

Mirrors the reference driver's operations (lu2.c):
  randomsetmat (lu2.c:70-85)   -> Context.randomsetmat
  lumcolumn    (lu2.c:1122)    -> Context.lumcolumn_dev / solve()
  mydgemm      (lu2lib.c:2079) -> Context.dgemm_dev
  showresult   (lu2.c:45-60)   -> Context.residual_dev
and the extensions: Context.solve_mixed (FP32 factors + FP64 refinement) and
Batch (many independent systems of uneven size in one graph-captured launch).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblu2mi.so")

PIV_COMPAT = 0
PIV_GEPP = 1
PREC_FP64 = 0
PREC_FP32 = 1
PREC_FP16 = 2
PREC_BF16 = 3
PREC_FP16X2 = 4
LOWP_FP16 = 1
LOWP_BF16 = 2
LOWP_FP16X2 = 3
OPT_PIVMODE = 1
OPT_LEAF_W = 2
OPT_TIMING = 3
OPT_LOOKAHEAD = 4
OPT_SUBPANEL = 5
OPT_TRAIL_KERNEL = 6
OPT_DGEMM_TRAILING = 7
OPT_STRICT_SYNC = 8
OPT_WIDE16 = 9
OPT_BACKSUB128 = 10
BATCH_EAGER = 1
BATCH_NMAX = 512
BATCH_RHSMAX = 8
RESID_NMAX = 131072
# lu2mi_get_subsystem_timing order (LU2MI_T_*)
SUBSYSTEMS = ("swap", "panel", "u12_trsm", "trailing_dgemm", "backsub")

_lib = None


class Lu2miError(RuntimeError):
    pass


class IRInfo(ctypes.Structure):
    """lu2mi_ir_info (include/lu2mi.h)."""
    _fields_ = [("iterations", ctypes.c_int), ("gmres_used", ctypes.c_int), ("gmres_inner", ctypes.c_int),
                ("resid_hpl2_initial", ctypes.c_double), ("resid_hpl2", ctypes.c_double),
                ("t_factor_ms", ctypes.c_double), ("t_refine_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchInfo(ctypes.Structure):
    """lu2mi_batch_info (include/lu2mi.h)."""
    _fields_ = [("max_resid_hpl2", ctypes.c_double), ("max_iterations", ctypes.c_int),
                ("n_failed", ctypes.c_int), ("t_ms", ctypes.c_double), ("nt", ctypes.c_int),
                ("wp_small", ctypes.c_int), ("wp_large", ctypes.c_int), ("split", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load liblu2mi.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise Lu2miError(f"{p} not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(p)
    c_int, c_long, c_double, c_size = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_size_t
    vp = ctypes.c_void_p
    sig = {
        "lu2mi_init": (c_int, [c_int, ctypes.POINTER(vp)]),
        "lu2mi_finalize": (c_int, [vp]),
        "lu2mi_last_error": (ctypes.c_char_p, []),
        "lu2mi_set_option": (c_int, [vp, c_int, c_int]),
        "lu2mi_synchronize": (c_int, [vp]),
        "lu2mi_device_info": (c_int, [vp, ctypes.POINTER(c_int), ctypes.c_char_p, c_int]),
        "lu2mi_malloc": (c_int, [vp, ctypes.POINTER(vp), c_size]),
        "lu2mi_free": (c_int, [vp, vp]),
        "lu2mi_memcpy_h2d": (c_int, [vp, vp, vp, c_size]),
        "lu2mi_memcpy_d2h": (c_int, [vp, vp, vp, c_size]),
        "lu2mi_memcpy_d2d": (c_int, [vp, vp, vp, c_size]),
        "lu2mi_randomsetmat": (c_int, [vp, c_int, c_int, vp, c_long]),
        "lu2mi_lumcolumn_dev": (c_int, [vp, c_int, vp, c_long, c_int, vp, vp]),
        "lu2mi_dgemm_dev": (c_int, [vp, c_int, c_int, c_int, c_double, vp, c_long, vp, c_long,
                                    c_double, vp, c_long]),
        "lu2mi_residual_dev": (c_int, [vp, c_int, vp, c_long, vp, vp]),
        "lu2mi_residual_dev_noalloc": (c_int, [vp, c_int, vp, c_long, vp, vp]),
        "lu2mi_get_timing": (c_int, [vp, vp, c_int]),
        "lu2mi_get_subsystem_timing": (c_int, [vp, vp, c_int]),
        "lu2mi_solve_mixed": (c_int, [vp, c_int, vp, c_long, c_int, c_int, vp, c_int, c_double,
                                      ctypes.POINTER(IRInfo)]),
        "lu2mi_sgemm_dev": (c_int, [vp, c_int, c_int, c_int, ctypes.c_float, vp, c_long, vp, c_long,
                                    ctypes.c_float, vp, c_long]),
        "lu2mi_make_diag_dominant": (c_int, [vp, c_int, vp, c_long]),
        "lu2mi_gemm16_dev": (c_int, [vp, c_int, c_int, c_int, c_int, vp, c_long, vp, c_long, vp, c_long]),
        "lu2mi_batch_create": (c_int, [vp, c_int, vp, vp, vp, vp, vp, c_int, c_int, c_double,
                                       ctypes.POINTER(vp)]),
        "lu2mi_batch_destroy": (c_int, [vp]),
        "lu2mi_batch_generate": (c_int, [vp, vp, c_long]),
        "lu2mi_batch_tune": (c_int, [vp, vp, vp, c_int]),
        "lu2mi_batch_set_config": (c_int, [vp, c_int, c_int, c_int, c_int]),
        "lu2mi_batch_get_config": (c_int, [vp, vp]),
        "lu2mi_batch_solve": (c_int, [vp, vp, vp, c_int, ctypes.POINTER(BatchInfo)]),
        "lu2mi_batch_results": (c_int, [vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = _lib.lu2mi_last_error().decode(errors="replace") if _lib else ""
        raise Lu2miError(f"liblu2mi error {rc}: {msg}")


class DeviceBuffer:
    """A device allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(_lib.lu2mi_malloc(ctx.handle, ctypes.byref(p), ctypes.c_size_t(self.nbytes)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        _check(_lib.lu2mi_memcpy_h2d(self.ctx.handle, self.ptr, a.ctypes.data, a.nbytes))

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        """count elements starting offset bytes into the buffer."""
        out = np.empty(count, dtype=dtype)
        assert offset >= 0 and offset + out.nbytes <= self.nbytes
        _check(_lib.lu2mi_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr + offset, out.nbytes))
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int | None = None) -> None:
        nb = other.nbytes if nbytes is None else nbytes
        _check(_lib.lu2mi_memcpy_d2d(self.ctx.handle, self.ptr, other.ptr, nb))

    def free(self) -> None:
        if self.ptr:
            _check(_lib.lu2mi_free(self.ctx.handle, self.ptr))
            self.ptr = None


class Context:
    """One GPU, one HIP stream, the factorization workspace."""

    def __init__(self, device: int = 0):
        load_library()
        h = ctypes.c_void_p()
        _check(_lib.lu2mi_init(device, ctypes.byref(h)))
        self.handle = h.value

    def close(self) -> None:
        if self.handle:
            _lib.lu2mi_finalize(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, opt: int, value: int) -> None:
        _check(_lib.lu2mi_set_option(self.handle, opt, value))

    def device_info(self):
        ncu = ctypes.c_int()
        arch = ctypes.create_string_buffer(64)
        _check(_lib.lu2mi_device_info(self.handle, ctypes.byref(ncu), arch, 64))
        return ncu.value, arch.value.decode()

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def randomsetmat(self, n: int, seed: int, dA: DeviceBuffer, lda: int) -> None:
        _check(_lib.lu2mi_randomsetmat(self.handle, n, seed, dA.ptr, lda))

    def lumcolumn_dev(self, n, dA, lda, nb, dpiv=None, dx=None) -> None:
        _check(_lib.lu2mi_lumcolumn_dev(self.handle, n, dA.ptr, lda, nb,
                                        dpiv.ptr if dpiv else None, dx.ptr if dx else None))

    def dgemm_dev(self, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, offs=(0, 0, 0)) -> None:
        _check(_lib.lu2mi_dgemm_dev(self.handle, m, n, k, alpha, dA.ptr + offs[0], lda,
                                    dB.ptr + offs[1], ldb, beta, dC.ptr + offs[2], ldc))

    def make_diag_dominant(self, n: int, dA: DeviceBuffer, lda: int) -> None:
        _check(_lib.lu2mi_make_diag_dominant(self.handle, n, dA.ptr, lda))

    def gemm16_dev(self, fmt, m, n, k, dL, ldl, dU, ldu, dC, ldc) -> None:
        """C -= L U with L, U rounded to FP16 (fmt 1) or BF16 (fmt 2) under
        per-128-block power-of-two scales, FP32 accumulation."""
        _check(_lib.lu2mi_gemm16_dev(self.handle, fmt, m, n, k, dL.ptr, ldl, dU.ptr, ldu, dC.ptr, ldc))

    def sgemm_dev(self, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc) -> None:
        _check(_lib.lu2mi_sgemm_dev(self.handle, m, n, k, alpha, dA.ptr, lda, dB.ptr, ldb, beta, dC.ptr, ldc))

    def solve_mixed(self, n, dA, lda, nb, dx, prec=PREC_FP32, maxiter=30, target=1.0,
                    raise_on_noconv=True) -> dict:
        """FP32 factorization + FP64 refinement; dA (FP64, b in column n) is not modified."""
        info = IRInfo()
        rc = _lib.lu2mi_solve_mixed(self.handle, n, dA.ptr, lda, nb, prec, dx.ptr, maxiter, target,
                                    ctypes.byref(info))
        if rc != 0 and not (rc == -7 and not raise_on_noconv):
            _check(rc)
        d = info.as_dict()
        d["converged"] = rc == 0
        return d

    def residual_dev(self, n, dA0, lda, dx) -> np.ndarray:
        out = np.zeros(10)
        _check(_lib.lu2mi_residual_dev(self.handle, n, dA0.ptr, lda, dx.ptr, out.ctypes.data))
        return out

    def subsystem_timing(self) -> dict:
        """Per-subsystem GPU ms, algorithmic flops and bytes of the last
        factorization with OPT_TIMING on (lu2mi_get_subsystem_timing)."""
        out = np.zeros(4 * len(SUBSYSTEMS) + 2)
        _check(_lib.lu2mi_get_subsystem_timing(self.handle, out.ctypes.data, len(out)))
        d = {name: {"ms": out[4 * c], "flops": out[4 * c + 1], "bytes": out[4 * c + 2], "groups": int(out[4 * c + 3])}
             for c, name in enumerate(SUBSYSTEMS)}
        d["wall_ms"] = out[-2]
        d["clock_hz"] = out[-1]
        return d

    def residual_dev_noalloc(self, n, dA0, lda, dx, dout) -> None:
        """The residual figures into dout (device, 10 doubles) without any
        allocation; stream-ordered, not synchronized."""
        _check(_lib.lu2mi_residual_dev_noalloc(self.handle, n, dA0.ptr, lda, dx.ptr, dout.ptr))

    def synchronize(self) -> None:
        _check(_lib.lu2mi_synchronize(self.handle))

    def timing(self) -> dict:
        out = np.zeros(4)
        _check(_lib.lu2mi_get_timing(self.handle, out.ctypes.data, 4))
        return {"gemm_ms": out[0], "gemm_launches": int(out[1]), "gemm_flops": out[2], "wall_ms": out[3]}


def solve(a_aug: np.ndarray, nb: int, pivmode: int = PIV_GEPP, ctx: Context | None = None):
    """Factor lu2's augmented matrix (n x lda, b in column n) on the GPU.

    Returns (x, piv, factored) where factored is lu2's stored factor form.
    """
    own = ctx is None
    ctx = ctx or Context()
    try:
        a = np.ascontiguousarray(a_aug, dtype=np.float64)
        n, lda = a.shape
        bufs = [ctx.alloc(a.nbytes), ctx.alloc(8 * n), ctx.alloc(4 * n)]
        dA, dx, dp = bufs
        try:
            dA.upload(a)
            ctx.set_option(OPT_PIVMODE, pivmode)
            ctx.lumcolumn_dev(n, dA, lda, nb, dp, dx)
            x = dx.download(np.float64, n)
            piv = dp.download(np.int32, n)
            fac = dA.download(np.float64, n * lda).reshape(n, lda)
        finally:
            for b in bufs:
                b.free()
        return x, piv, fac
    finally:
        if own:
            ctx.close()


class Batch:
    """A plan for many independent systems of uneven size (lu2mi_batch_*).

    System s is the augmented row-major [A | B] of n[s] rows and
    n[s] + nrhs[s] columns at element offset a_off[s] of one FP64 pool, with
    leading dimension lda[s]; its solution (n x nrhs, row-major) goes to
    element offset x_off[s] of the solution pool.  By default the systems
    are packed back to back (lda = n + nrhs)."""

    def __init__(self, ctx: Context, n, nrhs, prec: int = PREC_FP64, maxiter: int = 10,
                 target: float = 1.0, lda=None, a_off=None, x_off=None):
        self.ctx = ctx
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.nrhs = np.ascontiguousarray(nrhs, dtype=np.int32)
        ns = len(self.n)
        assert len(self.nrhs) == ns
        self.lda = np.ascontiguousarray(self.n + self.nrhs if lda is None else lda, dtype=np.int64)
        sizes = self.n.astype(np.int64) * self.lda
        self.a_off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)[:-1]]) if a_off is None
                                          else a_off, dtype=np.int64)
        xs = self.n.astype(np.int64) * self.nrhs
        self.x_off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(xs)[:-1]]) if x_off is None
                                          else x_off, dtype=np.int64)
        self.a_elems = int((self.a_off + sizes).max()) if ns else 0
        self.x_elems = int((self.x_off + xs).max()) if ns else 0
        self.prec = prec
        h = ctypes.c_void_p()
        _check(_lib.lu2mi_batch_create(ctx.handle, ns, self.n.ctypes.data, self.nrhs.ctypes.data,
                                       self.a_off.ctypes.data, self.lda.ctypes.data, self.x_off.ctypes.data,
                                       prec, maxiter, target, ctypes.byref(h)))
        self.handle = h.value

    def close(self) -> None:
        if self.handle:
            _check(_lib.lu2mi_batch_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def generate(self, dA: DeviceBuffer, base_seed: int) -> None:
        _check(_lib.lu2mi_batch_generate(self.handle, dA.ptr, base_seed))

    def tune(self, dA: DeviceBuffer, dX: DeviceBuffer, reps: int = 2) -> tuple:
        _check(_lib.lu2mi_batch_tune(self.handle, dA.ptr, dX.ptr, reps))
        return self.config()

    def set_config(self, nt: int, wp_small: int, wp_large: int, split: int) -> None:
        _check(_lib.lu2mi_batch_set_config(self.handle, nt, wp_small, wp_large, split))

    def config(self) -> tuple:
        out = np.zeros(4, dtype=np.int32)
        _check(_lib.lu2mi_batch_get_config(self.handle, out.ctypes.data))
        return tuple(int(v) for v in out)

    def solve(self, dA: DeviceBuffer, dX: DeviceBuffer, eager: bool = False) -> dict:
        info = BatchInfo()
        _check(_lib.lu2mi_batch_solve(self.handle, dA.ptr if dA else None, dX.ptr if dX else None,
                                      BATCH_EAGER if eager else 0, ctypes.byref(info)))
        return info.as_dict()

    def results(self):
        ns = len(self.n)
        resid = np.zeros(ns)
        iters = np.zeros(ns, dtype=np.int32)
        piv = np.zeros(max(1, int(self.n.sum())), dtype=np.int32)
        _check(_lib.lu2mi_batch_results(self.handle, resid.ctypes.data, iters.ctypes.data, piv.ctypes.data))
        return resid, iters, piv[:int(self.n.sum())]
