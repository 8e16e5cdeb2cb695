"""ctypes binding of the FP64 oracle and the restated CPU baseline.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker (or the timed CPU baseline),
never by the product package.

Arrays are numpy, column-major (Fortran order) 2-D views or flat buffers with
explicit leading dimensions, exactly like the BLAS signatures.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_flt = ctypes.c_float
_pd = ctypes.POINTER(ctypes.c_double)
_pf = ctypes.POINTER(ctypes.c_float)
_pl = ctypes.POINTER(ctypes.c_int64)
_cs = ctypes.c_char_p


def build():
    """Compile liboracle.so with the committed Makefile (gcc only)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _bind(L):
    """Declare the C signatures on a loaded build of this library."""
    L.orc_dgemm.argtypes = [_cs, _cs, _i64, _i64, _i64, _dbl, _pd, _i64, _pd, _i64, _dbl, _pd, _i64]
    L.orc_dtrsm.argtypes = [_cs, _cs, _cs, _cs, _i64, _i64, _dbl, _pd, _i64, _pd, _i64]
    L.orc_dtrmm.argtypes = [_cs, _cs, _cs, _cs, _i64, _i64, _dbl, _pd, _i64, _pd, _i64]
    L.orc_dsyrk.argtypes = [_cs, _cs, _i64, _i64, _dbl, _pd, _i64, _dbl, _pd, _i64]
    L.orc_dsymm.argtypes = [_cs, _cs, _i64, _i64, _dbl, _pd, _i64, _pd, _i64, _dbl, _pd, _i64]
    L.orc_dsyr2k.argtypes = [_cs, _cs, _i64, _i64, _dbl, _pd, _i64, _pd, _i64, _dbl, _pd, _i64]
    L.orc_dgemv.argtypes = [_cs, _i64, _i64, _dbl, _pd, _i64, _pd, _i64, _dbl, _pd, _i64]
    L.orc_dtrsv.argtypes = [_cs, _cs, _cs, _i64, _pd, _i64, _pd, _i64]
    for f in (L.orc_dgemm, L.orc_dtrsm, L.orc_dtrmm, L.orc_dsyrk, L.orc_dsymm, L.orc_dsyr2k):
        f.restype = ctypes.c_int
    L.orc_sgemm_sample_f64.argtypes = [
        ctypes.c_char, ctypes.c_char, _i64, _dbl, _pf, _i64, _pf, _i64, _dbl, _pf, _i64, _pl, _pl, _pd]
    L.orc_sgemm_sample_f64.restype = None
    L.orc_sgemm_sample_abs.argtypes = [
        ctypes.c_char, ctypes.c_char, _i64, _pf, _i64, _pf, _i64, _i64, _pl, _pl, _pd]
    L.orc_sgemm_sample_abs.restype = None
    L.orc_fill_f32.argtypes = [_pf, _i64, ctypes.c_uint64, ctypes.c_int, _dbl, _dbl]
    L.orc_fill_f32.restype = None
    L.base_threads.restype = ctypes.c_int
    L.base_sgemm.argtypes = [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, _flt, _pf, _i64, _pf, _i64,
                             _flt, _pf, _i64]
    L.base_sgemm.restype = None
    L.base_strsm_left.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_char, _i64, _i64, _flt, _pf, _i64,
                                  _pf, _i64]
    L.base_strsm_left.restype = None
    L.base_ssyrk_n.argtypes = [ctypes.c_char, _i64, _i64, _flt, _pf, _i64, _flt, _pf, _i64]
    L.base_ssyrk_n.restype = None
    L.base_ssymm.argtypes = [ctypes.c_char, ctypes.c_char, _i64, _i64, _flt, _pf, _i64, _pf, _i64, _flt, _pf,
                             _i64]
    L.base_ssymm.restype = None
    L.base_strmm.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_char, _i64, _i64, _flt, _pf,
                             _i64, _pf, _i64]
    L.base_strmm.restype = None
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


@contextlib.contextmanager
def swapped(path):
    """Run every wrapper below against another build of this library (the
    tests use it for builds with a planted mistake, to show that a check
    rejects them)."""
    global _lib
    old = lib()
    _lib = _bind(ctypes.CDLL(path))
    try:
        yield
    finally:
        _lib = old


def _p(x, t):
    if x is None:
        return None
    assert x.flags["F_CONTIGUOUS"] or x.ndim == 1 or x.size <= 1, "oracle wants Fortran-ordered buffers"
    return x.ctypes.data_as(ctypes.POINTER(t))


def _d(x):
    assert x.dtype == np.float64
    return _p(x, ctypes.c_double)


def _f(x):
    assert x.dtype == np.float32
    return _p(x, ctypes.c_float)


def _s(flag):
    return flag.encode() if isinstance(flag, str) else flag


# ---- FP64 BLAS-semantics routines (in place on b / c, return Netlib info) ----

def dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    return lib().orc_dgemm(_s(ta), _s(tb), m, n, k, alpha, _d(a), lda, _d(b), ldb, beta, _d(c), ldc)


def dtrsm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
    return lib().orc_dtrsm(_s(side), _s(uplo), _s(ta), _s(diag), m, n, alpha, _d(a), lda, _d(b), ldb)


def dtrmm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
    return lib().orc_dtrmm(_s(side), _s(uplo), _s(ta), _s(diag), m, n, alpha, _d(a), lda, _d(b), ldb)


def dsyrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
    return lib().orc_dsyrk(_s(uplo), _s(trans), n, k, alpha, _d(a), lda, beta, _d(c), ldc)


def dsymm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc):
    return lib().orc_dsymm(_s(side), _s(uplo), m, n, alpha, _d(a), lda, _d(b), ldb, beta, _d(c), ldc)


def dsyr2k(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    return lib().orc_dsyr2k(_s(uplo), _s(trans), n, k, alpha, _d(a), lda, _d(b), ldb, beta, _d(c), ldc)


def dgemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
    return lib().orc_dgemv(_s(trans), m, n, alpha, _d(a), lda, _d(x), incx, beta, _d(y), incy)


def dtrsv(uplo, trans, diag, n, a, lda, x, incx):
    return lib().orc_dtrsv(_s(uplo), _s(trans), _s(diag), n, _d(a), lda, _d(x), incx)


ROUTINES = {"gemm": dgemm, "trsm": dtrsm, "trmm": dtrmm, "syrk": dsyrk, "symm": dsymm, "syr2k": dsyr2k,
            "gemv": dgemv, "trsv": dtrsv}


# ---- sampled fp64 pricing of full-size fp32 results ----

def sgemm_sample(ta, tb, k, alpha, a, lda, b, ldb, beta, c0_samples, rows, cols):
    """FP64 value of alpha*op(A)op(B)+beta*C at (rows[s], cols[s]); a, b flat fp32."""
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.empty(rows.size, dtype=np.float64)
    c0 = None if c0_samples is None else np.ascontiguousarray(c0_samples, dtype=np.float32)
    lib().orc_sgemm_sample_f64(_s(ta), _s(tb), k, alpha, _f(a), lda, _f(b), ldb, beta,
                               _f(c0) if c0 is not None else None, rows.size,
                               _p(rows, ctypes.c_int64), _p(cols, ctypes.c_int64), _d(out))
    return out


def sgemm_sample_abs(ta, tb, k, a, lda, b, ldb, rows, cols):
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.empty(rows.size, dtype=np.float64)
    lib().orc_sgemm_sample_abs(_s(ta), _s(tb), k, _f(a), lda, _f(b), ldb, rows.size,
                               _p(rows, ctypes.c_int64), _p(cols, ctypes.c_int64), _d(out))
    return out


def fill_f32(count, seed, dist="uniform", lo=-1.0, hi=1.0):
    x = np.empty(count, dtype=np.float32)
    lib().orc_fill_f32(_f(x), count, seed, 0 if dist == "uniform" else 1, lo, hi)
    return x


# ---- restated mojoBLAS CPU baseline (timed) ----

def base_threads():
    return lib().base_threads()


def base_sgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    lib().base_sgemm(_s(ta), _s(tb), m, n, k, alpha, _f(a), lda, _f(b), ldb, beta, _f(c), ldc)


def base_strsm_left(uplo, trans, diag, m, n, alpha, a, lda, b, ldb):
    lib().base_strsm_left(_s(uplo), _s(trans), _s(diag), m, n, alpha, _f(a), lda, _f(b), ldb)


def base_ssyrk_n(uplo, n, k, alpha, a, lda, beta, c, ldc):
    lib().base_ssyrk_n(_s(uplo), n, k, alpha, _f(a), lda, beta, _f(c), ldc)


def base_ssymm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc):
    lib().base_ssymm(_s(side), _s(uplo), m, n, alpha, _f(a), lda, _f(b), ldb, beta, _f(c), ldc)


def base_strmm(side, uplo, trans, diag, m, n, alpha, a, lda, b, ldb):
    lib().base_strmm(_s(side), _s(uplo), _s(trans), _s(diag), m, n, alpha, _f(a), lda, _f(b), ldb)
