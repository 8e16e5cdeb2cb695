"""mojoblas_amd: MI355X (gfx950) level-3 BLAS hot path behind mojoBLAS's API.

The functions here mirror the reference's src/level3 Mojo functions
(docs/reference.md:264-327), and its level-2 gemv / trsv (src/level2): same names, same Netlib argument order and
meaning, flags as strings compared whole ("N"/"n"/"T"/"t"/"C"/"c", ...), and
the same error behaviour -- an invalid argument prints "<name>: Info <k>"
and returns without touching any output (gemm.mojo:211-213).  They return
the info code (0 on success) so tests can assert it.

Operands are either torch CUDA tensors (device pointers, stream-ordered on
torch's current stream) or numpy arrays (host pointers: copied to HBM, run,
copied back, synchronous).  Storage is column-major with explicit leading
dimensions exactly as in BLAS; a torch/numpy array is used as a flat buffer.
Every call goes through the C-ABI in libmojoblas_hip.so; there is no CPU
fallback.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceMatrix, Stream

PREC_FP32 = 0
PREC_BF16X3 = 1
PREC_FP8 = 2
PREC_BF16 = 3
PRECISIONS = {"fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "fp8": PREC_FP8, "bf16": PREC_BF16}


class HipBlasError(RuntimeError):
    """A runtime (non-Netlib-info) failure reported by the C-ABI."""


def _flag(s):
    # The reference compares whole strings (gemm.mojo:180-197): anything that
    # is not exactly one accepted character must yield the info code, so a
    # multi-character string maps to an invalid char.
    if isinstance(s, bytes):
        s = s.decode()
    return s.encode() if len(s) == 1 else b"\x00"


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ptr_and_kind(x):
    """(address, 'dev'|'host', dtype-name) of an operand buffer."""
    if x is None:
        return None, None, None
    if isinstance(x, DeviceMatrix):
        return x.ptr, "dev", str(x.dtype)
    if _is_torch(x):
        if not x.is_cuda:
            raise TypeError("torch operands must be CUDA tensors; pass numpy arrays for host memory")
        if not x.is_contiguous():
            raise TypeError("torch operands must be contiguous buffers (column-major via explicit ld)")
        return x.data_ptr(), "dev", str(x.dtype).replace("torch.", "")
    if isinstance(x, np.ndarray):
        if not (x.flags["C_CONTIGUOUS"] or x.flags["F_CONTIGUOUS"]):
            raise TypeError("numpy operands must be contiguous buffers")
        return x.ctypes.data, "host", str(x.dtype)
    raise TypeError("unsupported operand type %r" % type(x))


def _stream(stream):
    if stream is not None:
        return ctypes.c_void_p(stream)
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _prec(precision):
    if precision is None:
        return PREC_FP32
    if isinstance(precision, str):
        return PRECISIONS[precision.lower()]
    return int(precision)


def _finish(name, rc):
    if rc > 0:
        print(name + ": Info", rc)
    elif rc < 0:
        raise HipBlasError("%s failed (%d): %s" % (name, rc, _lib.last_error()))
    return rc


def _common(*xs, allow16=False):
    kinds = set()
    dtypes = set()
    for x in xs:
        p, kind, dt = _ptr_and_kind(x)
        if p is not None:
            kinds.add(kind)
            dtypes.add(dt)
    if len(kinds) > 1:
        raise TypeError("operands mix host and device memory")
    if len(dtypes) > 1:
        raise TypeError("operands mix dtypes %s" % sorted(dtypes))
    kind = kinds.pop() if kinds else "host"
    dt = dtypes.pop() if dtypes else "float32"
    if dt in ("float16", "bfloat16") and allow16:
        return kind, dt
    if dt not in ("float32", "float64"):
        raise TypeError("dtype %s not supported (float32, float64%s)" % (dt, ", float16, bfloat16" if allow16 else ""))
    return kind, dt


def _p(x):
    return ctypes.c_void_p(_ptr_and_kind(x)[0]) if x is not None else ctypes.c_void_p(0)


def gemm(trans_a, trans_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, precision=None, stream=None):
    """C := alpha*op(A)*op(B) + beta*C (src/level3/gemm.mojo:56-124).

    precision: "fp32" (default), "bf16x3" or "fp8" for float32 data; float64
    data always runs the fp64 matrix-core kernel.  float16 / bfloat16 data
    (torch CUDA tensors, or numpy float16 host arrays) runs the native 16-bit
    matrix-core kernel: fp32 accumulation, C rounded once to its format;
    `precision` must be None there.
    """
    L = _lib.lib()
    kind, dt = _common(a, b, c, allow16=True)
    ta, tb = _flag(trans_a), _flag(trans_b)
    if dt in ("float16", "bfloat16"):
        if precision is not None:
            raise ValueError("precision applies to float32 data only")
        fn = getattr(L, "mbh_%sgemm_%s" % ("h" if dt == "float16" else "b", kind))
        args = [ta, tb, m, n, k, float(alpha), _p(a), lda, _p(b), ldb, float(beta), _p(c), ldc]
        rc = fn(*(args + [_stream(stream)] if kind == "dev" else args))
    elif dt == "float32":
        if kind == "dev":
            rc = L.mbh_sgemm_dev(ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc, _prec(precision),
                                 _stream(stream))
        else:
            if precision is not None:
                old = L.mbh_set_default_precision(_prec(precision))
                try:
                    rc = L.mbh_sgemm_host(ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)
                finally:
                    if old >= 0:
                        L.mbh_set_default_precision(old)
            else:
                rc = L.mbh_sgemm_host(ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)
    else:
        if kind == "dev":
            rc = L.mbh_dgemm_dev(ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc, _stream(stream))
        else:
            rc = L.mbh_dgemm_host(ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)
    return _finish("gemm", rc)


def _call(name, prefix_args, tail_args, operands, precision, stream, dev_extra_float=(), dev_extra_double=()):
    """Route one level-3 call to the s/d, host/dev entry of the C-ABI.  The
    device entries take the trailing `dev_extra_*` arguments and a stream;
    the host entries take the precision from the library default."""
    L = _lib.lib()
    kind, dt = _common(*operands)
    t = "s" if dt == "float32" else "d"
    fn = getattr(L, "mbh_%s%s_%s" % (t, name, kind))
    args = list(prefix_args) + list(tail_args)
    if kind == "dev":
        extra = (list(dev_extra_float) if t == "s" else list(dev_extra_double)) + [_stream(stream)]
        return _finish(name, fn(*(args + extra)))
    if t == "s" and precision is not None:
        old = L.mbh_set_default_precision(_prec(precision))
        try:
            return _finish(name, fn(*args))
        finally:
            if old >= 0:
                L.mbh_set_default_precision(old)
    return _finish(name, fn(*args))


def trsm(side, uplo, trans_a, diag, m, n, alpha, a, lda, b, ldb, precision=None, refine=1, stream=None):
    """Solve op(A) X = alpha B (side L) or X op(A) = alpha B (side R); X
    overwrites B (src/level3/trsm.mojo:20-275, BLAS semantics).

    float32: `precision` ("fp32" default, "bf16x3") selects the gemm of the
    off-diagonal updates; diagonal 128x128 blocks are solved in fp64;
    refine=1 adds one fp64-residual refinement step (device entry; the host
    entry always refines once).
    """
    ptrs = [_p(a), lda, _p(b), ldb]
    return _call("trsm", [_flag(side), _flag(uplo), _flag(trans_a), _flag(diag), m, n, alpha],
                 ptrs, (a, b), precision, stream, dev_extra_float=[_prec(precision), int(refine)])


def trmm(side, uplo, trans_a, diag, m, n, alpha, a, lda, b, ldb, precision=None, stream=None):
    """B := alpha*op(A)*B (side L) or alpha*B*op(A) (side R), A triangular
    (src/level3/trmm.mojo:20-305)."""
    ptrs = [_p(a), lda, _p(b), ldb]
    return _call("trmm", [_flag(side), _flag(uplo), _flag(trans_a), _flag(diag), m, n, alpha],
                 ptrs, (a, b), precision, stream, dev_extra_float=[_prec(precision)])


def syrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, precision=None, stream=None):
    """C := alpha*A*A^T + beta*C (trans N) or alpha*A^T*A + beta*C on the
    `uplo` triangle (src/level3/syrk.mojo:20-200)."""
    return _call("syrk", [_flag(uplo), _flag(trans), n, k, alpha], [_p(a), lda, beta, _p(c), ldc], (a, c),
                 precision, stream, dev_extra_float=[_prec(precision)])


def syr2k(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc, precision=None, stream=None):
    """C := alpha*(A*B^T + B*A^T) + beta*C (trans N) or alpha*(A^T*B + B^T*A)
    + beta*C on the `uplo` triangle (src/level3/syr2k.mojo:20-190)."""
    return _call("syr2k", [_flag(uplo), _flag(trans), n, k, alpha],
                 [_p(a), lda, _p(b), ldb, beta, _p(c), ldc], (a, b, c), precision, stream,
                 dev_extra_float=[_prec(precision)])


def symm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, precision=None, stream=None):
    """C := alpha*A*B + beta*C (side L) or alpha*B*A + beta*C (side R), A
    symmetric from its `uplo` triangle (src/level3/symm.mojo:20-241)."""
    return _call("symm", [_flag(side), _flag(uplo), m, n, alpha],
                 [_p(a), lda, _p(b), ldb, beta, _p(c), ldc], (a, b, c), precision, stream,
                 dev_extra_float=[_prec(precision)])


def gemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, stream=None):
    """y := alpha*op(A)*x + beta*y (src/level2/gemv.mojo:23-212): trans is
    "N", "T" or "C" (uppercase only, as the reference compares them), x and y
    strided by incx / incy (negative increments as in BLAS)."""
    return _call("gemv", [_flag(trans), m, n, alpha], [_p(a), lda, _p(x), incx, beta, _p(y), incy], (a, x, y),
                 None, stream)


def trsv(uplo, trans, diag, n, a, lda, x, incx, stream=None):
    """x := inv(op(A))*x, A triangular (src/level2/trsv.mojo:20-214): one
    blocked-substitution kernel with fp64 residuals and fp64 diagonal-block
    inverses."""
    return _call("trsv", [_flag(uplo), _flag(trans), _flag(diag), n], [_p(a), lda, _p(x), incx], (a, x),
                 None, stream)


# The reference exports its gemm generations under their own names
# (src/level3/__init__.mojo:29-39); here they are the one GPU gemm.
gemm_v3 = gemm_v6 = gemm_v7 = gemm_v8 = gemm_v9 = gemm_v10 = gemm_v11 = gemm_v12 = gemm
