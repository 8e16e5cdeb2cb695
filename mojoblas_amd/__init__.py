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


# Precisions the gemm-based routines take; PREC_BF16 is the batched gemm's
# bf16-input mode only.
GEMM_PRECISIONS = (PREC_FP32, PREC_BF16X3, PREC_FP8)


def _prec(precision):
    if precision is None:
        return PREC_FP32
    if isinstance(precision, str):
        if precision.lower() not in PRECISIONS:
            raise ValueError("unknown precision %r (%s)" % (precision, ", ".join(sorted(PRECISIONS))))
        p = PRECISIONS[precision.lower()]
    else:
        p = int(precision)
    if p not in GEMM_PRECISIONS:
        raise ValueError("precision %r is not available here (fp32, bf16x3, fp8); bf16 is the batched gemm's "
                         "input mode" % (precision,))
    return p


class _ThreadPrecision:
    """Sets the host entries' precision for the calling thread only
    (mbh_set_thread_precision) for the duration of one call."""

    def __init__(self, precision):
        self.p = None if precision is None else _prec(precision)

    def __enter__(self):
        if self.p is not None:
            old = _lib.lib().mbh_set_thread_precision(self.p)
            if old < -1:
                raise HipBlasError("mbh_set_thread_precision failed (%d): %s" % (old, _lib.last_error()))
            self.old = old

    def __exit__(self, *exc):
        if self.p is not None:
            _lib.lib().mbh_set_thread_precision(self.old)


def _numel(x):
    if isinstance(x, DeviceMatrix):
        return x.ld * max(x.cols, 0)
    if _is_torch(x):
        return x.numel()
    return x.size


def _mat_need(rows, cols, ld):
    """Elements a column-major rows x cols matrix with leading dimension ld spans."""
    return 0 if rows <= 0 or cols <= 0 else ld * (cols - 1) + rows


def _vec_need(length, inc):
    return 0 if length <= 0 else (length - 1) * abs(inc) + 1


def _check_len(routine, pairs):
    """Raise before any native call when an operand buffer is shorter than
    the storage the call reads or writes (the C-ABI takes bare pointers)."""
    for name, x, need in pairs:
        if x is not None and need > 0 and _numel(x) < need:
            raise ValueError("%s: operand %s holds %d elements; the call reaches %d" % (routine, name, _numel(x), need))


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
    if ta in (b"N", b"n", b"T", b"t", b"C", b"c") and tb in (b"N", b"n", b"T", b"t", b"C", b"c") and min(m, n, k) >= 0:
        ar, ac = (m, k) if ta in (b"N", b"n") else (k, m)
        br, bc = (k, n) if tb in (b"N", b"n") else (n, k)
        reads_ab = alpha != 0 and k > 0 and m > 0 and n > 0
        _check_len("gemm", [("A", a, _mat_need(ar, ac, lda) if reads_ab and lda >= max(1, ar) else 0),
                            ("B", b, _mat_need(br, bc, ldb) if reads_ab and ldb >= max(1, br) else 0),
                            ("C", c, _mat_need(m, n, ldc) if ldc >= max(1, m) else 0)])
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
            with _ThreadPrecision(precision):
                rc = L.mbh_sgemm_host(ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)
    else:
        if kind == "dev":
            rc = L.mbh_dgemm_dev(ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc, _stream(stream))
        else:
            rc = L.mbh_dgemm_host(ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)
    return _finish("gemm", rc)


def _call(name, prefix_args, tail_args, operands, precision, stream, dev_extra_float=(), dev_extra_double=(),
          lengths=()):
    """Route one level-3 call to the s/d, host/dev entry of the C-ABI.  The
    device entries take the trailing `dev_extra_*` arguments and a stream;
    the host entries take `precision` as the calling thread's setting for the
    call.  `lengths`: (name, operand, elements the call reaches) checked
    first."""
    L = _lib.lib()
    kind, dt = _common(*operands)
    _check_len(name, lengths)
    t = "s" if dt == "float32" else "d"
    fn = getattr(L, "mbh_%s%s_%s" % (t, name, kind))
    args = list(prefix_args) + list(tail_args)
    if kind == "dev":
        extra = (list(dev_extra_float) if t == "s" else list(dev_extra_double)) + [_stream(stream)]
        return _finish(name, fn(*(args + extra)))
    if t == "s":
        with _ThreadPrecision(precision):
            return _finish(name, fn(*args))
    return _finish(name, fn(*args))


def _trxm_lengths(side, m, n, alpha, a, lda, b, ldb):
    if min(m, n) <= 0:
        return []
    na = m if side in ("L", "l") else n
    return [("A", a, _mat_need(na, na, lda) if alpha != 0 and lda >= max(1, na) else 0),
            ("B", b, _mat_need(m, n, ldb) if ldb >= max(1, m) else 0)]


def _syrk_lengths(trans, n, k, alpha, ops, c, ldc):
    if n <= 0 or k < 0:
        return []
    ar, ac = (n, k) if trans in ("N", "n") else (k, n)
    out = [(nm, x, _mat_need(ar, ac, ld) if alpha != 0 and k > 0 and ld >= max(1, ar) else 0) for nm, x, ld in ops]
    return out + [("C", c, _mat_need(n, n, ldc) if ldc >= max(1, n) else 0)]


def trsm(side, uplo, trans_a, diag, m, n, alpha, a, lda, b, ldb, precision=None, refine=1, stream=None):
    """Solve op(A) X = alpha B (side L) or X op(A) = alpha B (side R); X
    overwrites B (src/level3/trsm.mojo:20-275, BLAS semantics).

    float32: diagonal 128x128 blocks are solved in fp64.  refine=1 (default;
    the host entry always uses it) runs the whole solve on fp64 copies of
    alpha*B and A's triangle and rounds X to fp32 once; refine=0 solves in
    fp32 in place, `precision` ("fp32" default, "bf16x3") selecting the gemm
    of the off-diagonal updates.
    """
    ptrs = [_p(a), lda, _p(b), ldb]
    return _call("trsm", [_flag(side), _flag(uplo), _flag(trans_a), _flag(diag), m, n, alpha],
                 ptrs, (a, b), precision, stream, dev_extra_float=[_prec(precision), int(refine)],
                 lengths=_trxm_lengths(side, m, n, alpha, a, lda, b, ldb))


def trmm(side, uplo, trans_a, diag, m, n, alpha, a, lda, b, ldb, precision=None, stream=None):
    """B := alpha*op(A)*B (side L) or alpha*B*op(A) (side R), A triangular
    (src/level3/trmm.mojo:20-305)."""
    ptrs = [_p(a), lda, _p(b), ldb]
    return _call("trmm", [_flag(side), _flag(uplo), _flag(trans_a), _flag(diag), m, n, alpha],
                 ptrs, (a, b), precision, stream, dev_extra_float=[_prec(precision)],
                 lengths=_trxm_lengths(side, m, n, alpha, a, lda, b, ldb))


def syrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, precision=None, stream=None):
    """C := alpha*A*A^T + beta*C (trans N) or alpha*A^T*A + beta*C on the
    `uplo` triangle (src/level3/syrk.mojo:20-200)."""
    return _call("syrk", [_flag(uplo), _flag(trans), n, k, alpha], [_p(a), lda, beta, _p(c), ldc], (a, c),
                 precision, stream, dev_extra_float=[_prec(precision)],
                 lengths=_syrk_lengths(trans, n, k, alpha, [("A", a, lda)], c, ldc))


def syr2k(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc, precision=None, stream=None):
    """C := alpha*(A*B^T + B*A^T) + beta*C (trans N) or alpha*(A^T*B + B^T*A)
    + beta*C on the `uplo` triangle (src/level3/syr2k.mojo:20-190)."""
    return _call("syr2k", [_flag(uplo), _flag(trans), n, k, alpha],
                 [_p(a), lda, _p(b), ldb, beta, _p(c), ldc], (a, b, c), precision, stream,
                 dev_extra_float=[_prec(precision)],
                 lengths=_syrk_lengths(trans, n, k, alpha, [("A", a, lda), ("B", b, ldb)], c, ldc))


def symm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, precision=None, stream=None):
    """C := alpha*A*B + beta*C (side L) or alpha*B*A + beta*C (side R), A
    symmetric from its `uplo` triangle (src/level3/symm.mojo:20-241)."""
    na = m if side in ("L", "l") else n
    ok = min(m, n) >= 0
    reads = ok and alpha != 0
    return _call("symm", [_flag(side), _flag(uplo), m, n, alpha],
                 [_p(a), lda, _p(b), ldb, beta, _p(c), ldc], (a, b, c), precision, stream,
                 dev_extra_float=[_prec(precision)],
                 lengths=[("A", a, _mat_need(na, na, lda) if reads and m * n > 0 and lda >= max(1, na) else 0),
                          ("B", b, _mat_need(m, n, ldb) if reads and ldb >= max(1, m) else 0),
                          ("C", c, _mat_need(m, n, ldc) if ok and ldc >= max(1, m) else 0)])


def gemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, stream=None):
    """y := alpha*op(A)*x + beta*y (src/level2/gemv.mojo:23-212): trans is
    "N", "T" or "C" (uppercase only, as the reference compares them), x and y
    strided by incx / incy (negative increments as in BLAS)."""
    ok = min(m, n) > 0 and incx != 0 and incy != 0 and trans in ("N", "T", "C")
    lenx, leny = (n, m) if trans == "N" else (m, n)
    return _call("gemv", [_flag(trans), m, n, alpha], [_p(a), lda, _p(x), incx, beta, _p(y), incy], (a, x, y),
                 None, stream,
                 lengths=[("A", a, _mat_need(m, n, lda) if ok and alpha != 0 and lda >= max(1, m) else 0),
                          ("x", x, _vec_need(lenx, incx) if ok and alpha != 0 else 0),
                          ("y", y, _vec_need(leny, incy) if ok else 0)])


def trsv(uplo, trans, diag, n, a, lda, x, incx, stream=None):
    """x := inv(op(A))*x, A triangular (src/level2/trsv.mojo:20-214): one
    blocked-substitution kernel with fp64 residuals and fp64 diagonal-block
    inverses."""
    ok = n > 0 and incx != 0
    return _call("trsv", [_flag(uplo), _flag(trans), _flag(diag), n], [_p(a), lda, _p(x), incx], (a, x),
                 None, stream,
                 lengths=[("A", a, _mat_need(n, n, lda) if ok and lda >= max(1, n) else 0),
                          ("x", x, _vec_need(n, incx) if ok else 0)])


# The reference exports its gemm generations under their own names
# (src/level3/__init__.mojo:29-39); here they are the one GPU gemm.
gemm_v3 = gemm_v6 = gemm_v7 = gemm_v8 = gemm_v9 = gemm_v10 = gemm_v11 = gemm_v12 = gemm
