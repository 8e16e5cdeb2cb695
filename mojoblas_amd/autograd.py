"""torch.autograd.Function front-ends for gemm and trsm, forward and backward
on the HIP kernels of libmojoblas_hip.so (mbh_sgemm_dev / mbh_sgemm_backward,
mbh_strsm_dev / mbh_strsm_backward).

Tensors follow torch conventions (2-D, row-major, float32, CUDA).  A
contiguous r x c tensor is, as a flat buffer, the column-major c x r matrix
of its transpose, so
    C = alpha * A @ B                  runs as  C^T = alpha * B^T A^T,
    X = alpha * op(A)^-1 @ B (left)    runs as  X^T = alpha * B^T op(A)^-T,
i.e. a column-major right-side solve on the same buffers with the triangle
flipped, and the backward entries are called in that same frame.
"""
import ctypes

import torch

from . import _lib, _prec, HipBlasError


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _check(rc, what):
    if rc != 0:
        raise HipBlasError("%s failed (%d): %s" % (what, rc, _lib.last_error()))


def _as_operand(t, name):
    if not (t.is_cuda and t.dtype == torch.float32 and t.dim() == 2):
        raise TypeError("%s must be a 2-D float32 CUDA tensor" % name)
    return t.contiguous()


class GemmFunction(torch.autograd.Function):
    """C = alpha * A @ B with A (m x k), B (k x n)."""

    @staticmethod
    def forward(ctx, A, B, alpha=1.0, precision="fp32"):
        A = _as_operand(A, "A")
        B = _as_operand(B, "B")
        m, k = A.shape
        k2, n = B.shape
        if k2 != k:
            raise ValueError("inner dimensions differ: %s @ %s" % (tuple(A.shape), tuple(B.shape)))
        C = torch.empty(m, n, device=A.device, dtype=torch.float32)
        prec = _prec(precision)
        # column-major: C^T (n x m) = B^T (n x k, ld n) * A^T (k x m, ld k)
        rc = _lib.lib().mbh_sgemm_dev(b"N", b"N", n, m, k, float(alpha), _vp(B), max(n, 1), _vp(A), max(k, 1), 0.0,
                                      _vp(C), max(n, 1), prec, _stream())
        _check(rc, "mbh_sgemm_dev")
        ctx.save_for_backward(A, B)
        ctx.alpha, ctx.prec = float(alpha), prec
        return C

    @staticmethod
    def backward(ctx, dC):
        A, B = ctx.saved_tensors
        m, k = A.shape
        n = B.shape[1]
        dC = dC.contiguous()
        need_a, need_b = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        dA = torch.empty_like(A) if need_a else None
        dB = torch.empty_like(B) if need_b else None
        # In the column-major frame the forward was C' = A' B' with A' = B^T
        # (n x k), B' = A^T (k x m): grad(A') is dB's buffer, grad(B') is dA's.
        null = ctypes.c_void_p(0)
        rc = _lib.lib().mbh_sgemm_backward(b"N", b"N", n, m, k, ctx.alpha, _vp(B), max(n, 1), _vp(A), max(k, 1),
                                           _vp(dC), max(n, 1), _vp(dB) if need_b else null, max(n, 1),
                                           _vp(dA) if need_a else null, max(k, 1), ctx.prec, _stream())
        _check(rc, "mbh_sgemm_backward")
        return dA, dB, None, None


class TrsmFunction(torch.autograd.Function):
    """X = alpha * op(A)^-1 @ B (side "L") or alpha * B @ op(A)^-1 (side "R"),
    A triangular (only its `uplo` triangle is read; diag "U" = unit)."""

    @staticmethod
    def forward(ctx, A, B, side="L", uplo="L", trans="N", diag="N", alpha=1.0, precision="fp32"):
        A = _as_operand(A, "A")
        B = _as_operand(B, "B")
        rows, cols = B.shape
        na = A.shape[0]
        if A.shape[1] != na or na != (rows if side.upper() == "L" else cols):
            raise ValueError("A must be square and match B's %s dimension" % ("row" if side.upper() == "L" else
                                                                                "column"))
        # Column-major frame: X^T = alpha B^T op(A)^-T (left) or alpha op(A)^-T B^T
        # (right); the buffer of A is A^T, whose triangle is the other one.
        cside = b"R" if side.upper() == "L" else b"L"
        cuplo = b"U" if uplo.upper() == "L" else b"L"
        ctr = trans.upper().encode()
        cdiag = diag.upper().encode()
        X = B.clone()
        prec = _prec(precision)
        rc = _lib.lib().mbh_strsm_dev(cside, cuplo, ctr, cdiag, cols, rows, float(alpha), _vp(A), max(na, 1), _vp(X),
                                      max(cols, 1), prec, 1, _stream())
        _check(rc, "mbh_strsm_dev")
        ctx.save_for_backward(A, X)
        ctx.args = (cside, cuplo, ctr, cdiag, float(alpha), prec)
        return X

    @staticmethod
    def backward(ctx, dX):
        A, X = ctx.saved_tensors
        cside, cuplo, ctr, cdiag, alpha, prec = ctx.args
        rows, cols = X.shape
        na = A.shape[0]
        dX = dX.contiguous()
        dB = torch.empty_like(X)
        dA = torch.empty_like(A) if ctx.needs_input_grad[0] else None
        rc = _lib.lib().mbh_strsm_backward(cside, cuplo, ctr, cdiag, cols, rows, alpha, _vp(A), max(na, 1), _vp(X),
                                           max(cols, 1), _vp(dX), max(cols, 1), _vp(dB), max(cols, 1),
                                           _vp(dA) if dA is not None else ctypes.c_void_p(0), max(na, 1), prec,
                                           _stream())
        _check(rc, "mbh_strsm_backward")
        return dA, dB, None, None, None, None, None, None


def gemm(A, B, alpha=1.0, precision="fp32"):
    """Differentiable alpha * A @ B on the HIP gemm kernels."""
    return GemmFunction.apply(A, B, alpha, precision)


def trsm(A, B, side="L", uplo="L", trans="N", diag="N", alpha=1.0, precision="fp32"):
    """Differentiable triangular solve on the HIP trsm kernels."""
    return TrsmFunction.apply(A, B, side, uplo, trans, diag, alpha, precision)
