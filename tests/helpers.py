"""Shared helpers for the parity tests: padded column-major buffers, the
componentwise error bound, and oracle calls on fp64 copies."""
import numpy as np

import oracle

U32 = 2.0 ** -24


def rep_error(precision):
    """Bound on |a*b - computed product| / (|a||b|) from operand splitting.

    fp32: 0 (products of fp32 values are exact inside the MFMA).
    bf16x3: bf16 carries 8 significant bits, so hi = rn(a) has |a - hi| <=
    2^-8 |a| and lo = rn(a - hi) leaves a = hi + lo + r with |r| <= 2^-16 |a|;
    a*b - (hi*hi' + hi*lo' + lo*hi') = lo*lo' + r*b + a*r' - r*r', at most
    3 * 2^-16 |a||b| per product.
    """
    return {"fp32": 0.0, "bf16x3": 3.0 * 2.0 ** -16 * 1.01}[precision]


def gemm_bound(precision, k, alpha, abs_ab, beta, c0):
    """Componentwise bound  (e_rep + (2k + 2) u)|alpha| |A||B| + 4u|beta C| + tiny:
    k products and k additions in fp32 (gamma_2k), plus the epilogue's
    rounding of alpha*acc and of the beta*C addition."""
    e = rep_error(precision) + (2.0 * max(k, 1) + 2.0) * U32
    b = e * abs(alpha) * abs_ab + 4.0 * U32 * np.abs(beta * c0)
    return b + 1e-30


def col_major(rng, rows, cols, ld, dtype=np.float32, fill=None, dist="uniform"):
    """Flat buffer of ld*cols holding a rows x cols column-major matrix; the
    padding rows hold NaN.  Returns (flat buffer, fp64 logical matrix)."""
    buf = np.full(ld * max(cols, 1), np.nan, dtype=dtype)
    view = buf[: ld * cols].reshape(cols, ld).T[:rows]
    if fill is not None:
        view[...] = fill
    elif dist == "uniform":
        view[...] = rng.uniform(-1, 1, size=(rows, cols))
    else:
        view[...] = rng.standard_normal(size=(rows, cols))
    return buf, view.astype(np.float64)


def logical(buf, rows, cols, ld):
    return buf[: ld * cols].reshape(cols, ld).T[:rows]


def oracle_gemm(ta, tb, m, n, k, alpha, abuf, lda, bbuf, ldb, beta, cbuf, ldc):
    """Run the fp64 oracle on fp64 copies; returns the fp64 m x n result."""
    a64 = abuf.astype(np.float64)
    b64 = bbuf.astype(np.float64)
    c64 = cbuf.astype(np.float64)
    info = oracle.dgemm(ta, tb, m, n, k, alpha, a64, lda, b64, ldb, beta, c64, ldc)
    assert info == 0
    return logical(c64, m, n, ldc)


def abs_product(ta, tb, amat, bmat):
    a = np.abs(amat if ta == "N" else amat.T)
    b = np.abs(bmat if tb == "N" else bmat.T)
    return a @ b


def round_bf16(x):
    """fp32 values -> bf16 bit patterns (uint16), round to nearest, ties to
    even, on the fp32 bits: add 0x7fff plus the lowest kept bit, drop 16
    bits.  Finite inputs only."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_to_f64(bits):
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def f16_to_f64(bits):
    return np.asarray(bits, dtype=np.uint16).view(np.float16).astype(np.float64)


def round_f16(x):
    """fp32 values -> f16 bit patterns (numpy's conversion rounds to nearest
    even)."""
    return np.asarray(x, dtype=np.float32).astype(np.float16).view(np.uint16)
