"""GPU parity of the native 16-bit gemm (hgemm: f16, bgemm: bf16) through the
C-ABI, device and host entries.

The kernel multiplies 16-bit operands on the MFMA with fp32 accumulation
and rounds alpha*acc + beta*C once, to nearest even, into C's format.
  * Integer-valued data (every partial sum exact in fp32): the result must
    equal round16(alpha*A*B + beta*C) bit for bit, at every tile size and
    transpose, with NaN sentinels in the ld padding left untouched.
  * Random data: |C - C_fp64| <= u16 |C_fp64| + (1 + u16) ((2k + 2) u32
    |alpha| |A||B| + 4 u32 |beta C|) + tiny, u16 = 2^-11 (f16) or 2^-8
    (bf16), u32 = 2^-24, C_fp64 from the fp64 oracle on the decoded inputs.
  * The reference's gemm known answers (small integers) in both formats.
"""
import ctypes
import itertools
import json
import os

import numpy as np
import pytest

from tests import helpers as H

torch = pytest.importorskip("torch")
import mojoblas_amd as mb
from mojoblas_amd import _lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "level3_kat.json")))

FMTS = ["f16", "bf16"]
U16 = {"f16": 2.0 ** -11, "bf16": 2.0 ** -8}
TINY = {"f16": 2.0 ** -25, "bf16": 2.0 ** -134}
NAN16 = {"f16": 0x7E00, "bf16": 0x7FC0}


def enc(fmt, x):
    return H.round_f16(x) if fmt == "f16" else H.round_bf16(x)


def dec(fmt, bits):
    return H.f16_to_f64(bits) if fmt == "f16" else H.bf16_to_f64(bits)


def tdev(fmt, bits):
    t = torch.from_numpy(np.ascontiguousarray(bits).view(np.int16)).cuda()
    return t.view(torch.float16 if fmt == "f16" else torch.bfloat16)


def tbits(t):
    return t.cpu().view(torch.int16).numpy().view(np.uint16)


def buf16(fmt, vals, rows, cols, ld):
    """Flat 16-bit buffer of ld*cols with a rows x cols column-major matrix
    of `vals` (fp64, representable) and NaN in the padding rows."""
    b = np.full(ld * max(cols, 1), NAN16[fmt], np.uint16)
    if rows and cols:
        b[: ld * cols].reshape(cols, ld).T[:rows] = enc(fmt, vals)
    return b


def run(fmt, entry, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """One call; returns (rc, C bits)."""
    if entry == "dev":
        A, B, C = tdev(fmt, a), tdev(fmt, b), tdev(fmt, c)
        rc = mb.gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
        torch.cuda.synchronize()
        return rc, tbits(C)
    c = c.copy()
    L = _lib.lib()
    fn = L.mbh_hgemm_host if fmt == "f16" else L.mbh_bgemm_host
    P = lambda x: ctypes.c_void_p(x.ctypes.data)
    rc = fn(ta.encode(), tb.encode(), m, n, k, alpha, P(a), lda, P(b), ldb, beta, P(c), ldc)
    return rc, c


def dims(t, rows, cols):
    return (rows, cols) if t == "N" else (cols, rows)


def int_case(rng, fmt, ta, tb, m, n, k, pad=3, lim=4):
    ar, ac = dims(ta, m, k)
    br, bc = dims(tb, k, n)
    av = rng.integers(-lim, lim + 1, size=(ar, ac)).astype(np.float64)
    bv = rng.integers(-lim, lim + 1, size=(br, bc)).astype(np.float64)
    cv = rng.integers(-lim, lim + 1, size=(m, n)).astype(np.float64)
    lda, ldb, ldc = ar + pad, br + pad, m + pad
    return (av, bv, cv, lda, ldb, ldc, buf16(fmt, av, ar, ac, lda), buf16(fmt, bv, br, bc, ldb),
            buf16(fmt, cv, m, n, ldc))


def expect_exact(fmt, ta, tb, alpha, beta, av, bv, cv):
    opa = av if ta == "N" else av.T
    opb = bv if tb == "N" else bv.T
    exact = alpha * (opa @ opb) + beta * cv
    assert np.abs(exact).max(initial=0) < 2.0 ** 24
    return enc(fmt, exact.astype(np.float32))


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("case", [t for t in KAT["tests"] if t["routine"] == "gemm"], ids=lambda t: t["name"])
def test_known_answers(case, fmt):
    arrays = {name: enc(fmt, np.array(v, np.float32)) for name, v in case["arrays"].items()}
    darr = {name: tdev(fmt, v) for name, v in arrays.items()}
    args = [darr[a["array"]] if isinstance(a, dict) else a for a in case["args"]]
    assert mb.gemm(*args) == 0
    torch.cuda.synchronize()
    for e in case["expects"]:
        got = dec(fmt, tbits(darr[e["array"]]))[e["index"]]
        assert got == e["value"], (case["name"], e, got)


SHAPES = [(1, 1, 1), (7, 13, 5), (128, 128, 16), (129, 65, 33), (64, 300, 257), (300, 257, 190)]
SCALARS = [(1.0, 0.0), (-2.0, 1.0), (0.5, -0.5)]


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_integer_data_bit_exact(ta, tb, fmt):
    rng = np.random.default_rng(21)
    for (m, n, k), (alpha, beta) in itertools.product(SHAPES, SCALARS):
        av, bv, cv, lda, ldb, ldc, a, b, c = int_case(rng, fmt, ta, tb, m, n, k)
        rc, got = run(fmt, "dev", ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
        assert rc == 0
        want = expect_exact(fmt, ta, tb, alpha, beta, av, bv, cv)
        view = got[: ldc * n].reshape(n, ldc).T
        assert (view[:m] == want).all(), (m, n, k, alpha, beta)
        assert (view[m:] == NAN16[fmt]).all(), "ldc padding written"


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_in_place_operands_bit_exact(ta, tb, fmt):
    # k % 64 == 0 and ld % 8 == 0: a k-contiguous operand (op(A) when A is
    # transposed, op(B)^T when B is not) is read in place, with its rows past
    # the last clamped.  The other one is read in place too (k-rows in LDS,
    # transposing fragment reads) when its row count is a multiple of 8
    # (8, 136, 72, 264, 520, 1000, 1016, 64), else packed (1, 129, 65, 300,
    # 257).  A base one element off the 16-byte alignment, or k = 96, falls
    # back to packing.
    rng = np.random.default_rng(26)
    for (m, n, k) in [(1, 1, 64), (129, 65, 64), (300, 257, 128), (64, 300, 256), (65, 33, 96), (8, 8, 64),
                      (136, 72, 64), (264, 520, 128), (1000, 1016, 192)]:
        for shift in (0, 1):
            av, bv, cv, lda, ldb, ldc, a, b, c = int_case(rng, fmt, ta, tb, m, n, k, pad=8)
            a = np.concatenate([np.zeros(shift, np.uint16), a])
            b = np.concatenate([np.zeros(shift, np.uint16), b])
            A, B, C = tdev(fmt, a)[shift:], tdev(fmt, b)[shift:], tdev(fmt, c)
            assert mb.gemm(ta, tb, m, n, k, 1.0, A, lda, B, ldb, 1.0, C, ldc) == 0
            torch.cuda.synchronize()
            got = tbits(C)
            want = expect_exact(fmt, ta, tb, 1.0, 1.0, av, bv, cv)
            view = got[: ldc * n].reshape(n, ldc).T
            assert (view[:m] == want).all(), (m, n, k, shift)
            assert (view[m:] == NAN16[fmt]).all()


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("shape,ta,tb", [((2048, 1024, 128), "N", "N"), ((2048, 1024, 128), "T", "T"),
                                         ((2048, 1024, 96), "T", "N"), ((4096, 4096, 64), "N", "N"),
                                         ((4096, 4096, 192), "T", "N"), ((4096, 4096, 72), "T", "N"),
                                         ((2048, 1024, 64), "N", "T"), ((4104, 4088, 128), "N", "T"),
                                         ((4104, 4088, 128), "T", "T"), ((2056, 1032, 192), "N", "N")])
def test_large_tiles_bit_exact(shape, ta, tb, fmt):
    # 2048 x 1024 (and 2056 x 1032) take the 128 x 128 tiles, 4096^2 (and
    # 4104 x 4088) the 256 x 256 ones; with k a multiple of 64 (ld padded by
    # 8) every operand is read in place (those whose rows are contiguous
    # through transposing LDS reads; 4104, 4088, 2056, 1032: a last tile
    # with 8 valid rows, chunks past them clamped), with k = 96 or 72 both
    # are packed (a k-tail inside the last 64-deep k-tile).
    m, n, k = shape
    rng = np.random.default_rng(22)
    av, bv, cv, lda, ldb, ldc, a, b, c = int_case(rng, fmt, ta, tb, m, n, k, pad=8)
    rc, got = run(fmt, "dev", ta, tb, m, n, k, 1.0, a, lda, b, ldb, 1.0, c, ldc)
    assert rc == 0
    want = expect_exact(fmt, ta, tb, 1.0, 1.0, av, bv, cv)
    assert (got[: ldc * n].reshape(n, ldc).T[:m] == want).all()


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("entry", ["dev", "host"])
@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_random_data_vs_oracle(ta, tb, entry, fmt):
    rng = np.random.default_rng(23)
    for (m, n, k) in [(129, 65, 33), (300, 257, 190), (64, 300, 1000)]:
        ar, ac = dims(ta, m, k)
        br, bc = dims(tb, k, n)
        lda, ldb, ldc = ar + 2, br + 1, m + 3
        a = buf16(fmt, dec(fmt, enc(fmt, rng.uniform(-1, 1, (ar, ac)))), ar, ac, lda)
        b = buf16(fmt, dec(fmt, enc(fmt, rng.uniform(-1, 1, (br, bc)))), br, bc, ldb)
        c = buf16(fmt, dec(fmt, enc(fmt, rng.uniform(-1, 1, (m, n)))), m, n, ldc)
        alpha, beta = 1.5, 0.5
        rc, got = run(fmt, entry, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
        assert rc == 0
        a64, b64, c64 = (np.where(x == NAN16[fmt], 0, x) for x in (a, b, c))
        ref = H.oracle_gemm(ta, tb, m, n, k, alpha, dec(fmt, a64), lda, dec(fmt, b64), ldb, beta, dec(fmt, c64),
                            ldc)
        got_v = dec(fmt, got[: ldc * n].reshape(n, ldc).T[:m])
        absab = H.abs_product(ta, tb, dec(fmt, H.logical(a64, ar, ac, lda)), dec(fmt, H.logical(b64, br, bc, ldb)))
        c0 = dec(fmt, H.logical(c64, m, n, ldc))
        acc_bound = (2.0 * k + 2.0) * H.U32 * abs(alpha) * absab + 4.0 * H.U32 * np.abs(beta * c0)
        bound = U16[fmt] * np.abs(ref) + (1 + U16[fmt]) * acc_bound + TINY[fmt]
        err = np.abs(got_v - ref)
        assert (err <= bound).all(), (m, n, k, float((err / bound).max()))
        assert (got[: ldc * n].reshape(n, ldc).T[m:] == NAN16[fmt]).all()


@pytest.mark.parametrize("fmt", FMTS)
def test_scale_paths_and_special_values(fmt):
    m, n = 37, 19
    rng = np.random.default_rng(24)
    cv = rng.integers(-8, 9, size=(m, n)).astype(np.float64)
    c = buf16(fmt, cv, m, n, m)
    a = buf16(fmt, np.ones((m, 4)), m, 4, m)
    b = buf16(fmt, np.ones((4, n)), 4, n, 4)
    # alpha == 0: C := beta*C without reading A or B; k == 0 likewise.
    rc, got = run(fmt, "dev", "N", "N", m, n, 4, 0.0, a, m, b, 4, 2.0, c, m)
    assert rc == 0 and (dec(fmt, got) == 2 * cv.T.ravel()).all()
    rc, got = run(fmt, "dev", "N", "N", m, n, 0, 1.0, a, m, b, 1, -1.0, c, m)
    assert rc == 0 and (dec(fmt, got) == -cv.T.ravel()).all()
    # beta == 0 never reads C: NaN in C disappears.
    cn = np.full(m * n, NAN16[fmt], np.uint16)
    rc, got = run(fmt, "dev", "N", "N", m, n, 0, 1.0, a, m, b, 1, 0.0, cn, m)
    assert rc == 0 and (got == 0).all()
    rc, got = run(fmt, "dev", "N", "N", m, n, 4, 1.0, a, m, b, 4, 0.0, cn, m)
    assert rc == 0 and (dec(fmt, got) == 4.0).all()
    # Inf times an exact zero gives NaN (no zero-skip, as in BLAS); an f16
    # result past 65520 rounds to inf.
    av = np.zeros((m, 4))
    av[0, 0] = np.inf
    ai = buf16(fmt, av, m, 4, m)
    bz = buf16(fmt, np.zeros((4, n)), 4, n, 4)
    rc, got = run(fmt, "dev", "N", "N", m, n, 4, 1.0, ai, m, bz, 4, 0.0, c, m)
    res = dec(fmt, got).reshape(n, m).T
    assert rc == 0 and np.isnan(res[0]).all() and (res[1:] == 0).all()
    if fmt == "f16":
        big = buf16(fmt, np.full((m, 1), 256.0), m, 1, m)
        rc, got = run(fmt, "dev", "N", "N", m, n, 1, 1.0, big, m, big[:n], 1, 0.0, c, m)
        assert rc == 0 and np.isposinf(dec(fmt, got)).all()
    # m == 0: quick return, nothing touched.
    rc, got = run(fmt, "dev", "N", "N", 0, n, 4, 1.0, a, 1, b, 4, 0.0, c, 1)
    assert rc == 0 and (got == c).all()


def test_python_host_float16_arrays():
    rng = np.random.default_rng(25)
    m, n, k = 70, 40, 50
    a = rng.integers(-3, 4, size=(m, k)).astype(np.float16)
    b = rng.integers(-3, 4, size=(k, n)).astype(np.float16)
    c = np.zeros((m, n), np.float16, order="F")
    af, bf = np.asfortranarray(a), np.asfortranarray(b)
    assert mb.gemm("N", "N", m, n, k, 1.0, af, m, bf, k, 0.0, c, m) == 0
    assert (c.astype(np.float64) == a.astype(np.float64) @ b.astype(np.float64)).all()
