"""Pins the FP64 oracle's level-2 routines (orc_dgemv, orc_dtrsv in
oracle/oracle.c) before the GPU gemv / trsv are checked against them.

* the 13 gemv / trsv known answers of mojoBLAS tests/test_level2.mojo and
  tests/test_level2_extended.mojo (tests/golden/level2_kat.json, written by
  tests/golden/extract_kat.py);
* OpenBLAS through scipy.linalg.blas on random inputs for every flag, positive
  and negative increments, padded leading dimensions, and NaN sentinels in
  the storage the routine must not read (matrix padding, the unused
  triangle, a junk unit diagonal) or write (the gaps of strided vectors);
* the reference's info numbering (gemv.mojo:76-91, trsv.mojo:57-80), with
  the output untouched, and its quick returns;
* closed forms: the trsv residual op(A) x = b, a negative increment equals the
  reversed vector, gemv 'N' and 'T' agree through A^T.
A mutation test shows that the known answers catch a transposed operand.
"""
import itertools
import json
import os

import numpy as np
import pytest
import scipy.linalg.blas as sb

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "level2_kat.json")))


def _almost_equal(got, want, atol, rtol):
    return abs(got - want) <= atol + rtol * abs(want)


def _run_kat(case, routine=None, trans=None):
    arrays = {k: np.array(v, dtype=np.float64) for k, v in case["arrays"].items()}
    args = [arrays[a["array"]] if isinstance(a, dict) else a for a in case["args"]]
    if trans is not None:
        args[0] = trans
    info = oracle.ROUTINES[routine or case["routine"]](*args)
    return info, arrays


@pytest.mark.parametrize("case", KAT["tests"], ids=[t["name"] for t in KAT["tests"]])
def test_known_answers(case):
    info, arrays = _run_kat(case)
    assert info == 0
    for e in case["expects"]:
        got = arrays[e["array"]][e["index"]]
        assert _almost_equal(got, e["value"], e["atol"], e["rtol"]), (case["name"], e, got)


def test_kat_count():
    assert KAT["count"] == 13
    assert sum(t["routine"] == "gemv" for t in KAT["tests"]) == 8
    assert sum(t["routine"] == "trsv" for t in KAT["tests"]) == 5


def test_mutation_transposed_operand_is_caught():
    """Running the non-square gemv known answers with the other transpose
    (a planted mistake) must miss an expected value."""
    caught = 0
    for case in KAT["tests"]:
        if case["routine"] != "gemv" or case["args"][1] == case["args"][2]:
            continue
        other = "T" if case["args"][0] == "N" else "N"
        # the swapped call sees the same buffers with m, n exchanged
        case2 = json.loads(json.dumps(case))
        case2["args"][1], case2["args"][2] = case["args"][2], case["args"][1]
        case2["args"][5] = case2["args"][1]
        info, arrays = _run_kat(case2, trans=other)
        assert info == 0
        if any(not _almost_equal(arrays[e["array"]][e["index"]], e["value"], e["atol"], e["rtol"])
               for e in case["expects"]):
            caught += 1
    assert caught >= 3


# ---------------------------------------------------------------------------

def strided(rng, length, inc, fill_gaps=np.nan):
    """Vector of `length` logical elements stored with increment |inc| (gaps
    hold `fill_gaps`); returns (buffer, logical values in BLAS order)."""
    buf = np.full(abs(inc) * max(length - 1, 0) + 1, fill_gaps)
    vals = rng.uniform(-1, 1, length)
    for i, v in enumerate(vals):
        buf[i * inc if inc > 0 else (length - 1 - i) * (-inc)] = v
    return buf, vals


def logical(buf, length, inc):
    return np.array([buf[i * inc if inc > 0 else (length - 1 - i) * (-inc)] for i in range(length)])


INCS = [1, 2, -1, -3]
SCALARS = [(1.0, 0.0), (1.5, 0.5), (-0.5, 1.0), (0.0, 2.0), (2.0, 1.0)]


@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_gemv_vs_openblas(trans):
    rng = np.random.default_rng(50)
    for m, n in itertools.product([1, 3, 17], [1, 4, 16]):
        for incx, incy in itertools.product(INCS, [1, 3, -2]):
            alpha, beta = SCALARS[(m + n + incx + incy) % len(SCALARS)]
            lda = m + (m % 3)
            abuf = np.full(lda * n, np.nan)
            A = rng.uniform(-1, 1, (m, n))
            abuf.reshape(n, lda).T[:m] = A
            lx, ly = (n, m) if trans == "N" else (m, n)
            xbuf, xv = strided(rng, lx, incx)
            ybuf, yv = strided(rng, ly, incy)
            got = ybuf.copy()
            assert oracle.dgemv(trans, m, n, alpha, abuf, lda, xbuf, incx, beta, got, incy) == 0
            opA = A if trans == "N" else A.T
            want = alpha * (opA @ xv) + beta * yv
            np.testing.assert_allclose(logical(got, ly, incy), want, rtol=1e-13, atol=1e-13)
            # scipy (OpenBLAS) on the same storage
            ys = sb.dgemv(alpha, A, np.nan_to_num(xbuf), beta=beta, y=np.nan_to_num(ybuf), incx=incx, incy=incy,
                          trans=0 if trans == "N" else 1)
            np.testing.assert_allclose(logical(ys, ly, incy), logical(got, ly, incy), rtol=1e-13, atol=1e-13)
            # gaps of y untouched (NaN stays NaN)
            gaps = np.ones(ybuf.size, bool)
            for i in range(ly):
                gaps[i * incy if incy > 0 else (ly - 1 - i) * (-incy)] = False
            assert np.isnan(got[gaps]).all()


def test_gemv_beta_zero_does_not_read_y():
    rng = np.random.default_rng(51)
    A = np.asfortranarray(rng.uniform(-1, 1, (5, 4)))
    x = rng.uniform(-1, 1, 4)
    y = np.full(5, np.nan)
    assert oracle.dgemv("N", 5, 4, 1.0, A.ravel(order="F"), 5, x, 1, 0.0, y, 1) == 0
    np.testing.assert_allclose(y, A @ x, rtol=1e-14)
    y = np.full(5, np.nan)
    assert oracle.dgemv("N", 5, 4, 0.0, A.ravel(order="F"), 5, x, 1, 0.0, y, 1) == 0
    assert (y == 0).all()


def test_gemv_n_and_t_agree_through_transpose():
    rng = np.random.default_rng(52)
    A = rng.uniform(-1, 1, (6, 9))
    x = rng.uniform(-1, 1, 9)
    y1, y2 = np.zeros(6), np.zeros(6)
    oracle.dgemv("N", 6, 9, 1.0, np.asfortranarray(A).ravel(order="F"), 6, x, 1, 0.0, y1, 1)
    oracle.dgemv("T", 9, 6, 1.0, np.asfortranarray(A.T).ravel(order="F"), 9, x, 1, 0.0, y2, 1)
    np.testing.assert_allclose(y1, y2, rtol=1e-14)


def tri_matrix(rng, n, upper, unit_junk):
    A = rng.uniform(-1, 1, (n, n)) / n
    A[np.diag_indices(n)] = rng.uniform(1, 2, n)
    T = np.triu(A) if upper else np.tril(A)
    store = T.copy()
    store[np.tril_indices(n, -1) if upper else np.triu_indices(n, 1)] = np.nan  # unused triangle
    if unit_junk:
        store[np.diag_indices(n)] = 99.0
        T[np.diag_indices(n)] = 1.0
    return T, store


@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product("UL", "NTC", "NU")))
def test_trsv_vs_openblas_and_residual(uplo, trans, diag):
    rng = np.random.default_rng(53)
    for n in (1, 2, 5, 33):
        T, store = tri_matrix(rng, n, uplo == "U", diag == "U")
        lda = n + 2
        abuf = np.full(lda * n, np.nan)
        abuf.reshape(n, lda).T[:n] = store
        opT = T if trans == "N" else T.T
        for incx in INCS:
            xbuf, bv = strided(rng, n, incx)
            got = xbuf.copy()
            assert oracle.dtrsv(uplo, trans, diag, n, abuf, lda, got, incx) == 0
            xv = logical(got, n, incx)
            np.testing.assert_allclose(opT @ xv, bv, rtol=1e-12, atol=1e-12)  # residual
            if incx == 1:
                xs = sb.dtrsv(np.nan_to_num(store, nan=0.0) if diag == "N" else np.where(np.isnan(store), 0.0, store),
                              bv.copy(), lower=int(uplo == "L"), trans=0 if trans == "N" else 1,
                              diag=int(diag == "U"))
                np.testing.assert_allclose(xv, xs, rtol=1e-12, atol=1e-12)
            else:  # negative / strided increments = the contiguous solve of the logical vector
                ref = bv.copy()
                oracle.dtrsv(uplo, trans, diag, n, abuf, lda, ref, 1)
                np.testing.assert_allclose(xv, ref, rtol=0, atol=0)
            gaps = np.ones(xbuf.size, bool)
            for i in range(n):
                gaps[i * incx if incx > 0 else (n - 1 - i) * (-incx)] = False
            assert np.isnan(got[gaps]).all()


GEMV_INFO = [
    (("n", 2, 2, 1.0, 2, 1, 0.0, 1), 1),  # gemv.mojo:77 compares uppercase only
    (("No", 2, 2, 1.0, 2, 1, 0.0, 1), 1),
    (("N", -1, 2, 1.0, 2, 1, 0.0, 1), 2),
    (("N", 2, -1, 1.0, 2, 1, 0.0, 1), 3),
    (("N", 3, 2, 1.0, 2, 1, 0.0, 1), 6),
    (("T", 2, 2, 1.0, 2, 0, 0.0, 1), 8),
    (("T", 2, 2, 1.0, 2, 1, 0.0, 0), 11),
]


@pytest.mark.parametrize("args,info", GEMV_INFO)
def test_gemv_info_codes_leave_y_untouched(args, info):
    trans, m, n, alpha, lda, incx, beta, incy = args
    a, x = np.ones(16), np.ones(8)
    y = np.full(8, 7.0)
    assert oracle.dgemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy) == info
    assert (y == 7.0).all()


TRSV_INFO = [
    (("X", "N", "N", 2, 2, 1), 1), (("U", "X", "N", 2, 2, 1), 2), (("U", "N", "X", 2, 2, 1), 3),
    (("U", "N", "N", -1, 2, 1), 4), (("U", "N", "N", 3, 2, 1), 6), (("U", "N", "N", 2, 2, 0), 8),
    (("u", "t", "n", 2, 1, 1), 6),  # lowercase flags are valid for trsv (trsv.mojo:58-70)
]


@pytest.mark.parametrize("args,info", TRSV_INFO)
def test_trsv_info_codes_leave_x_untouched(args, info):
    uplo, trans, diag, n, lda, incx = args
    a, x = np.ones(16), np.full(8, 7.0)
    assert oracle.dtrsv(uplo, trans, diag, n, a, lda, x, incx) == info
    assert (x == 7.0).all()


def test_quick_returns():
    a, x = np.ones(4), np.ones(2)
    y = np.full(2, np.nan)
    assert oracle.dgemv("N", 0, 2, 1.0, a, 1, x, 1, 0.0, y, 1) == 0 and np.isnan(y).all()
    assert oracle.dgemv("N", 2, 2, 0.0, a, 2, x, 1, 1.0, y, 1) == 0 and np.isnan(y).all()
    assert oracle.dtrsv("U", "N", "N", 0, a, 1, y, 1) == 0 and np.isnan(y).all()
