"""Pins the FP64 oracle (oracle/oracle.c) before anything is checked against it.

* the 31 known answers of mojoBLAS tests/test_level3.mojo (via
  tests/golden/level3_kat.json, produced by tests/golden/extract_kat.py);
* OpenBLAS through scipy.linalg.blas (the library tests/reference.c:88-161
  binds for its level-3 cases, reference.c:519-672) on random inputs for
  every flag combination, padded leading dimensions and NaN sentinels in
  storage the routine must not read or write;
* the Netlib info codes of the reference's validation blocks;
* closed forms: trsm undoes trmm; identity operands.
A mutation test shows the OpenBLAS comparison fails for a transposed operand.
"""
import itertools
import json
import os

import numpy as np
import pytest
import scipy.linalg.blas as sb

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "level3_kat.json")))


def _almost_equal(got, want, atol, rtol):
    return abs(got - want) <= atol + rtol * abs(want)


@pytest.mark.parametrize("case", KAT["tests"], ids=[t["name"] for t in KAT["tests"]])
def test_known_answers(case):
    arrays = {k: np.array(v, dtype=np.float64) for k, v in case["arrays"].items()}
    args = [arrays[a["array"]] if isinstance(a, dict) else a for a in case["args"]]
    info = oracle.ROUTINES[case["routine"]](*args)
    assert info == 0
    for e in case["expects"]:
        got = arrays[e["array"]][e["index"]]
        assert _almost_equal(got, e["value"], e["atol"], e["rtol"]), (case["name"], e, got)


def test_kat_count():
    # 31 level-3 tests in the reference file, 26 of them on the hot path
    # routines (gemm/trsm/trmm/syrk/symm) and 2 on syr2k.
    assert KAT["count"] == 31
    assert sum(1 for t in KAT["tests"] if t["routine"] != "syr2k") == 29


# ---------------------------------------------------------------------------
# Padded column-major storage with NaN sentinels.

def padded(rng, rows, cols, ld, fill=None):
    buf = np.full(ld * max(cols, 1), np.nan)
    view = buf[: ld * cols].reshape(cols, ld).T[:rows]  # column-major view
    if fill is None:
        view[...] = rng.uniform(-1, 1, size=(rows, cols))
    else:
        view[...] = fill
    return buf, view


def fview(buf, rows, cols, ld):
    return buf[: ld * cols].reshape(cols, ld).T[:rows]


DIMS = [0, 1, 3, 17]
SCALARS = [(1.0, 0.0), (1.5, 0.5), (-0.5, 1.0), (0.0, 2.0), (0.0, 0.0), (2.0, 1.0)]


@pytest.mark.parametrize("ta,tb", list(itertools.product("NTC", "NTc")))
def test_gemm_vs_openblas(ta, tb):
    rng = np.random.default_rng(1)
    for m, n, k in itertools.product(DIMS, DIMS, [0, 1, 5, 16]):
        for pad in (0, 3):
            alpha, beta = SCALARS[(m + n + k + pad) % len(SCALARS)]
            ar, ac = (m, k) if ta == "N" else (k, m)
            br, bc = (k, n) if tb == "N" else (n, k)
            lda, ldb, ldc = max(1, ar) + pad, max(1, br) + pad, max(1, m) + pad
            abuf, av = padded(rng, ar, ac, lda)
            bbuf, bv = padded(rng, br, bc, ldb)
            cbuf, cv = padded(rng, m, n, ldc, fill=np.nan if beta == 0 else None)
            c0 = cv.copy()
            info = oracle.dgemm(ta, tb, m, n, k, alpha, abuf, lda, bbuf, ldb, beta, cbuf, ldc)
            assert info == 0
            if m == 0 or n == 0:
                continue
            want = sb.dgemm(alpha, np.asfortranarray(av) if ar and ac else np.zeros((ar, ac)),
                            np.asfortranarray(bv) if br and bc else np.zeros((br, bc)),
                            beta=beta, c=np.nan_to_num(c0) if beta != 0 else None,
                            trans_a=int(ta != "N"), trans_b=int(tb not in "N")) if k > 0 else None
            if want is None:  # k == 0: C := beta*C
                want = np.zeros((m, n)) if beta == 0 else beta * c0
            got = fview(cbuf, m, n, ldc)
            np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
            # padding rows never written
            full = cbuf[: ldc * n].reshape(n, ldc).T
            assert np.isnan(full[m:]).all()


def _tri_matrix(rng, n, upper, cond=True):
    a = rng.uniform(-1, 1, size=(n, n)) / max(n, 1)
    a[np.diag_indices(n)] = rng.uniform(1, 2, size=n)
    junk = np.triu(np.full((n, n), np.nan), 1) if not upper else np.tril(np.full((n, n), np.nan), -1)
    keep = np.triu(a) if upper else np.tril(a)
    return np.where(np.isnan(junk), np.nan, keep)


SIDE_UPLO_TRANS_DIAG = list(itertools.product("LR", "UL", "NT", "NU"))


@pytest.mark.parametrize("side,uplo,ta,diag", SIDE_UPLO_TRANS_DIAG)
def test_trsm_vs_openblas(side, uplo, ta, diag):
    rng = np.random.default_rng(2)
    for m, n in [(1, 1), (5, 3), (17, 9), (4, 16)]:
        na = m if side == "L" else n
        a = _tri_matrix(rng, na, uplo == "U")
        if diag == "U":
            a[np.diag_indices(na)] = np.nan  # unit diagonal must not be read
        lda = na + 2
        abuf = np.full(lda * na, np.nan)
        fview(abuf, na, na, lda)[...] = a
        alpha = 1.5
        bbuf, bv = padded(rng, m, n, m + 1)
        b0 = bv.copy()
        assert oracle.dtrsm(side, uplo, ta, diag, m, n, alpha, abuf, lda, bbuf, m + 1) == 0
        a_clean = np.nan_to_num(a, nan=0.0)
        want = sb.dtrsm(alpha, np.asfortranarray(a_clean), np.asfortranarray(b0), side=int(side == "R"),
                        lower=int(uplo == "L"), trans_a=int(ta == "T"), diag=int(diag == "U"))
        np.testing.assert_allclose(fview(bbuf, m, n, m + 1), want, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("side,uplo,ta,diag", SIDE_UPLO_TRANS_DIAG)
def test_trmm_vs_openblas_and_inverse(side, uplo, ta, diag):
    rng = np.random.default_rng(3)
    for m, n in [(1, 1), (6, 4), (13, 17)]:
        na = m if side == "L" else n
        a = _tri_matrix(rng, na, uplo == "U")
        if diag == "U":
            a[np.diag_indices(na)] = np.nan
        abuf = np.full(na * na, np.nan)
        fview(abuf, na, na, na)[...] = a
        bbuf, bv = padded(rng, m, n, m)
        b0 = bv.copy()
        assert oracle.dtrmm(side, uplo, ta, diag, m, n, -0.75, abuf, na, bbuf, m) == 0
        a_clean = np.nan_to_num(a, nan=0.0)
        want = sb.dtrmm(-0.75, np.asfortranarray(a_clean), np.asfortranarray(b0), side=int(side == "R"),
                        lower=int(uplo == "L"), trans_a=int(ta == "T"), diag=int(diag == "U"))
        np.testing.assert_allclose(fview(bbuf, m, n, m), want, rtol=1e-12, atol=1e-13)
        # closed form: trsm with the same operand undoes trmm
        assert oracle.dtrsm(side, uplo, ta, diag, m, n, 1.0 / -0.75, abuf, na, bbuf, m) == 0
        np.testing.assert_allclose(fview(bbuf, m, n, m), b0, rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize("uplo,trans", list(itertools.product("UL", "NT")))
def test_syrk_syr2k_vs_openblas(uplo, trans):
    rng = np.random.default_rng(4)
    for n, k in [(1, 1), (7, 3), (16, 21), (5, 0)]:
        ar, ac = (n, k) if trans == "N" else (k, n)
        lda = max(1, ar) + 1
        abuf, av = padded(rng, ar, ac, lda)
        bbuf, bv = padded(rng, ar, ac, lda)
        for alpha, beta in [(1.0, 0.0), (0.5, 2.0), (0.0, 0.5)]:
            cbuf, cv = padded(rng, n, n, n)
            tri = np.triu(np.ones((n, n), bool)) if uplo == "U" else np.tril(np.ones((n, n), bool))
            cv[~tri] = np.nan  # other triangle: neither read nor written
            c0 = np.nan_to_num(cv.copy())
            assert oracle.dsyrk(uplo, trans, n, k, alpha, abuf, lda, beta, cbuf, n) == 0
            got = fview(cbuf, n, n, n)
            if k > 0:
                want = sb.dsyrk(alpha, np.asfortranarray(av), beta=beta, c=np.asfortranarray(c0),
                                trans=int(trans == "T"), lower=int(uplo == "L"))
            else:
                want = beta * c0
            np.testing.assert_allclose(got[tri], want[tri], rtol=1e-12, atol=1e-13)
            assert np.isnan(got[~tri]).all()

            cbuf, cv = padded(rng, n, n, n)
            cv[~tri] = np.nan
            c0 = np.nan_to_num(cv.copy())
            assert oracle.dsyr2k(uplo, trans, n, k, alpha, abuf, lda, bbuf, lda, beta, cbuf, n) == 0
            got = fview(cbuf, n, n, n)
            if k > 0:
                want = sb.dsyr2k(alpha, np.asfortranarray(av), np.asfortranarray(bv), beta=beta,
                                 c=np.asfortranarray(c0), trans=int(trans == "T"), lower=int(uplo == "L"))
            else:
                want = beta * c0
            np.testing.assert_allclose(got[tri], want[tri], rtol=1e-12, atol=1e-13)
            assert np.isnan(got[~tri]).all()


@pytest.mark.parametrize("side,uplo", list(itertools.product("LR", "UL")))
def test_symm_vs_openblas(side, uplo):
    rng = np.random.default_rng(5)
    for m, n in [(1, 1), (6, 9), (17, 4)]:
        na = m if side == "L" else n
        s = rng.uniform(-1, 1, size=(na, na))
        s = s + s.T
        stored = np.triu(s) if uplo == "U" else np.tril(s)
        stored[(np.tril(np.ones((na, na)), -1) if uplo == "U" else np.triu(np.ones((na, na)), 1)) > 0] = np.nan
        abuf = np.full(na * na, np.nan)
        fview(abuf, na, na, na)[...] = stored
        bbuf, bv = padded(rng, m, n, m)
        bv[0, 0] = 0.0  # exact zero in B: the reference's zero-skip must not matter
        for alpha, beta in [(1.0, 0.0), (-2.0, 0.5)]:
            cbuf, cv = padded(rng, m, n, m, fill=np.nan if beta == 0 else None)
            c0 = np.nan_to_num(cv.copy())
            assert oracle.dsymm(side, uplo, m, n, alpha, abuf, na, bbuf, m, beta, cbuf, m) == 0
            want = sb.dsymm(alpha, np.asfortranarray(s), np.asfortranarray(bv), beta=beta, c=np.asfortranarray(c0),
                            side=int(side == "R"), lower=int(uplo == "L"))
            np.testing.assert_allclose(fview(cbuf, m, n, m), want, rtol=1e-12, atol=1e-13)


# Netlib info codes in the reference's validation order.
INFO_CASES = [
    # gemm.mojo:179-213
    ("gemm", ("X", "N", 2, 2, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 1),
    ("gemm", ("No", "N", 2, 2, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 1),
    ("gemm", ("N", "x", 2, 2, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 2),
    ("gemm", ("N", "N", -1, 2, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 3),
    ("gemm", ("N", "N", 2, -1, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 4),
    ("gemm", ("N", "N", 2, 2, -1, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 5),
    ("gemm", ("N", "N", 2, 2, 2, 1.0, "A4", 1, "B4", 2, 0.0, "C4", 2), 8),
    ("gemm", ("T", "N", 2, 2, 3, 1.0, "A4", 2, "B4", 3, 0.0, "C4", 2), 8),
    ("gemm", ("N", "N", 2, 2, 2, 1.0, "A4", 2, "B4", 1, 0.0, "C4", 2), 10),
    ("gemm", ("N", "T", 2, 3, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 10),
    ("gemm", ("N", "N", 2, 2, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 1), 13),
    # trsm.mojo:62-89 (trmm.mojo:62-89 identical)
    ("trsm", ("X", "U", "N", "N", 2, 2, 1.0, "A4", 2, "B4", 2), 1),
    ("trsm", ("L", "X", "N", "N", 2, 2, 1.0, "A4", 2, "B4", 2), 2),
    ("trsm", ("L", "U", "X", "N", 2, 2, 1.0, "A4", 2, "B4", 2), 3),
    ("trsm", ("L", "U", "N", "X", 2, 2, 1.0, "A4", 2, "B4", 2), 4),
    ("trsm", ("L", "U", "N", "N", -1, 2, 1.0, "A4", 2, "B4", 2), 5),
    ("trsm", ("L", "U", "N", "N", 2, -1, 1.0, "A4", 2, "B4", 2), 6),
    ("trsm", ("L", "U", "N", "N", 2, 2, 1.0, "A4", 1, "B4", 2), 9),
    ("trsm", ("R", "U", "N", "N", 1, 3, 1.0, "A4", 2, "B4", 1), 9),
    ("trsm", ("L", "U", "N", "N", 2, 2, 1.0, "A4", 2, "B4", 1), 11),
    ("trmm", ("l", "u", "t", "x", 2, 2, 1.0, "A4", 2, "B4", 2), 4),
    ("trmm", ("L", "U", "N", "N", 2, 2, 1.0, "A4", 2, "B4", 1), 11),
    # syrk.mojo:60-83
    ("syrk", ("X", "N", 2, 2, 1.0, "A4", 2, 0.0, "C4", 2), 1),
    ("syrk", ("U", "X", 2, 2, 1.0, "A4", 2, 0.0, "C4", 2), 2),
    ("syrk", ("U", "N", -1, 2, 1.0, "A4", 2, 0.0, "C4", 2), 3),
    ("syrk", ("U", "N", 2, -1, 1.0, "A4", 2, 0.0, "C4", 2), 4),
    ("syrk", ("U", "N", 2, 2, 1.0, "A4", 1, 0.0, "C4", 2), 7),
    ("syrk", ("U", "T", 2, 3, 1.0, "A4", 2, 0.0, "C4", 2), 7),
    ("syrk", ("U", "N", 2, 2, 1.0, "A4", 2, 0.0, "C4", 1), 10),
    # symm.mojo:68-86
    ("symm", ("X", "U", 2, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 1),
    ("symm", ("L", "X", 2, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 2),
    ("symm", ("L", "U", -1, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 3),
    ("symm", ("L", "U", 2, -1, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 4),
    ("symm", ("R", "U", 2, 3, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 2), 7),
    ("symm", ("L", "U", 2, 2, 1.0, "A4", 2, "B4", 1, 0.0, "C4", 2), 9),
    ("symm", ("L", "U", 2, 2, 1.0, "A4", 2, "B4", 2, 0.0, "C4", 1), 12),
]


@pytest.mark.parametrize("routine,args,info", INFO_CASES)
def test_info_codes_leave_outputs_untouched(routine, args, info):
    bufs = {}
    real = []
    for a in args:
        if isinstance(a, str) and a[:1] in "ABC" and a[1:].isdigit():
            bufs[a] = np.full(int(a[1:]) * 4, 7.0)
            real.append(bufs[a])
        else:
            real.append(a)
    assert oracle.ROUTINES[routine](*real) == info
    for v in bufs.values():
        assert (v == 7.0).all()


def test_quick_returns():
    c = np.array([1.0, np.nan, 3.0, 4.0])
    a = np.full(4, np.nan)
    # alpha == 0, beta == 0: C zeroed without reading A, B or C (gemm.mojo:218-226)
    assert oracle.dgemm("N", "N", 2, 2, 2, 0.0, a, 2, a, 2, 0.0, c, 2) == 0
    assert (c == 0).all()
    # trmm alpha == 0 zeroes B (trmm.mojo:94-98)
    b = np.full(4, np.nan)
    assert oracle.dtrmm("L", "U", "N", "N", 2, 2, 0.0, a, 2, b, 2) == 0
    assert (b == 0).all()
    # syrk alpha == 0 and beta == 1 returns without touching C (syrk.mojo:85)
    c = np.full(4, np.nan)
    assert oracle.dsyrk("U", "N", 2, 2, 0.0, a, 2, 1.0, c, 2) == 0
    assert np.isnan(c).all()


def test_mutation_transposed_operand_is_caught():
    """The OpenBLAS comparison has teeth: a transposed operand fails it."""
    rng = np.random.default_rng(7)
    a = np.asfortranarray(rng.uniform(-1, 1, (5, 5)))
    b = np.asfortranarray(rng.uniform(-1, 1, (5, 5)))
    c = np.zeros(25)
    oracle.dgemm("T", "N", 5, 5, 5, 1.0, a.ravel(order="F"), 5, b.ravel(order="F"), 5, 0.0, c, 5)
    want = sb.dgemm(1.0, a, b)
    assert not np.allclose(c.reshape(5, 5, order="F"), want)


def test_fill_is_reproducible_and_distributed():
    x = oracle.fill_f32(200000, 42, "uniform", -1, 1)
    y = oracle.fill_f32(200000, 42, "uniform", -1, 1)
    assert (x == y).all() and x.min() >= -1 and x.max() < 1 and abs(x.mean()) < 0.01
    z = oracle.fill_f32(200000, 43, "normal")
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert not (oracle.fill_f32(10, 1) == oracle.fill_f32(10, 2)).all()


def test_sample_matches_full():
    rng = np.random.default_rng(8)
    for ta, tb in itertools.product("NT", "NT"):
        m, n, k = 7, 5, 9
        a = rng.standard_normal((m, k) if ta == "N" else (k, m)).astype(np.float32)
        b = rng.standard_normal((k, n) if tb == "N" else (n, k)).astype(np.float32)
        full = (a.astype(np.float64) if ta == "N" else a.astype(np.float64).T) @ (
            b.astype(np.float64) if tb == "N" else b.astype(np.float64).T)
        rows, cols = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
        af, bf = np.asfortranarray(a).ravel(order="F"), np.asfortranarray(b).ravel(order="F")
        s = oracle.sgemm_sample(ta, tb, k, 2.0, af, a.shape[0], bf, b.shape[0], 0.0, None, rows.ravel(), cols.ravel())
        np.testing.assert_allclose(s.reshape(m, n), 2.0 * full, rtol=1e-14)


def test_cpu_baseline_matches_oracle():
    rng = np.random.default_rng(9)
    for ta, tb in itertools.product("NT", "NT"):
        for m, n, k in [(33, 70, 1030), (130, 17, 5)]:
            a = np.asfortranarray(rng.uniform(-1, 1, (m, k) if ta == "N" else (k, m)).astype(np.float32))
            b = np.asfortranarray(rng.uniform(-1, 1, (k, n) if tb == "N" else (n, k)).astype(np.float32))
            c = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
            want = c.astype(np.float64).ravel(order="F").copy()
            oracle.dgemm(ta, tb, m, n, k, 1.5, a.astype(np.float64).ravel(order="F"), a.shape[0],
                         b.astype(np.float64).ravel(order="F"), b.shape[0], 0.5, want, m)
            got = c.ravel(order="F").copy()
            oracle.base_sgemm(ta, tb, m, n, k, 1.5, a.ravel(order="F"), a.shape[0], b.ravel(order="F"), b.shape[0],
                              0.5, got, m)
            np.testing.assert_allclose(got, want, rtol=0, atol=2e-4 * np.sqrt(k))
