"""Pins the oracle functions that the reference's known answers do not reach,
and shows that every check here (and the OpenBLAS comparisons of the FP64
routines) rejects a build of the oracle with a planted mistake.

What pins what:
* sgemm_sample_abs: numpy |op(A)| @ |op(B)| at sampled entries, all four
  transposes, padded leading dimensions with NaN in the padding;
* sgemm_sample with alpha and beta != 0: numpy alpha*op(A)op(B) + beta*C0;
* fill_f32: the published first output of splitmix64 (Steele, Lea, Flood
  2014; state 0 -> 0xE220A8397B1DCDAF) and a Kolmogorov-Smirnov test of the
  normal draw;
* the restated CPU baseline (oracle/baseline.c), against the pinned FP64
  routines on the same fp32 inputs: base_sgemm (config 1 itself: 512^3,
  alpha 1.5, beta 0.5), base_strsm_left (all uplo / trans / diag),
  base_ssyrk_n, base_ssymm (all side / uplo, plus the reference's skip of
  alpha*temp2 where B(l,j) == 0, symm.mojo:114,133, as a closed form) and
  base_strmm (all 16 side / uplo / trans / diag), unused triangles NaN.
The registry at the end requires every public function of the oracle
package to be named here or in the other oracle test files.
"""
import inspect
import itertools

import numpy as np
import pytest
from scipy import stats

import oracle
import oracle_mutants
import test_oracle
import test_oracle_level2

U32 = 2.0 ** -24


def _padded32(rng, rows, cols, ld, extra_cols=0, dist="normal"):
    """fp32 flat buffer of ld*(cols+extra_cols) holding a rows x cols matrix,
    NaN elsewhere; returns (buffer, fp64 logical matrix)."""
    buf = np.full(ld * (cols + extra_cols), np.nan, dtype=np.float32)
    view = buf[: ld * cols].reshape(cols, ld).T[:rows]
    view[...] = rng.standard_normal((rows, cols)) if dist == "normal" else rng.uniform(-1, 1, (rows, cols))
    return buf, view.astype(np.float64)


def _op(ta, x):
    return x if ta == "N" else x.T


def _sample_entries(rng, m, n, count):
    rows = np.concatenate([[0, m - 1, 0, m - 1], rng.integers(0, m, count)])
    cols = np.concatenate([[0, 0, n - 1, n - 1], rng.integers(0, n, count)])
    return rows, cols


# ---------------------------------------------------------------------------
# sampled pricing of full-size results

def check_sample_abs():
    rng = np.random.default_rng(11)
    for ta, tb in itertools.product("NT", "NT"):
        m, n, k = 37, 29, 53
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        big = max(m, n, k)
        abuf, a = _padded32(rng, ar, ac, ar + 3, extra_cols=big)
        bbuf, b = _padded32(rng, br, bc, br + 5, extra_cols=big)
        rows, cols = _sample_entries(rng, m, n, 200)
        got = oracle.sgemm_sample_abs(ta, tb, k, abuf, ar + 3, bbuf, br + 5, rows, cols)
        want = (np.abs(_op(ta, a)) @ np.abs(_op(tb, b)))[rows, cols]
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)


def check_sample_alpha_beta():
    rng = np.random.default_rng(12)
    for ta, tb in itertools.product("NT", "NT"):
        m, n, k = 31, 41, 19
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        big = max(m, n, k)
        abuf, a = _padded32(rng, ar, ac, ar + 2, extra_cols=big)
        bbuf, b = _padded32(rng, br, bc, br + 1, extra_cols=big)
        c = rng.uniform(-1, 1, (m, n)).astype(np.float32)
        rows, cols = _sample_entries(rng, m, n, 150)
        got = oracle.sgemm_sample(ta, tb, k, 1.5, abuf, ar + 2, bbuf, br + 1, -0.75, c[rows, cols], rows, cols)
        want = (1.5 * (_op(ta, a) @ _op(tb, b)) - 0.75 * c.astype(np.float64))[rows, cols]
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def test_sample_abs_pinned_and_mutants_rejected(tmp_path):
    oracle_mutants.assert_rejects(tmp_path, check_sample_abs, [
        ("oracle.c", "fabs(opa_f(transa, a, lda, i, l))", "(opa_f(transa, a, lda, i, l))"),
        ("oracle.c", "fabs(opb_f(transb, b, ldb, l, j))", "fabs(opb_f(transb, b, ldb, l, i))"),
        ("oracle.c", "(double)a[i + l * lda] : (double)a[l + i * lda]", "(double)a[l + i * lda] : (double)a[i + l * lda]"),
    ])


def test_sample_alpha_beta_pinned_and_mutants_rejected(tmp_path):
    oracle_mutants.assert_rejects(tmp_path, check_sample_alpha_beta, [
        ("oracle.c", "beta * (double)c0[s]", "(double)c0[s]"),
        ("oracle.c", "out[s] = alpha * acc + cv;", "out[s] = acc + cv;"),
        ("oracle.c", "(double)b[l + j * ldb] : (double)b[j + l * ldb]", "(double)b[j + l * ldb] : (double)b[l + j * ldb]"),
    ])


# ---------------------------------------------------------------------------
# input generator

def check_fill():
    # splitmix64 from state 0 gives 0xE220A8397B1DCDAF first (the published
    # test vector); u01 keeps its top 53 bits and centres them in their cell.
    want = np.float32(((0xE220A8397B1DCDAF >> 11) + 0.5) / 2.0 ** 53)
    assert oracle.fill_f32(1, 0, "uniform", 0.0, 1.0)[0] == want
    lo, hi = -3.0, 5.0
    assert oracle.fill_f32(1, 0, "uniform", lo, hi)[0] == np.float32(lo + (hi - lo) * float(want))
    z = oracle.fill_f32(200000, 5, "normal").astype(np.float64)
    assert stats.kstest(z, "norm").pvalue > 1e-3
    u = oracle.fill_f32(200000, 6, "uniform", -1.0, 1.0).astype(np.float64)
    assert stats.kstest(u, "uniform", args=(-1.0, 2.0)).pvalue > 1e-3


def test_fill_pinned_and_mutants_rejected(tmp_path):
    oracle_mutants.assert_rejects(tmp_path, check_fill, [
        ("oracle.c", "z += 0x9E3779B97F4A7C15ull;", "z += 0x9E3779B97F4A7C17ull;"),
        ("oracle.c", "sqrt(-2.0 * log(u1))", "sqrt(-1.0 * log(u1))"),
    ])


# ---------------------------------------------------------------------------
# restated CPU baseline against the FP64 routines

def _fp32_gemm_bound(k, alpha, abs_ab, beta, c0):
    return (2.0 * k + 4.0) * U32 * abs(alpha) * abs_ab + 4.0 * U32 * np.abs(beta * c0) + 1e-30


def check_base_sgemm_config1():
    """Config 1 exactly: N,N 512^3, lda = ldb = ldc = 512, alpha 1.5, beta 0.5,
    A/B/C uniform[-1,1] from the shared generator."""
    m = n = k = 512
    a = oracle.fill_f32(m * k, 101, "uniform")
    b = oracle.fill_f32(k * n, 102, "uniform")
    c = oracle.fill_f32(m * n, 103, "uniform")
    want = c.astype(np.float64)
    assert oracle.dgemm("N", "N", m, n, k, 1.5, a.astype(np.float64), m, b.astype(np.float64), k, 0.5, want, m) == 0
    got = c.copy()
    oracle.base_sgemm("N", "N", m, n, k, 1.5, a, m, b, k, 0.5, got, m)
    a64 = a.astype(np.float64).reshape(k, m).T
    b64 = b.astype(np.float64).reshape(n, k).T
    bound = _fp32_gemm_bound(k, 1.5, np.abs(a64) @ np.abs(b64), 0.5, c.astype(np.float64).reshape(n, m).T)
    err = np.abs(got.astype(np.float64) - want).reshape(n, m).T
    assert (err <= bound).all(), float((err / bound).max())


def check_base_sgemm_transposes():
    rng = np.random.default_rng(13)
    for ta, tb in itertools.product("NT", "NT"):
        m, n, k = 45, 70, 38
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        a = np.asfortranarray(rng.uniform(-1, 1, (ar, ac)).astype(np.float32))
        b = np.asfortranarray(rng.uniform(-1, 1, (br, bc)).astype(np.float32))
        c = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
        got = c.ravel(order="F").copy()
        oracle.base_sgemm(ta, tb, m, n, k, -0.5, a.ravel(order="F"), ar, b.ravel(order="F"), br, 2.0, got, m)
        want = -0.5 * _op(ta, a.astype(np.float64)) @ _op(tb, b.astype(np.float64)) + 2.0 * c.astype(np.float64)
        bound = _fp32_gemm_bound(k, -0.5, np.abs(_op(ta, a.astype(np.float64))) @ np.abs(_op(tb, b.astype(np.float64))),
                                 2.0, c.astype(np.float64))
        err = np.abs(got.reshape(n, m).T - want)
        assert (err <= bound).all(), (ta, tb, float((err / bound).max()))


def test_base_sgemm_pinned_and_mutants_rejected(tmp_path):
    oracle_mutants.assert_rejects(tmp_path, check_base_sgemm_config1, [
        ("baseline.c", "dst[l * TILE + r] = alpha * a[mc0 + t + r + (k0 + l) * lda];",
         "dst[l * TILE + r] = a[mc0 + t + r + (k0 + l) * lda];"),
        ("baseline.c", "dst[l * TILE + r] = b[(k0 + l) + (j0 + r) * ldb];", "dst[l * TILE + r] = b[(k0 + l) + j0 * ldb];"),
        ("baseline.c", "for (int64_t i = 0; i < m; ++i) c[i + j * ldc] *= beta;",
         "for (int64_t i = 0; i < m; ++i) c[i + j * ldc] *= 1.0f;"),
    ])
    oracle_mutants.assert_rejects(tmp_path / "t", check_base_sgemm_transposes, [
        ("baseline.c", "for (int64_t i = 0; i < m; ++i) cj[i] += temp * al[i * lda];",
         "for (int64_t i = 0; i < m; ++i) cj[i] -= temp * al[i * lda];"),
        ("baseline.c", "float bv = ntb ? b[l + j * ldb] : b[j + l * ldb];", "float bv = b[l + j * ldb];"),
    ])


def _tri32(rng, n, upper, unit):
    """Well-conditioned fp32 triangle (diagonal in [1,2], off-diagonal
    uniform[-1,1]/n), NaN in the unused triangle and, when unit, on the
    diagonal (neither may be read)."""
    a = rng.uniform(-1, 1, (n, n)) / n
    a[np.diag_indices(n)] = rng.uniform(1, 2, n)
    a = np.triu(a) if upper else np.tril(a)
    a[(np.tril(np.ones((n, n)), -1) if upper else np.triu(np.ones((n, n)), 1)) > 0] = np.nan
    if unit:
        a[np.diag_indices(n)] = np.nan
    return np.asfortranarray(a.astype(np.float32))


def check_base_strsm_left():
    rng = np.random.default_rng(14)
    m, n = 67, 23
    for uplo, trans, diag in itertools.product("UL", "NT", "NU"):
        a = _tri32(rng, m, uplo == "U", diag == "U")
        b = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
        b[5, :] = 0.0  # exercises the zero-pivot skip (trsm.mojo:116,134)
        want = b.astype(np.float64).ravel(order="F")
        assert oracle.dtrsm("L", uplo, trans, diag, m, n, 1.25, a.astype(np.float64).ravel(order="F"), m, want, m) == 0
        got = b.ravel(order="F").copy()
        oracle.base_strsm_left(uplo, trans, diag, m, n, 1.25, a.ravel(order="F"), m, got, m)
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-5, err_msg=str((uplo, trans, diag)))


def test_base_strsm_left_pinned_and_mutants_rejected(tmp_path):
    oracle_mutants.assert_rejects(tmp_path, check_base_strsm_left, [
        ("baseline.c", "for (int64_t i = 0; i < m; ++i) b[i + j * ldb] *= alpha;",
         "for (int64_t i = 0; i < m; ++i) b[i + j * ldb] *= 1.0f;"),
        ("baseline.c", "for (int64_t i = l + 1; i < m; ++i) bj[i] -= p * al[i];",
         "for (int64_t i = l + 1; i < m; ++i) bj[i] += p * al[i];"),
        ("baseline.c", "for (int64_t l = 0; l < i; ++l) bj[i] -= a[l + i * lda] * bj[l];",
         "for (int64_t l = 0; l < i; ++l) bj[i] -= a[i + l * lda] * bj[l];"),
        ("baseline.c", "if (nounit) bj[i] /= a[i + i * lda];", "if (nounit) bj[i] *= a[i + i * lda];",
         "} else {\n            for (int64_t i = m - 1; i >= 0; --i) {"),
    ])


def check_base_ssyrk_n():
    rng = np.random.default_rng(15)
    n, k = 70, 33
    for uplo in "UL":
        a = np.asfortranarray(rng.standard_normal((n, k)).astype(np.float32))
        c = rng.uniform(-1, 1, (n, n))
        tri = np.triu(np.ones((n, n), bool)) if uplo == "U" else np.tril(np.ones((n, n), bool))
        c[~tri] = np.nan
        c = np.asfortranarray(c.astype(np.float32))
        want = c.astype(np.float64).ravel(order="F")
        assert oracle.dsyrk(uplo, "N", n, k, 0.75, a.astype(np.float64).ravel(order="F"), n, -0.5, want, n) == 0
        got = c.ravel(order="F").copy()
        oracle.base_ssyrk_n(uplo, n, k, 0.75, a.ravel(order="F"), n, -0.5, got, n)
        g = got.reshape(n, n).T
        w = want.reshape(n, n).T
        np.testing.assert_allclose(g[tri], w[tri], rtol=0, atol=1e-4 * np.sqrt(k))
        assert np.isnan(g[~tri]).all()


def test_base_ssyrk_n_pinned_and_mutants_rejected(tmp_path):
    oracle_mutants.assert_rejects(tmp_path, check_base_ssyrk_n, [
        ("baseline.c", "c[i + j * ldc] = (beta == 0.0f) ? 0.0f : beta * c[i + j * ldc];",
         "c[i + j * ldc] = (beta == 0.0f) ? 0.0f : c[i + j * ldc];", "void base_ssyrk_n("),
        ("baseline.c", "int64_t i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;",
         "int64_t i0 = upper ? j : 0, i1 = upper ? n : j + 1;", "#pragma omp parallel for schedule(dynamic, 16)"),
        ("baseline.c", "float temp = alpha * av;", "float temp = av;", "void base_ssyrk_n("),
    ])


def _sym32(rng, na, upper):
    s = rng.uniform(-1, 1, (na, na))
    s = s + s.T
    stored = np.triu(s) if upper else np.tril(s)
    stored[(np.tril(np.ones((na, na)), -1) if upper else np.triu(np.ones((na, na)), 1)) > 0] = np.nan
    return np.asfortranarray(stored.astype(np.float32))


def check_base_ssymm():
    rng = np.random.default_rng(16)
    m, n = 70, 66
    for side, uplo in itertools.product("LR", "UL"):
        na = m if side == "L" else n
        a = _sym32(rng, na, uplo == "U")
        b = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
        c = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
        want = c.astype(np.float64).ravel(order="F")
        assert oracle.dsymm(side, uplo, m, n, -1.5, a.astype(np.float64).ravel(order="F"), na,
                            b.astype(np.float64).ravel(order="F"), m, 0.5, want, m) == 0
        got = c.ravel(order="F").copy()
        oracle.base_ssymm(side, uplo, m, n, -1.5, a.ravel(order="F"), na, b.ravel(order="F"), m, 0.5, got, m)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-4 * np.sqrt(na), err_msg=str((side, uplo)))


def check_base_ssymm_zero_skip():
    """Where B(l,j) == 0 the reference skips alpha*temp2 for C(l,j)
    (symm.mojo:114,133): the restatement equals the BLAS result minus
    alpha * sum over the stored off-diagonal part of column l of A(i,l)*B(i,j)."""
    rng = np.random.default_rng(17)
    m, n, l, j = 40, 6, 17, 3
    for uplo in "UL":
        a = _sym32(rng, m, uplo == "U")
        b = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
        b[l, j] = 0.0
        c = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
        want = c.astype(np.float64).ravel(order="F")
        assert oracle.dsymm("L", uplo, m, n, 2.0, a.astype(np.float64).ravel(order="F"), m,
                            b.astype(np.float64).ravel(order="F"), m, 1.0, want, m) == 0
        rows = np.arange(l) if uplo == "U" else np.arange(l + 1, m)
        skipped = 2.0 * np.dot(a[rows, l].astype(np.float64), b[rows, j].astype(np.float64))
        want = want.reshape(n, m).T.copy()
        want[l, j] -= skipped
        got = c.ravel(order="F").copy()
        oracle.base_ssymm("L", uplo, m, n, 2.0, a.ravel(order="F"), m, b.ravel(order="F"), m, 1.0, got, m)
        np.testing.assert_allclose(got.reshape(n, m).T, want, rtol=0, atol=1e-4 * np.sqrt(m))
        assert abs(skipped) > 1e-2  # the skipped term is not negligible


def test_base_ssymm_pinned_and_mutants_rejected(tmp_path):
    oracle_mutants.assert_rejects(tmp_path, check_base_ssymm, [
        ("baseline.c", "cj[l] = cj[l] + t1 * al[l] + alpha * t2;", "cj[l] = cj[l] + t1 * al[l];"),
        ("baseline.c", "stored_kj ? a[k + j * lda] : a[j + k * lda]", "stored_kj ? a[j + k * lda] : a[k + j * lda]"),
        ("baseline.c", "float td = alpha * a[j + j * lda];", "float td = a[j + j * lda];"),
    ])
    oracle_mutants.assert_rejects(tmp_path / "z", check_base_ssymm_zero_skip, [
        ("baseline.c", "if (bj[l] == 0.0f) continue;\n                    float t1", "float t1"),
    ])


def check_base_strmm():
    rng = np.random.default_rng(18)
    m, n = 37, 70
    for side, uplo, trans, diag in itertools.product("LR", "UL", "NT", "NU"):
        na = m if side == "L" else n
        a = _tri32(rng, na, uplo == "U", diag == "U")
        a[np.isfinite(a) & (rng.uniform(size=a.shape) < 0.05)] = 0.0  # the reference's zero skips
        b = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
        b[3, :] = 0.0
        want = b.astype(np.float64).ravel(order="F")
        assert oracle.dtrmm(side, uplo, trans, diag, m, n, 0.8, a.astype(np.float64).ravel(order="F"), na, want,
                            m) == 0
        got = b.ravel(order="F").copy()
        oracle.base_strmm(side, uplo, trans, diag, m, n, 0.8, a.ravel(order="F"), na, got, m)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-5 * na, err_msg=str((side, uplo, trans, diag)))


def test_base_strmm_pinned_and_mutants_rejected(tmp_path):
    oracle_mutants.assert_rejects(tmp_path, check_base_strmm, [
        ("baseline.c", "if (nounit) t *= a[k + k * lda];\n                    bj[k] = t;", "bj[k] = t;"),
        ("baseline.c", "trmm_axpy_col(bj + k + 1, a + k * lda + k + 1, m - k - 1, t);",
         "trmm_axpy_col(bj + k + 1, a + k * lda + k + 1, m - k - 1, -t);"),
        ("baseline.c", "for (int64_t kk = 0; kk < i; ++kk) t += a[kk + i * lda] * bj[kk];",
         "for (int64_t kk = 0; kk < i; ++kk) t += a[i + kk * lda] * bj[kk];"),
        ("baseline.c", "if (a[j + k * lda] != 0.0f) trmm_axpy_col(b + j * ldb, bk, m, alpha * a[j + k * lda]);",
         "if (a[j + k * lda] != 0.0f) trmm_axpy_col(b + j * ldb, bk, m, a[j + k * lda]);"),
        ("baseline.c", "trmm_scale_col(bj, m, nounit ? alpha * a[j + j * lda] : alpha);",
         "trmm_scale_col(bj, m, alpha);", "for (int64_t j = 0; j < n; ++j) {\n                float* bj = b + j * ldb;\n                trmm_scale_col"),
    ])


# ---------------------------------------------------------------------------
# the FP64 routines: their OpenBLAS comparisons reject a planted mistake

def test_fp64_routine_checks_reject_mutants(tmp_path):
    cases = [
        (lambda: test_oracle.test_gemm_vs_openblas("T", "T"), ("oracle.c", "s += av * bv;", "s -= av * bv;")),
        (lambda: test_oracle.test_gemm_vs_openblas("N", "N"),
         ("oracle.c", "double cv = (beta == 0.0) ? 0.0 : beta * C_(i, j);", "double cv = (beta == 0.0) ? 0.0 : C_(i, j);")),
        (lambda: test_oracle.test_trsm_vs_openblas("L", "L", "N", "N"),
         ("oracle.c", "B_(i, j) = s / OPA(i, i);", "B_(i, j) = s * OPA(i, i);")),
        (lambda: test_oracle.test_trsm_vs_openblas("R", "U", "N", "N"),
         ("oracle.c", "for (int64_t i = 0; i < m; ++i) B_(i, j) -= t * B_(i, l);",
          "for (int64_t i = 0; i < m; ++i) B_(i, j) += t * B_(i, l);")),
        (lambda: test_oracle.test_trmm_vs_openblas_and_inverse("L", "U", "T", "U"),
         ("oracle.c", "if (i == j) return unit ? 1.0 : A_(i, j);", "if (i == j) return unit ? 0.0 : A_(i, j);")),
        (lambda: test_oracle.test_syrk_syr2k_vs_openblas("L", "T"),
         ("oracle.c", "s += notr ? A_(i, l) * A_(j, l) : A_(l, i) * A_(l, j);",
          "s += notr ? A_(i, l) * A_(j, l) : A_(l, i) * A_(l, i);")),
        (lambda: test_oracle.test_syrk_syr2k_vs_openblas("U", "N"),
         ("oracle.c", "s += notr ? (A_(i, l) * B_(j, l) + B_(i, l) * A_(j, l))",
          "s += notr ? (A_(i, l) * B_(j, l) + A_(i, l) * B_(j, l))")),
        (lambda: test_oracle.test_symm_vs_openblas("R", "L"),
         ("oracle.c", "? A_((i), (j)) : A_((j), (i)))", "? A_((j), (i)) : A_((i), (j)))")),
        (lambda: test_oracle_level2.test_gemv_vs_openblas("T"),
         ("oracle.c", "y[vidx(j, leny, incy)] += alpha * t;", "y[vidx(j, leny, incy)] += t;")),
        (lambda: test_oracle_level2.test_trsv_vs_openblas_and_residual("U", "N", "N"),
         ("oracle.c", "X(i) = unit ? t : t / OPA(i, i);", "X(i) = unit ? t : t * OPA(i, i);", "} else {\n        for (int64_t i = n - 1;")),
    ]
    for i, (check, *muts) in enumerate(cases):
        oracle_mutants.assert_rejects(tmp_path / ("r%d" % i), check, muts)


# ---------------------------------------------------------------------------
# registry: every public oracle function has a pin

PINNED = {
    "dgemm": "test_oracle.py: 31 KATs, OpenBLAS all flag pairs; mutants here",
    "dtrsm": "test_oracle.py: KATs, OpenBLAS 16 combos, trmm inverse; mutants here",
    "dtrmm": "test_oracle.py: KATs, OpenBLAS 16 combos; mutants here",
    "dsyrk": "test_oracle.py: KATs, OpenBLAS; mutants here",
    "dsymm": "test_oracle.py: KATs, OpenBLAS; mutants here",
    "dsyr2k": "test_oracle.py: KATs, OpenBLAS; mutants here",
    "dgemv": "test_oracle_level2.py: KATs, OpenBLAS; mutants here",
    "dtrsv": "test_oracle_level2.py: KATs, OpenBLAS, residual; mutants here",
    "sgemm_sample": "check_sample_alpha_beta (+ test_oracle.test_sample_matches_full)",
    "sgemm_sample_abs": "check_sample_abs",
    "fill_f32": "check_fill",
    "base_threads": "reports omp_get_max_threads(); checked against os.sched_getaffinity below",
    "base_sgemm": "check_base_sgemm_config1, check_base_sgemm_transposes",
    "base_strsm_left": "check_base_strsm_left",
    "base_ssyrk_n": "check_base_ssyrk_n",
    "base_ssymm": "check_base_ssymm, check_base_ssymm_zero_skip",
    "base_strmm": "check_base_strmm",
}
NOT_ROUTINES = {"build", "lib", "swapped"}


def test_every_oracle_function_is_pinned():
    public = {name for name, obj in vars(oracle).items()
              if inspect.isfunction(obj) and not name.startswith("_") and obj.__module__ == "oracle"}
    assert public - NOT_ROUTINES == set(PINNED), sorted((public - NOT_ROUTINES) ^ set(PINNED))


def test_base_threads_matches_affinity():
    import os
    assert 1 <= oracle.base_threads() <= max(len(os.sched_getaffinity(0)), os.cpu_count() or 1)
