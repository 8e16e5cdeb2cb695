/*
 * FP64 CPU oracle for the level-3 hot path.  TEST INFRASTRUCTURE ONLY: see
 * oracle.h.  Every routine restates the BLAS semantics of the reference file
 * cited above it; accumulation is in fp64 so fp32-input products are exact
 * and the oracle's own rounding error (about k * 2^-53) is far below the fp32
 * errors it prices.
 *
 * Pinned by tests/test_oracle.py against: the 31 known answers of
 * tests/test_level3.mojo (extracted to tests/golden/level3_kat.json),
 * scipy.linalg.blas (OpenBLAS, the library tests/reference.c:88-161 binds)
 * on random inputs for every flag combination, closed forms (identity,
 * inverse round trips trsm(trmm(B)) == B) and numpy float64 matmul.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Whole-string flag tests, as gemm.mojo:180-197 / trsm.mojo:63-76 compare. */
static int eq(const char* s, const char* t) { return s && strcmp(s, t) == 0; }
static int flag_n(const char* s) { return eq(s, "N") || eq(s, "n"); }
static int flag_t(const char* s) { return eq(s, "T") || eq(s, "t") || eq(s, "C") || eq(s, "c"); }
static int flag_l(const char* s) { return eq(s, "L") || eq(s, "l"); }
static int flag_r(const char* s) { return eq(s, "R") || eq(s, "r"); }
static int flag_u(const char* s) { return eq(s, "U") || eq(s, "u"); }
static int64_t max1(int64_t v) { return v > 1 ? v : 1; }

#define A_(i, j) a[(i) + (j) * lda]
#define B_(i, j) b[(i) + (j) * ldb]
#define C_(i, j) c[(i) + (j) * ldc]

/* C := alpha*op(A)*op(B) + beta*C.  Validation and quick returns follow
 * gemm.mojo:179-237; beta == 0 writes without reading C (gemm.mojo:218-226).
 * The product is formed as an fp64 dot product per entry (no zero-skip). */
int orc_dgemm(const char* transa, const char* transb, int64_t m, int64_t n, int64_t k,
              double alpha, const double* a, int64_t lda, const double* b, int64_t ldb,
              double beta, double* c, int64_t ldc) {
    int na = flag_n(transa), nb = flag_n(transb);
    int info = 0;
    if (!na && !flag_t(transa)) info = 1;
    else if (!nb && !flag_t(transb)) info = 2;
    else if (m < 0) info = 3;
    else if (n < 0) info = 4;
    else if (k < 0) info = 5;
    else if (lda < max1(na ? m : k)) info = 8;
    else if (ldb < max1(nb ? k : n)) info = 10;
    else if (ldc < max1(m)) info = 13;
    if (info) return info;
    if (m == 0 || n == 0) return 0;
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < m; ++i) {
            double s = 0.0;
            if (alpha != 0.0) {
                for (int64_t l = 0; l < k; ++l) {
                    double av = na ? A_(i, l) : A_(l, i);
                    double bv = nb ? B_(l, j) : B_(j, l);
                    s += av * bv;
                }
            }
            double cv = (beta == 0.0) ? 0.0 : beta * C_(i, j);
            C_(i, j) = (alpha == 0.0 || k == 0) ? cv : alpha * s + cv;
        }
    }
    return 0;
}

/* Element of the triangular operand as stored: zero outside the `upper`
 * triangle, 1 on the diagonal when unit. */
static double tri_el(const double* a, int64_t lda, int upper, int unit, int64_t i, int64_t j) {
    if (i == j) return unit ? 1.0 : A_(i, j);
    if (upper ? (i < j) : (i > j)) return A_(i, j);
    return 0.0;
}

/* B := alpha*inv(op(A))*B or alpha*B*inv(op(A)).  Validation trsm.mojo:62-89,
 * alpha pre-scale trsm.mojo:96-103, left substitutions trsm.mojo:110-168,
 * right no-trans trsm.mojo:170-222.  Right transposed solves X*op(A)=B with
 * op(A)=A^T (Netlib dtrsm), where the reference reuses its no-trans loops
 * (trsm.mojo:223-273). */
int orc_dtrsm(const char* side, const char* uplo, const char* transa, const char* diag,
              int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
              double* b, int64_t ldb) {
    int left = flag_l(side), upper = flag_u(uplo), notr = flag_n(transa);
    int info = 0;
    if (!left && !flag_r(side)) info = 1;
    else if (!upper && !flag_l(uplo)) info = 2;
    else if (!notr && !flag_t(transa)) info = 3;
    else if (!flag_u(diag) && !flag_n(diag)) info = 4;
    else if (m < 0) info = 5;
    else if (n < 0) info = 6;
    else if (lda < max1(left ? m : n)) info = 9;
    else if (ldb < max1(m)) info = 11;
    if (info) return info;
    if (m == 0 || n == 0) return 0;
    int unit = flag_u(diag);
    if (alpha != 1.0)
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) B_(i, j) = alpha * B_(i, j);
    /* op(A) as an explicit triangle: lower-triangular iff (lower) xor (trans). */
    int eff_lower = (!upper) ^ (!notr);
#define OPA(i, j) (notr ? tri_el(a, lda, upper, unit, (i), (j)) : tri_el(a, lda, upper, unit, (j), (i)))
    if (left) {
        /* op(A) X = B, column by column. */
        for (int64_t j = 0; j < n; ++j) {
            if (eff_lower) {
                for (int64_t i = 0; i < m; ++i) {
                    double s = B_(i, j);
                    for (int64_t l = 0; l < i; ++l) s -= OPA(i, l) * B_(l, j);
                    B_(i, j) = s / OPA(i, i);
                }
            } else {
                for (int64_t i = m - 1; i >= 0; --i) {
                    double s = B_(i, j);
                    for (int64_t l = i + 1; l < m; ++l) s -= OPA(i, l) * B_(l, j);
                    B_(i, j) = s / OPA(i, i);
                }
            }
        }
    } else {
        /* X op(A) = B: column j of X couples to columns l with op(A)(l,j) != 0. */
        if (!eff_lower) { /* op(A) upper: x_j = (b_j - sum_{l<j} x_l opA(l,j)) / opA(j,j) */
            for (int64_t j = 0; j < n; ++j) {
                for (int64_t l = 0; l < j; ++l) {
                    double t = OPA(l, j);
                    for (int64_t i = 0; i < m; ++i) B_(i, j) -= t * B_(i, l);
                }
                double d = OPA(j, j);
                for (int64_t i = 0; i < m; ++i) B_(i, j) /= d;
            }
        } else { /* op(A) lower: backward over j */
            for (int64_t j = n - 1; j >= 0; --j) {
                for (int64_t l = j + 1; l < n; ++l) {
                    double t = OPA(l, j);
                    for (int64_t i = 0; i < m; ++i) B_(i, j) -= t * B_(i, l);
                }
                double d = OPA(j, j);
                for (int64_t i = 0; i < m; ++i) B_(i, j) /= d;
            }
        }
    }
#undef OPA
    return 0;
}

/* B := alpha*op(A)*B or alpha*B*op(A), A triangular.  Validation
 * trmm.mojo:62-89, alpha == 0 zeroes B (trmm.mojo:94-98); the product is the
 * definition with an explicit triangle, accumulated in fp64 into a copy. */
int orc_dtrmm(const char* side, const char* uplo, const char* transa, const char* diag,
              int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
              double* b, int64_t ldb) {
    int left = flag_l(side), upper = flag_u(uplo), notr = flag_n(transa);
    int info = 0;
    if (!left && !flag_r(side)) info = 1;
    else if (!upper && !flag_l(uplo)) info = 2;
    else if (!notr && !flag_t(transa)) info = 3;
    else if (!flag_u(diag) && !flag_n(diag)) info = 4;
    else if (m < 0) info = 5;
    else if (n < 0) info = 6;
    else if (lda < max1(left ? m : n)) info = 9;
    else if (ldb < max1(m)) info = 11;
    if (info) return info;
    if (m == 0 || n == 0) return 0;
    if (alpha == 0.0) {
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) B_(i, j) = 0.0;
        return 0;
    }
    int unit = flag_u(diag);
    double* t = (double*)malloc(sizeof(double) * (size_t)(m * n));
#define OPA(i, j) (notr ? tri_el(a, lda, upper, unit, (i), (j)) : tri_el(a, lda, upper, unit, (j), (i)))
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < m; ++i) {
            double s = 0.0;
            if (left)
                for (int64_t l = 0; l < m; ++l) s += OPA(i, l) * B_(l, j);
            else
                for (int64_t l = 0; l < n; ++l) s += B_(i, l) * OPA(l, j);
            t[i + j * m] = alpha * s;
        }
    }
#undef OPA
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) B_(i, j) = t[i + j * m];
    free(t);
    return 0;
}

/* C := alpha*A*A^T + beta*C (trans N) or alpha*A^T*A + beta*C, only the
 * `uplo` triangle of C referenced.  Validation syrk.mojo:60-83, quick return
 * syrk.mojo:85, beta handling syrk.mojo:89-110. */
int orc_dsyrk(const char* uplo, const char* trans, int64_t n, int64_t k, double alpha,
              const double* a, int64_t lda, double beta, double* c, int64_t ldc) {
    int upper = flag_u(uplo), notr = flag_n(trans);
    int info = 0;
    if (!upper && !flag_l(uplo)) info = 1;
    else if (!notr && !flag_t(trans)) info = 2;
    else if (n < 0) info = 3;
    else if (k < 0) info = 4;
    else if (lda < max1(notr ? n : k)) info = 7;
    else if (ldc < max1(n)) info = 10;
    if (info) return info;
    if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return 0;
    for (int64_t j = 0; j < n; ++j) {
        int64_t i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
        for (int64_t i = i0; i < i1; ++i) {
            double s = 0.0;
            if (alpha != 0.0)
                for (int64_t l = 0; l < k; ++l)
                    s += notr ? A_(i, l) * A_(j, l) : A_(l, i) * A_(l, j);
            double cv = (beta == 0.0) ? 0.0 : beta * C_(i, j);
            C_(i, j) = (alpha == 0.0) ? cv : alpha * s + cv;
        }
    }
    return 0;
}

/* C := alpha*A*B + beta*C (side L) or alpha*B*A + beta*C (side R), A
 * symmetric, only its `uplo` triangle referenced.  Validation symm.mojo:68-86,
 * quick return symm.mojo:88; beta == 0 writes without reading C
 * (symm.mojo:101-102). */
int orc_dsymm(const char* side, const char* uplo, int64_t m, int64_t n, double alpha,
              const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
              double* c, int64_t ldc) {
    int left = flag_l(side), upper = flag_u(uplo);
    int info = 0;
    if (!left && !flag_r(side)) info = 1;
    else if (!upper && !flag_l(uplo)) info = 2;
    else if (m < 0) info = 3;
    else if (n < 0) info = 4;
    else if (lda < max1(left ? m : n)) info = 7;
    else if (ldb < max1(m)) info = 9;
    else if (ldc < max1(m)) info = 12;
    if (info) return info;
    if (m == 0 || n == 0 || (alpha == 0.0 && beta == 1.0)) return 0;
#define SYA(i, j) ((upper ? ((i) <= (j)) : ((i) >= (j))) ? A_((i), (j)) : A_((j), (i)))
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < m; ++i) {
            double s = 0.0;
            if (alpha != 0.0) {
                if (left)
                    for (int64_t l = 0; l < m; ++l) s += SYA(i, l) * B_(l, j);
                else
                    for (int64_t l = 0; l < n; ++l) s += B_(i, l) * SYA(l, j);
            }
            double cv = (beta == 0.0) ? 0.0 : beta * C_(i, j);
            C_(i, j) = (alpha == 0.0) ? cv : alpha * s + cv;
        }
    }
#undef SYA
    return 0;
}

/* C := alpha*(A*B^T + B*A^T) + beta*C (trans N) or alpha*(A^T*B + B^T*A) +
 * beta*C on the `uplo` triangle (src/level3/syr2k.mojo:20-190). */
int orc_dsyr2k(const char* uplo, const char* trans, int64_t n, int64_t k, double alpha,
               const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
               double* c, int64_t ldc) {
    int upper = flag_u(uplo), notr = flag_n(trans);
    int info = 0;
    if (!upper && !flag_l(uplo)) info = 1;
    else if (!notr && !flag_t(trans)) info = 2;
    else if (n < 0) info = 3;
    else if (k < 0) info = 4;
    else if (lda < max1(notr ? n : k)) info = 7;
    else if (ldb < max1(notr ? n : k)) info = 9;
    else if (ldc < max1(n)) info = 12;
    if (info) return info;
    if (n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return 0;
    for (int64_t j = 0; j < n; ++j) {
        int64_t i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
        for (int64_t i = i0; i < i1; ++i) {
            double s = 0.0;
            if (alpha != 0.0)
                for (int64_t l = 0; l < k; ++l)
                    s += notr ? (A_(i, l) * B_(j, l) + B_(i, l) * A_(j, l))
                              : (A_(l, i) * B_(l, j) + B_(l, i) * A_(l, j));
            double cv = (beta == 0.0) ? 0.0 : beta * C_(i, j);
            C_(i, j) = (alpha == 0.0) ? cv : alpha * s + cv;
        }
    }
    return 0;
}

#undef A_
#undef B_
#undef C_

/* ---- level 2 ------------------------------------------------------------ */

/* Logical element i of a strided vector of length len (BLAS convention). */
static inline int64_t vidx(int64_t i, int64_t len, int64_t inc) { return inc > 0 ? i * inc : (len - 1 - i) * (-inc); }

int orc_dgemv(const char* trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
              const double* x, int64_t incx, double beta, double* y, int64_t incy) {
    /* gemv.mojo:76-91: trans compared against "N", "T", "C" only */
    const int tn = eq(trans, "N"), tt = eq(trans, "T") || eq(trans, "C");
    int info = 0;
    if (!tn && !tt) info = 1;
    else if (m < 0) info = 2;
    else if (n < 0) info = 3;
    else if (lda < max1(m)) info = 6;
    else if (incx == 0) info = 8;
    else if (incy == 0) info = 11;
    if (info) return info;
    if (m == 0 || n == 0 || (alpha == 0.0 && beta == 1.0)) return 0;
    const int64_t lenx = tn ? n : m, leny = tn ? m : n;
    for (int64_t i = 0; i < leny; ++i) {
        double* yi = y + vidx(i, leny, incy);
        *yi = (beta == 0.0) ? 0.0 : beta * *yi;
    }
    if (alpha == 0.0) return 0;
    if (tn) {
        for (int64_t j = 0; j < n; ++j) {
            const double t = alpha * x[vidx(j, lenx, incx)];
            for (int64_t i = 0; i < m; ++i) y[vidx(i, leny, incy)] += t * a[i + j * lda];
        }
    } else {
        for (int64_t j = 0; j < n; ++j) {
            double t = 0.0;
            for (int64_t i = 0; i < m; ++i) t += a[i + j * lda] * x[vidx(i, lenx, incx)];
            y[vidx(j, leny, incy)] += alpha * t;
        }
    }
    return 0;
}

int orc_dtrsv(const char* uplo, const char* trans, const char* diag, int64_t n, const double* a,
              int64_t lda, double* x, int64_t incx) {
    int info = 0;
    if (!flag_u(uplo) && !flag_l(uplo)) info = 1;
    else if (!flag_n(trans) && !flag_t(trans)) info = 2;
    else if (!flag_u(diag) && !flag_n(diag)) info = 3;
    else if (n < 0) info = 4;
    else if (lda < max1(n)) info = 6;
    else if (incx == 0) info = 8;
    if (info) return info;
    if (n == 0) return 0;
    const int upper = flag_u(uplo), unit = flag_u(diag), tr = flag_t(trans);
    /* op(A) is lower iff (lower and 'N') or (upper and 'T'); solve op(A) x = b by
     * forward (lower) or backward (upper) substitution on op(A)(i,j). */
    const int eff_lower = (!upper) != tr;
#define OPA(i, j) (tr ? a[(j) + (i) * lda] : a[(i) + (j) * lda])
#define X(i) x[vidx((i), n, incx)]
    if (eff_lower) {
        for (int64_t i = 0; i < n; ++i) {
            double t = X(i);
            for (int64_t j = 0; j < i; ++j) t -= OPA(i, j) * X(j);
            X(i) = unit ? t : t / OPA(i, i);
        }
    } else {
        for (int64_t i = n - 1; i >= 0; --i) {
            double t = X(i);
            for (int64_t j = i + 1; j < n; ++j) t -= OPA(i, j) * X(j);
            X(i) = unit ? t : t / OPA(i, i);
        }
    }
#undef OPA
#undef X
    return 0;
}

static inline double opa_f(char ta, const float* a, int64_t lda, int64_t i, int64_t l) {
    return (ta == 'N' || ta == 'n') ? (double)a[i + l * lda] : (double)a[l + i * lda];
}
static inline double opb_f(char tb, const float* b, int64_t ldb, int64_t l, int64_t j) {
    return (tb == 'N' || tb == 'n') ? (double)b[l + j * ldb] : (double)b[j + l * ldb];
}

void orc_sgemm_sample_f64(char transa, char transb, int64_t k, double alpha,
                          const float* a, int64_t lda, const float* b, int64_t ldb,
                          double beta, const float* c0, int64_t nsamp,
                          const int64_t* rows, const int64_t* cols, double* out) {
#pragma omp parallel for schedule(dynamic, 64) if (nsamp * k > 1000000)
    for (int64_t s = 0; s < nsamp; ++s) {
        int64_t i = rows[s], j = cols[s];
        double acc = 0.0;
        for (int64_t l = 0; l < k; ++l) acc += opa_f(transa, a, lda, i, l) * opb_f(transb, b, ldb, l, j);
        double cv = (beta == 0.0 || !c0) ? 0.0 : beta * (double)c0[s];
        out[s] = alpha * acc + cv;
    }
}

void orc_sgemm_sample_abs(char transa, char transb, int64_t k,
                          const float* a, int64_t lda, const float* b, int64_t ldb,
                          int64_t nsamp, const int64_t* rows, const int64_t* cols,
                          double* out_abs) {
#pragma omp parallel for schedule(dynamic, 64) if (nsamp * k > 1000000)
    for (int64_t s = 0; s < nsamp; ++s) {
        int64_t i = rows[s], j = cols[s];
        double acc = 0.0;
        for (int64_t l = 0; l < k; ++l)
            acc += fabs(opa_f(transa, a, lda, i, l)) * fabs(opb_f(transb, b, ldb, l, j));
        out_abs[s] = acc;
    }
}

/* splitmix64 (Steele, Lea, Flood 2014): a counter-based 64-bit mixer. */
static inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t seed, uint64_t ctr) {
    /* 53 random bits -> (0,1) open interval (avoid log(0) in Box-Muller). */
    uint64_t r = splitmix64(seed * 0x632BE59BD9B4E019ull + ctr);
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void orc_fill_f32(float* x, int64_t count, uint64_t seed, int dist, double lo, double hi) {
    for (int64_t i = 0; i < count; ++i) {
        if (dist == 0) {
            x[i] = (float)(lo + (hi - lo) * u01(seed, (uint64_t)i));
        } else {
            double u1 = u01(seed, 2 * (uint64_t)i), u2 = u01(seed, 2 * (uint64_t)i + 1);
            x[i] = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
        }
    }
}
