/*
 * FP64 CPU oracle for the level-3 hot path (gemm x4 transposes, trsm, trmm,
 * syrk, symm, syr2k) and the level-2 gemv / trsv built next to it.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.  The product path
 * (mojoblas_amd/libmojoblas_hip.so) does not link or call it.
 *
 * Semantics are BLAS (Netlib) semantics as the reference states them in its
 * docstrings and validation code:
 *   gemm  src/level3/gemm.mojo:127-377   (_gemm_naive: info codes, quick returns)
 *   trsm  src/level3/trsm.mojo:20-275
 *   trmm  src/level3/trmm.mojo:20-305
 *   syrk  src/level3/syrk.mojo:20-200
 *   symm  src/level3/symm.mojo:20-241
 *   syr2k src/level3/syr2k.mojo:20-190
 *   gemv  src/level2/gemv.mojo:23-212   (trans is "N"/"T"/"C" only, :77)
 *   trsv  src/level2/trsv.mojo:20-214
 * Flags are compared as whole strings exactly as the reference does
 * ("N"/"n"/"T"/"t"/"C"/"c", gemm.mojo:180-197), so "No" is info 1.
 * Deliberate divergences from the reference's arithmetic (documented where
 * they occur): no zero-skips (gemm.mojo:263 etc.), trsm side=R trans=T solves
 * the transposed system (the reference solves the untransposed one,
 * trsm.mojo:223-273), symm side=L does not skip alpha*temp2 when B(l,j)==0
 * (symm.mojo:114,133), gemm always applies alpha (gemm_v7 drops it,
 * gemm.mojo:923-936).
 *
 * Return value: 0 on success, the Netlib argument index (info) on invalid
 * arguments; outputs are untouched when info != 0.  Dimensions are int64.
 */
#ifndef MOJOBLAS_ORACLE_H
#define MOJOBLAS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

int orc_dgemm(const char* transa, const char* transb, int64_t m, int64_t n, int64_t k,
              double alpha, const double* a, int64_t lda, const double* b, int64_t ldb,
              double beta, double* c, int64_t ldc);

int orc_dtrsm(const char* side, const char* uplo, const char* transa, const char* diag,
              int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
              double* b, int64_t ldb);

int orc_dtrmm(const char* side, const char* uplo, const char* transa, const char* diag,
              int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
              double* b, int64_t ldb);

int orc_dsyrk(const char* uplo, const char* trans, int64_t n, int64_t k, double alpha,
              const double* a, int64_t lda, double beta, double* c, int64_t ldc);

int orc_dsymm(const char* side, const char* uplo, int64_t m, int64_t n, double alpha,
              const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
              double* c, int64_t ldc);

int orc_dsyr2k(const char* uplo, const char* trans, int64_t n, int64_t k, double alpha,
               const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
               double* c, int64_t ldc);

/* y := alpha*op(A)*x + beta*y, A m x n, strided x / y (incx, incy != 0;
 * negative increments walk the vector from its far end, as in BLAS).  Info
 * numbering of gemv.mojo:76-91.  Divergence: no zero-skip on x(j)
 * (gemv.mojo:136-141). */
int orc_dgemv(const char* trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
              const double* x, int64_t incx, double beta, double* y, int64_t incy);

/* x := op(A)^-1 x, A n x n triangular.  Info numbering of trsv.mojo:57-80.
 * Divergences: no zero-skip on x(j) (trsv.mojo:95,120), and for upper, 'N',
 * incx != 1 element j of x is x((j)*incx) as in BLAS (trsv.mojo:108-111
 * decrements the index before its first use, reading one stride off). */
int orc_dtrsv(const char* uplo, const char* trans, const char* diag, int64_t n, const double* a,
              int64_t lda, double* x, int64_t incx);

/* Sampled FP64 gemm: out[s] = alpha*sum_l op(A)(rows[s],l)*op(B)(l,cols[s]) + beta*c0[s]
 * (c0 may be NULL when beta == 0).  Inputs are fp32 (the hot path's data type);
 * every product is exact in fp64.  Used to price full-size results on a sample
 * of entries without an O(mnk) CPU pass. */
void orc_sgemm_sample_f64(char transa, char transb, int64_t k, double alpha,
                          const float* a, int64_t lda, const float* b, int64_t ldb,
                          double beta, const float* c0, int64_t nsamp,
                          const int64_t* rows, const int64_t* cols, double* out);

/* Same sample, also returns sum_l |op(A)(i,l)|*|op(B)(l,j)| per entry (the
 * |A||B| term of the componentwise error bound). */
void orc_sgemm_sample_abs(char transa, char transb, int64_t k,
                          const float* a, int64_t lda, const float* b, int64_t ldb,
                          int64_t nsamp, const int64_t* rows, const int64_t* cols,
                          double* out_abs);

/* Reproducible input generator shared by tests and bench (splitmix64 counter
 * PRNG): dist 0 = uniform[lo,hi), 1 = normal(0,1) by Box-Muller in fp64,
 * rounded to fp32.  Element i depends only on (seed, i). */
void orc_fill_f32(float* x, int64_t count, uint64_t seed, int dist, double lo, double hi);

#ifdef __cplusplus
}
#endif
#endif
