// Level-3 drivers (device pointers, validated arguments, asynchronous on s).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbh {

// B := alpha*inv(op(A))*B or alpha*B*inv(op(A)).  `prec` selects the gemm
// of the off-diagonal updates (fp32 data), `refine` the number of fp64-residual
// refinement steps (0 or 1; fp32 data only).
template <typename T>
hipError_t trsm(char side, char uplo, char transa, char diag, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                T* B, int64_t ldb, int prec, int refine, hipStream_t s);
// fp64 inverses of the 128x128 diagonal blocks of op(A) (A na x na, op(A)
// lower iff eff_lower), into inv + b*128*128 (column-major, ld 128, zero
// padded past na).  Used by trsm and trsv.
constexpr int TRI_NB = 128;
template <typename T>
hipError_t tri_block_inverses(int64_t na, const T* A, int64_t lda, bool eff_lower, bool trans, bool unit, double* inv,
                              hipStream_t s);
template <typename T>
hipError_t trmm(char side, char uplo, char transa, char diag, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                T* B, int64_t ldb, int prec, hipStream_t s);
template <typename T>
hipError_t syrk(char uplo, char trans, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, T beta, T* C,
                int64_t ldc, int prec, hipStream_t s);
template <typename T>
hipError_t syr2k(char uplo, char trans, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B,
                 int64_t ldb, T beta, T* C, int64_t ldc, int prec, hipStream_t s);
template <typename T>
hipError_t symm(char side, char uplo, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb,
                T beta, T* C, int64_t ldc, int prec, hipStream_t s);

}  // namespace mbh
