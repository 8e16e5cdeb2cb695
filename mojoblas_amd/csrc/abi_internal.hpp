// Helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

namespace mbh {

void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* where);
bool flag_n(char c);
bool flag_t(char c);
bool flag_l(char c);
bool flag_r(char c);
bool flag_u(char c);
int check_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc);
int check_trxm(char side, char uplo, char ta, char diag, int64_t m, int64_t n, int64_t lda, int64_t ldb);
int check_syrk(char uplo, char trans, int64_t n, int64_t k, int64_t lda, int64_t ldc);
int check_syr2k(char uplo, char trans, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc);
int check_symm(char side, char uplo, int64_t m, int64_t n, int64_t lda, int64_t ldb, int64_t ldc);
int default_precision();
bool gemm_precision_ok(int prec);
hipError_t slot_get(int i, size_t bytes, void** out);
std::unique_lock<std::mutex> stage_lock();
hipError_t put_matrix(void* dev, const void* host, int64_t r, int64_t c, int64_t ld, size_t es);
hipError_t get_matrix(void* host, const void* dev, int64_t r, int64_t c, int64_t ld, size_t es);

}  // namespace mbh
