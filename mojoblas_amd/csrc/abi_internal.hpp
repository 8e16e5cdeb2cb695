// Helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <mutex>

#include "dispatch.hpp"

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
int host_precision();

// Host-pointer entries: the calling thread's own staging stream on the
// current device, with its workspace leased for the call.
class HostStage {
   public:
    HostStage();
    hipError_t error() const { return err_; }
    hipStream_t stream() const { return s_; }

   private:
    hipStream_t s_ = nullptr;
    hipError_t err_ = hipSuccess;
    std::unique_ptr<StreamLease> lease_;
};
// Grow-only staging buffer i of the calling thread (current device).
hipError_t slot_get(int i, size_t bytes, void** out);
hipError_t put_matrix(void* dev, const void* host, int64_t r, int64_t c, int64_t ld, size_t es, hipStream_t s);
// Copies back and synchronises s.
hipError_t get_matrix(void* host, const void* dev, int64_t r, int64_t c, int64_t ld, size_t es, hipStream_t s);

}  // namespace mbh
