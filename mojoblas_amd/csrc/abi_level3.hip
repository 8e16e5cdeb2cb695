// C-ABI entries for trsm, trmm, syrk, syr2k and symm: validation with the
// reference's info codes, quick returns, host-pointer staging, dispatch to
// the level-3 drivers.
#include "../../include/mojoblas_hip.h"
#include "abi_internal.hpp"
#include "dispatch.hpp"
#include "level3.hpp"

namespace mbh {
namespace {

int64_t max1(int64_t v) { return v > 1 ? v : 1; }

int finish(hipError_t e, const char* where) { return e == hipSuccess ? 0 : hip_fail(e, where); }

// Host staging: operands are copied to packed device buffers (ld = rows),
// the device routine runs on the null stream, outputs are copied back.
struct Staged {
    void* p = nullptr;
    int64_t r = 0, c = 0;
};

template <typename T>
hipError_t stage_in(int slot, const T* host, int64_t r, int64_t c, int64_t ld, Staged& out, hipStream_t s) {
    out.r = r;
    out.c = c;
    hipError_t e = slot_get(slot, (size_t)max1(r) * max1(c) * sizeof(T), &out.p);
    if (e != hipSuccess) return e;
    return put_matrix(out.p, host, r, c, ld, sizeof(T), s);
}

template <typename T>
hipError_t stage_out(T* host, const Staged& st, int64_t ld, hipStream_t s) {
    return get_matrix(host, st.p, st.r, st.c, ld, sizeof(T), s);
}

#define MBH_LEASE(name)                                                               \
    StreamLease lease((hipStream_t)stream);                                           \
    if (lease.error() != hipSuccess) return hip_fail(lease.error(), name)
#define MBH_HOST_STAGE(name)                                                          \
    HostStage st;                                                                     \
    if (st.error() != hipSuccess) return hip_fail(st.error(), name ": host staging"); \
    const hipStream_t hs = st.stream()

template <typename T>
int trsm_dev(char side, char uplo, char ta, char diag, int64_t m, int64_t n, T alpha, const T* a, int64_t lda, T* b,
             int64_t ldb, int prec, int refine, void* stream) {
    int info = check_trxm(side, uplo, ta, diag, m, n, lda, ldb);
    if (info) return info;
    if (sizeof(T) == 4 && !gemm_precision_ok(prec)) {
        set_error("trsm: invalid precision %d", prec);
        return MBH_E_PRECISION;
    }
    if (refine < 0 || refine > 1) {
        set_error("trsm: refine must be 0 or 1 (got %d)", refine);
        return MBH_E_ARG;
    }
    if (m == 0 || n == 0) return 0;
    MBH_LEASE("trsm");
    return finish(trsm<T>(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, prec, refine, (hipStream_t)stream),
                  "trsm");
}

template <typename T>
int trsm_host(char side, char uplo, char ta, char diag, int64_t m, int64_t n, T alpha, const T* a, int64_t lda, T* b,
              int64_t ldb) {
    int info = check_trxm(side, uplo, ta, diag, m, n, lda, ldb);
    if (info) return info;
    if (m == 0 || n == 0) return 0;
    const int64_t na = flag_l(side) ? m : n;
    MBH_HOST_STAGE("trsm");
    Staged A, B;
    hipError_t e;
    if ((e = stage_in(0, a, na, na, lda, A, hs)) != hipSuccess) return hip_fail(e, "trsm: stage A");
    if ((e = stage_in(1, b, m, n, ldb, B, hs)) != hipSuccess) return hip_fail(e, "trsm: stage B");
    if ((e = trsm<T>(side, uplo, ta, diag, m, n, alpha, (const T*)A.p, max1(na), (T*)B.p, m, host_precision(), 1,
                     hs)) != hipSuccess)
        return hip_fail(e, "trsm");
    return finish(stage_out(b, B, ldb, hs), "trsm: copy B back");
}

template <typename T>
int trmm_dev(char side, char uplo, char ta, char diag, int64_t m, int64_t n, T alpha, const T* a, int64_t lda, T* b,
             int64_t ldb, int prec, void* stream) {
    int info = check_trxm(side, uplo, ta, diag, m, n, lda, ldb);
    if (info) return info;
    if (sizeof(T) == 4 && !gemm_precision_ok(prec)) {
        set_error("trmm: invalid precision %d", prec);
        return MBH_E_PRECISION;
    }
    if (m == 0 || n == 0) return 0;
    MBH_LEASE("trmm");
    return finish(trmm<T>(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, prec, (hipStream_t)stream), "trmm");
}

template <typename T>
int trmm_host(char side, char uplo, char ta, char diag, int64_t m, int64_t n, T alpha, const T* a, int64_t lda, T* b,
              int64_t ldb) {
    int info = check_trxm(side, uplo, ta, diag, m, n, lda, ldb);
    if (info) return info;
    if (m == 0 || n == 0) return 0;
    const int64_t na = flag_l(side) ? m : n;
    MBH_HOST_STAGE("trmm");
    Staged A, B;
    hipError_t e;
    if (alpha != T(0))
        if ((e = stage_in(0, a, na, na, lda, A, hs)) != hipSuccess) return hip_fail(e, "trmm: stage A");
    if ((e = stage_in(1, b, m, n, ldb, B, hs)) != hipSuccess) return hip_fail(e, "trmm: stage B");
    if ((e = trmm<T>(side, uplo, ta, diag, m, n, alpha, (const T*)A.p, max1(na), (T*)B.p, m, host_precision(),
                     hs)) != hipSuccess)
        return hip_fail(e, "trmm");
    return finish(stage_out(b, B, ldb, hs), "trmm: copy B back");
}

template <typename T>
int syrk_dev(char uplo, char trans, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, T beta, T* c, int64_t ldc,
             int prec, void* stream) {
    int info = check_syrk(uplo, trans, n, k, lda, ldc);
    if (info) return info;
    if (sizeof(T) == 4 && !gemm_precision_ok(prec)) {
        set_error("syrk: invalid precision %d", prec);
        return MBH_E_PRECISION;
    }
    if (n == 0) return 0;
    MBH_LEASE("syrk");
    return finish(syrk<T>(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, prec, (hipStream_t)stream), "syrk");
}

template <typename T>
int syrk_host(char uplo, char trans, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, T beta, T* c,
              int64_t ldc) {
    int info = check_syrk(uplo, trans, n, k, lda, ldc);
    if (info) return info;
    if (n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return 0;
    const bool tr = !flag_n(trans);
    const int64_t ar = tr ? k : n, ac = tr ? n : k;
    MBH_HOST_STAGE("syrk");
    Staged A, C;
    hipError_t e;
    if ((e = stage_in(0, a, ar, ac, lda, A, hs)) != hipSuccess) return hip_fail(e, "syrk: stage A");
    if ((e = stage_in(2, c, n, n, ldc, C, hs)) != hipSuccess) return hip_fail(e, "syrk: stage C");
    if ((e = syrk<T>(uplo, trans, n, k, alpha, (const T*)A.p, max1(ar), beta, (T*)C.p, n, host_precision(),
                     hs)) != hipSuccess)
        return hip_fail(e, "syrk");
    return finish(stage_out(c, C, ldc, hs), "syrk: copy C back");
}

template <typename T>
int syr2k_dev(char uplo, char trans, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb,
              T beta, T* c, int64_t ldc, int prec, void* stream) {
    int info = check_syr2k(uplo, trans, n, k, lda, ldb, ldc);
    if (info) return info;
    if (sizeof(T) == 4 && !gemm_precision_ok(prec)) {
        set_error("syr2k: invalid precision %d", prec);
        return MBH_E_PRECISION;
    }
    if (n == 0) return 0;
    MBH_LEASE("syr2k");
    return finish(syr2k<T>(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc, prec, (hipStream_t)stream),
                  "syr2k");
}

template <typename T>
int syr2k_host(char uplo, char trans, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb,
               T beta, T* c, int64_t ldc) {
    int info = check_syr2k(uplo, trans, n, k, lda, ldb, ldc);
    if (info) return info;
    if (n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return 0;
    const bool tr = !flag_n(trans);
    const int64_t ar = tr ? k : n, ac = tr ? n : k;
    MBH_HOST_STAGE("syr2k");
    Staged A, B, C;
    hipError_t e;
    if ((e = stage_in(0, a, ar, ac, lda, A, hs)) != hipSuccess) return hip_fail(e, "syr2k: stage A");
    if ((e = stage_in(1, b, ar, ac, ldb, B, hs)) != hipSuccess) return hip_fail(e, "syr2k: stage B");
    if ((e = stage_in(2, c, n, n, ldc, C, hs)) != hipSuccess) return hip_fail(e, "syr2k: stage C");
    if ((e = syr2k<T>(uplo, trans, n, k, alpha, (const T*)A.p, max1(ar), (const T*)B.p, max1(ar), beta, (T*)C.p, n,
                      host_precision(), hs)) != hipSuccess)
        return hip_fail(e, "syr2k");
    return finish(stage_out(c, C, ldc, hs), "syr2k: copy C back");
}

template <typename T>
int symm_dev(char side, char uplo, int64_t m, int64_t n, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb,
             T beta, T* c, int64_t ldc, int prec, void* stream) {
    int info = check_symm(side, uplo, m, n, lda, ldb, ldc);
    if (info) return info;
    if (sizeof(T) == 4 && !gemm_precision_ok(prec)) {
        set_error("symm: invalid precision %d", prec);
        return MBH_E_PRECISION;
    }
    if (m == 0 || n == 0) return 0;
    MBH_LEASE("symm");
    return finish(symm<T>(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, prec, (hipStream_t)stream), "symm");
}

template <typename T>
int symm_host(char side, char uplo, int64_t m, int64_t n, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb,
              T beta, T* c, int64_t ldc) {
    int info = check_symm(side, uplo, m, n, lda, ldb, ldc);
    if (info) return info;
    if (m == 0 || n == 0 || (alpha == T(0) && beta == T(1))) return 0;
    const int64_t na = flag_l(side) ? m : n;
    MBH_HOST_STAGE("symm");
    Staged A, B, C;
    hipError_t e;
    if ((e = stage_in(0, a, na, na, lda, A, hs)) != hipSuccess) return hip_fail(e, "symm: stage A");
    if ((e = stage_in(1, b, m, n, ldb, B, hs)) != hipSuccess) return hip_fail(e, "symm: stage B");
    if ((e = stage_in(2, c, m, n, ldc, C, hs)) != hipSuccess) return hip_fail(e, "symm: stage C");
    if ((e = symm<T>(side, uplo, m, n, alpha, (const T*)A.p, max1(na), (const T*)B.p, m, beta, (T*)C.p, m,
                     host_precision(), hs)) != hipSuccess)
        return hip_fail(e, "symm");
    return finish(stage_out(c, C, ldc, hs), "symm: copy C back");
}

}  // namespace
}  // namespace mbh

using namespace mbh;

extern "C" {

int mbh_strsm_host(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float alpha, const float* a,
                   int64_t lda, float* b, int64_t ldb) {
    return trsm_host<float>(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}
int mbh_dtrsm_host(char side, char uplo, char transa, char diag, int64_t m, int64_t n, double alpha, const double* a,
                   int64_t lda, double* b, int64_t ldb) {
    return trsm_host<double>(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}
int mbh_strsm_dev(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float alpha, const float* a,
                  int64_t lda, float* b, int64_t ldb, int precision, int refine, void* stream) {
    return trsm_dev<float>(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, precision, refine, stream);
}
int mbh_dtrsm_dev(char side, char uplo, char transa, char diag, int64_t m, int64_t n, double alpha, const double* a,
                  int64_t lda, double* b, int64_t ldb, void* stream) {
    return trsm_dev<double>(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, MBH_PREC_FP32, 0, stream);
}

int mbh_strmm_host(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float alpha, const float* a,
                   int64_t lda, float* b, int64_t ldb) {
    return trmm_host<float>(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}
int mbh_dtrmm_host(char side, char uplo, char transa, char diag, int64_t m, int64_t n, double alpha, const double* a,
                   int64_t lda, double* b, int64_t ldb) {
    return trmm_host<double>(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}
int mbh_strmm_dev(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float alpha, const float* a,
                  int64_t lda, float* b, int64_t ldb, int precision, void* stream) {
    return trmm_dev<float>(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, precision, stream);
}
int mbh_dtrmm_dev(char side, char uplo, char transa, char diag, int64_t m, int64_t n, double alpha, const double* a,
                  int64_t lda, double* b, int64_t ldb, void* stream) {
    return trmm_dev<double>(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, MBH_PREC_FP32, stream);
}

int mbh_ssyrk_host(char uplo, char trans, int64_t n, int64_t k, float alpha, const float* a, int64_t lda, float beta,
                   float* c, int64_t ldc) {
    return syrk_host<float>(uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}
int mbh_dsyrk_host(char uplo, char trans, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                   double beta, double* c, int64_t ldc) {
    return syrk_host<double>(uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}
int mbh_ssyrk_dev(char uplo, char trans, int64_t n, int64_t k, float alpha, const float* a, int64_t lda, float beta,
                  float* c, int64_t ldc, int precision, void* stream) {
    return syrk_dev<float>(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, precision, stream);
}
int mbh_dsyrk_dev(char uplo, char trans, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                  double beta, double* c, int64_t ldc, void* stream) {
    return syrk_dev<double>(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, MBH_PREC_FP32, stream);
}

int mbh_ssyr2k_host(char uplo, char trans, int64_t n, int64_t k, float alpha, const float* a, int64_t lda,
                    const float* b, int64_t ldb, float beta, float* c, int64_t ldc) {
    return syr2k_host<float>(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
int mbh_dsyr2k_host(char uplo, char trans, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                    const double* b, int64_t ldb, double beta, double* c, int64_t ldc) {
    return syr2k_host<double>(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
int mbh_ssyr2k_dev(char uplo, char trans, int64_t n, int64_t k, float alpha, const float* a, int64_t lda,
                   const float* b, int64_t ldb, float beta, float* c, int64_t ldc, int precision, void* stream) {
    return syr2k_dev<float>(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc, precision, stream);
}
int mbh_dsyr2k_dev(char uplo, char trans, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                   const double* b, int64_t ldb, double beta, double* c, int64_t ldc, void* stream) {
    return syr2k_dev<double>(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc, MBH_PREC_FP32, stream);
}

int mbh_ssymm_host(char side, char uplo, int64_t m, int64_t n, float alpha, const float* a, int64_t lda,
                   const float* b, int64_t ldb, float beta, float* c, int64_t ldc) {
    return symm_host<float>(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);
}
int mbh_dsymm_host(char side, char uplo, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
                   const double* b, int64_t ldb, double beta, double* c, int64_t ldc) {
    return symm_host<double>(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);
}
int mbh_ssymm_dev(char side, char uplo, int64_t m, int64_t n, float alpha, const float* a, int64_t lda,
                  const float* b, int64_t ldb, float beta, float* c, int64_t ldc, int precision, void* stream) {
    return symm_dev<float>(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, precision, stream);
}
int mbh_dsymm_dev(char side, char uplo, int64_t m, int64_t n, double alpha, const double* a, int64_t lda,
                  const double* b, int64_t ldb, double beta, double* c, int64_t ldc, void* stream) {
    return symm_dev<double>(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, MBH_PREC_FP32, stream);
}

}  // extern "C"
