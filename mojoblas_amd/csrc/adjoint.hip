// Adjoints (reverse-mode derivatives) of gemm and trsm on the same kernels.
//
// gemm  C = alpha*op(A)*op(B) + beta*C0:
//   d op(A) = alpha * dC * op(B)^T,  d op(B) = alpha * op(A)^T * dC
//   (beta*dC for C0 is left to the caller); each case is one gemm whose
//   transpose flags absorb op() on both sides, so no operand is transposed in
//   memory.
// trsm  op(A) X = alpha B (side L) or X op(A) = alpha B (side R):
//   G = op(A)^-T dX (L) or dX op(A)^-T (R)   -- a trsm with the opposite
//                                               transpose flag
//   dB = alpha * G
//   d op(A) = -G X^T (L) or -X^T G (R), kept on A's stored triangle (zero
//   elsewhere, and on the diagonal when it is unit) -- a gemm over the
//   block-triangle tile grid with a triangle-masked store.
#include "../../include/mojoblas_hip.h"
#include "abi_internal.hpp"
#include "common.hpp"
#include "dispatch.hpp"
#include "level3.hpp"

namespace mbh {
namespace {

int64_t max1(int64_t v) { return v > 1 ? v : 1; }

__global__ void zero_diag_kernel(int64_t n, float* a, int64_t lda) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i + i * lda] = 0.0f;
}

__global__ void copy_f32_kernel(int64_t m, int64_t n, const float* in, int64_t ldi, float* out, int64_t ldo) {
    const int64_t total = m * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % m, j = idx / m;
        out[i + j * ldo] = in[i + j * ldi];
    }
}

dim3 elems_grid(int64_t count) {
    int64_t b = (count + 255) / 256;
    return dim3((unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b)));
}

}  // namespace
}  // namespace mbh

using namespace mbh;

extern "C" {

int mbh_sgemm_backward(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float* a,
                       int64_t lda, const float* b, int64_t ldb, const float* dc, int64_t lddc, float* da,
                       int64_t ldda, float* db, int64_t lddb, int precision, void* stream) {
    int info = check_gemm(transa, transb, m, n, k, lda, ldb, max1(m));
    if (info) return info;
    const bool at = !flag_n(transa), bt = !flag_n(transb);
    if (lddc < max1(m) || (da && ldda < max1(at ? k : m)) || (db && lddb < max1(bt ? n : k))) {
        set_error("mbh_sgemm_backward: leading dimension of dC, dA or dB too small");
        return MBH_E_ARG;
    }
    if (!gemm_precision_ok(precision)) {
        set_error("mbh_sgemm_backward: invalid precision %d", precision);
        return MBH_E_PRECISION;
    }
    hipStream_t s = (hipStream_t)stream;
    StreamLease lease(s);
    if (lease.error() != hipSuccess) return hip_fail(lease.error(), "mbh_sgemm_backward");
    hipError_t e = hipSuccess;
    if (da) {
        if (!at)  // dA (m x k) = alpha * dC * op(B)^T
            e = sgemm(false, !bt, m, k, n, alpha, dc, lddc, b, ldb, 0.0f, da, ldda, precision, s);
        else      // dA (k x m) = alpha * op(B) * dC^T
            e = sgemm(bt, true, k, m, n, alpha, b, ldb, dc, lddc, 0.0f, da, ldda, precision, s);
        if (e != hipSuccess) return hip_fail(e, "mbh_sgemm_backward dA");
    }
    if (db) {
        if (!bt)  // dB (k x n) = alpha * op(A)^T * dC
            e = sgemm(!at, false, k, n, m, alpha, a, lda, dc, lddc, 0.0f, db, lddb, precision, s);
        else      // dB (n x k) = alpha * dC^T * op(A)
            e = sgemm(true, at, n, k, m, alpha, dc, lddc, a, lda, 0.0f, db, lddb, precision, s);
        if (e != hipSuccess) return hip_fail(e, "mbh_sgemm_backward dB");
    }
    return 0;
}

int mbh_strsm_backward(char side, char uplo, char transa, char diag, int64_t m, int64_t n, float alpha,
                       const float* a, int64_t lda, const float* x, int64_t ldx, const float* dx, int64_t lddx,
                       float* db, int64_t lddb, float* da, int64_t ldda, int precision, void* stream) {
    int info = check_trxm(side, uplo, transa, diag, m, n, lda, max1(m));
    if (info) return info;
    const bool left = flag_l(side), upper = flag_u(uplo), trans = !flag_n(transa), unit = flag_u(diag);
    const int64_t na = left ? m : n;
    if (!db) {
        set_error("mbh_strsm_backward: dB is required (it holds op(A)^-T dX for dA)");
        return MBH_E_ARG;
    }
    if (ldx < max1(m) || lddx < max1(m) || lddb < max1(m) || (da && ldda < max1(na))) {
        set_error("mbh_strsm_backward: leading dimension of X, dX, dB or dA too small");
        return MBH_E_ARG;
    }
    if (!gemm_precision_ok(precision)) {
        set_error("mbh_strsm_backward: invalid precision %d", precision);
        return MBH_E_PRECISION;
    }
    StreamLease lease((hipStream_t)stream);
    if (lease.error() != hipSuccess) return hip_fail(lease.error(), "mbh_strsm_backward");
    if (m == 0 || n == 0) {
        if (da && na > 0) {
            hipError_t e = scale<float>(na, na, 0.0f, da, ldda, (hipStream_t)stream);
            if (e != hipSuccess) return hip_fail(e, "mbh_strsm_backward zero dA");
        }
        return 0;
    }
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    // G = op(A)^-T dX (L) / dX op(A)^-T (R), formed in dB.
    copy_f32_kernel<<<elems_grid(m * n), 256, 0, s>>>(m, n, dx, lddx, db, lddb);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "mbh_strsm_backward copy");
    const char tflip = trans ? 'N' : 'T';
    if ((e = trsm<float>(side, uplo, tflip, diag, m, n, 1.0f, a, lda, db, lddb, precision, 1, s)) != hipSuccess)
        return hip_fail(e, "mbh_strsm_backward solve");
    if (da) {
        const int tri = upper ? TRI_UPPER : TRI_LOWER, grid = upper ? 2 : 1;
        if ((e = scale<float>(na, na, 0.0f, da, ldda, s)) != hipSuccess) return hip_fail(e, "zero dA");
        if (left) {
            if (!trans)  // dA = -G X^T
                e = sgemm(false, true, m, m, n, -1.0f, db, lddb, x, ldx, 0.0f, da, ldda, precision, s, tri, grid);
            else         // dA = (-G X^T)^T = -X G^T
                e = sgemm(false, true, m, m, n, -1.0f, x, ldx, db, lddb, 0.0f, da, ldda, precision, s, tri, grid);
        } else {
            if (!trans)  // dA = -X^T G
                e = sgemm(true, false, n, n, m, -1.0f, x, ldx, db, lddb, 0.0f, da, ldda, precision, s, tri, grid);
            else         // dA = (-X^T G)^T = -G^T X
                e = sgemm(true, false, n, n, m, -1.0f, db, lddb, x, ldx, 0.0f, da, ldda, precision, s, tri, grid);
        }
        if (e != hipSuccess) return hip_fail(e, "mbh_strsm_backward dA");
        if (unit) {
            zero_diag_kernel<<<elems_grid(na), 256, 0, s>>>(na, da, ldda);
            if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "mbh_strsm_backward diag");
        }
    }
    if (alpha != 1.0f)
        if ((e = scale<float>(m, n, alpha, db, lddb, s)) != hipSuccess) return hip_fail(e, "mbh_strsm_backward dB");
    return 0;
}

}  // extern "C"
