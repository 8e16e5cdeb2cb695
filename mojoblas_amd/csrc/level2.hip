// Level-2 gemv and trsv on the GPU (mojoBLAS src/level2/gemv.mojo:23-212,
// src/level2/trsv.mojo:20-214), with the reference's argument checks.
//
//   gemv 'N'  y := alpha*A*x + beta*y.  A workgroup owns 256 rows and a chunk
//             of columns; each thread walks its row through the chunk, so the
//             64 lanes of a wave read 256 contiguous bytes of a column per
//             load (any lda, any alignment), x(j) is a uniform load, eight
//             columns are in flight per thread.  Column chunks are sized for
//             about 2048 workgroups (HBM-bound: A is read once); their partial
//             sums go to a workspace and a second kernel adds them in a fixed
//             order and applies alpha and beta.
//   gemv 'T'  y := alpha*A^T*x + beta*y.  A wave owns four columns and a chunk
//             of rows; lanes stride the rows (coalesced), x(i) is read once for
//             the four columns, each column is reduced across the wave by
//             shuffles; row chunks give the workgroup count, reduced as above.
//   trsv      x := op(A)^-1 x, the n x 1 case of the blocked trsm (fp64
//             diagonal-block inverses, gemm updates, one fp64-residual
//             refinement for fp32 data), on a contiguous copy when incx != 1.
// Strided vectors follow BLAS: for inc < 0 logical element i is stored at
// (len-1-i)*|inc|.  The results are deterministic (fixed summation order).
#include <algorithm>
#include <climits>

#include "../../include/mojoblas_hip.h"
#include "abi_internal.hpp"
#include "common.hpp"
#include "dispatch.hpp"
#include "level3.hpp"

namespace mbh {
namespace {

constexpr int GV_NT = 256, GV_UNR = 8;  // 'N': rows per workgroup, columns in flight
constexpr int GT_COLS = 4;              // 'T': columns per wave (16 per workgroup)
constexpr int64_t GV_TARGET_WGS = 2048;

__host__ __device__ inline int64_t vidx(int64_t i, int64_t len, int64_t inc) {
    return inc > 0 ? i * inc : (len - 1 - i) * (-inc);
}

template <typename T>
__device__ inline void gv_store(T* y, int64_t i, int64_t len, int64_t incy, T alpha, T beta, T s) {
    T* yp = y + vidx(i, len, incy);
    *yp = (beta == T(0)) ? alpha * s : alpha * s + beta * *yp;
}

template <typename T>
__global__ __launch_bounds__(GV_NT) void gemv_n_kernel(int64_t m, int64_t n, const T* __restrict__ A, int64_t lda,
                                                       const T* __restrict__ x, int64_t incx, int64_t chunk, T alpha,
                                                       T beta, T* y, int64_t incy, T* part) {
    const int64_t i = (int64_t)blockIdx.x * GV_NT + threadIdx.x;
    if (i >= m) return;
    const int64_t j0 = (int64_t)blockIdx.y * chunk, j1 = (j0 + chunk < n) ? j0 + chunk : n;
    const MBH_GLOBAL T* a = (const MBH_GLOBAL T*)A + i;
    const MBH_GLOBAL T* xg = (const MBH_GLOBAL T*)x;
    T acc = T(0);
    int64_t j = j0;
    for (; j + GV_UNR <= j1; j += GV_UNR) {
        T av[GV_UNR], xv[GV_UNR];
#pragma unroll
        for (int u = 0; u < GV_UNR; ++u) {
            av[u] = a[(j + u) * lda];
            xv[u] = xg[vidx(j + u, n, incx)];
        }
#pragma unroll
        for (int u = 0; u < GV_UNR; ++u) acc += av[u] * xv[u];
    }
    for (; j < j1; ++j) acc += a[j * lda] * xg[vidx(j, n, incx)];
    if (part) part[(int64_t)blockIdx.y * m + i] = acc;
    else gv_store(y, i, m, incy, alpha, beta, acc);
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename T>
__global__ __launch_bounds__(256) void gemv_t_kernel(int64_t m, int64_t n, const T* __restrict__ A, int64_t lda,
                                                     const T* __restrict__ x, int64_t incx, int64_t chunk, T alpha,
                                                     T beta, T* y, int64_t incy, T* part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t jb = ((int64_t)blockIdx.x * 4 + wave) * GT_COLS;
    if (jb >= n) return;  // whole wave
    const int nc = (n - jb < GT_COLS) ? (int)(n - jb) : GT_COLS;
    const int64_t i0 = (int64_t)blockIdx.y * chunk, i1 = (i0 + chunk < m) ? i0 + chunk : m;
    const MBH_GLOBAL T* a = (const MBH_GLOBAL T*)A + jb * lda;
    const MBH_GLOBAL T* xg = (const MBH_GLOBAL T*)x;
    // Column c of the wave reads column min(c, nc-1) (in range) and drops the
    // result when c >= nc, so the loads carry no branches.
    int64_t coff[GT_COLS];
#pragma unroll
    for (int c = 0; c < GT_COLS; ++c) coff[c] = (int64_t)(c < nc ? c : nc - 1) * lda;
    T acc[GT_COLS];
#pragma unroll
    for (int c = 0; c < GT_COLS; ++c) acc[c] = T(0);
    int64_t base = i0;
    for (; base + 256 <= i1; base += 256) {  // four rows per lane in flight
        T xv[4], av[4][GT_COLS];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t i = base + lane + 64 * u;
            xv[u] = xg[vidx(i, m, incx)];
#pragma unroll
            for (int c = 0; c < GT_COLS; ++c) av[u][c] = a[coff[c] + i];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < GT_COLS; ++c) acc[c] += av[u][c] * xv[u];
    }
    for (int64_t i = base + lane; i < i1; i += 64) {
        const T xi = xg[vidx(i, m, incx)];
#pragma unroll
        for (int c = 0; c < GT_COLS; ++c) acc[c] += a[coff[c] + i] * xi;
    }
#pragma unroll
    for (int c = 0; c < GT_COLS; ++c) acc[c] = wave_sum(acc[c]);
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < GT_COLS; ++c) {
            if (c >= nc) break;
            if (part) part[(int64_t)blockIdx.y * n + jb + c] = acc[c];
            else gv_store(y, jb + c, n, incy, alpha, beta, acc[c]);
        }
    }
}

template <typename T>
__global__ void gemv_reduce_kernel(int64_t len, int splits, const T* __restrict__ part, T alpha, T beta, T* y,
                                   int64_t incy) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        T s = T(0);
        for (int k = 0; k < splits; ++k) s += part[(int64_t)k * len + i];
        gv_store(y, i, len, incy, alpha, beta, s);
    }
}

// y := beta*y (beta == 0 writes zeros without reading).
template <typename T>
__global__ void vec_scale_kernel(int64_t len, T beta, T* y, int64_t incy) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        T* yp = y + vidx(i, len, incy);
        *yp = (beta == T(0)) ? T(0) : beta * *yp;
    }
}

// dst(i) = src(i) for logical elements of two strided vectors.
template <typename T>
__global__ void vec_copy_kernel(int64_t len, const T* src, int64_t incs, T* dst, int64_t incd) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
        dst[vidx(i, len, incd)] = src[vidx(i, len, incs)];
}

dim3 grid1(int64_t len) {
    int64_t b = (len + 255) / 256;
    return dim3((unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b)));
}

}  // namespace

template <typename T>
hipError_t gemv(bool trans, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, const T* x, int64_t incx, T beta,
                T* y, int64_t incy, hipStream_t s) {
    const int64_t leny = trans ? n : m;
    if (m <= 0 || n <= 0 || (alpha == T(0) && beta == T(1))) return hipSuccess;
    if (alpha == T(0)) {
        vec_scale_kernel<T><<<grid1(leny), 256, 0, s>>>(leny, beta, y, incy);
        return hipGetLastError();
    }
    int64_t splits, chunk, groups;
    if (!trans) {
        groups = (m + GV_NT - 1) / GV_NT;
        splits = std::max<int64_t>(1, std::min((GV_TARGET_WGS + groups - 1) / groups, (n + 63) / 64));
        chunk = (n + splits - 1) / splits;
        chunk = (chunk + GV_UNR - 1) / GV_UNR * GV_UNR;
    } else {
        groups = (n + 4 * GT_COLS - 1) / (4 * GT_COLS);
        splits = std::max<int64_t>(1, std::min((GV_TARGET_WGS + groups - 1) / groups, (m + 255) / 256));
        chunk = (m + splits - 1) / splits;
        chunk = (chunk + 255) / 256 * 256;
    }
    splits = (((trans ? m : n) + chunk - 1) / chunk);
    if (groups > INT_MAX || splits > 65535) return hipErrorInvalidValue;
    T* part = nullptr;
    if (splits > 1) {
        void* ws = nullptr;
        hipError_t e = workspace(s, 2, (size_t)splits * leny * sizeof(T), &ws);
        if (e != hipSuccess) return e;
        part = (T*)ws;
    }
    const dim3 grid((unsigned)groups, (unsigned)splits);
    if (!trans) gemv_n_kernel<T><<<grid, GV_NT, 0, s>>>(m, n, A, lda, x, incx, chunk, alpha, beta, y, incy, part);
    else gemv_t_kernel<T><<<grid, 256, 0, s>>>(m, n, A, lda, x, incx, chunk, alpha, beta, y, incy, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || splits == 1) return e;
    gemv_reduce_kernel<T><<<grid1(leny), 256, 0, s>>>(leny, (int)splits, part, alpha, beta, y, incy);
    return hipGetLastError();
}

template <typename T>
hipError_t trsv(char uplo, char trans, char diag, int64_t n, const T* A, int64_t lda, T* x, int64_t incx, int prec,
                hipStream_t s) {
    if (n <= 0) return hipSuccess;
    T* xc = x;
    if (incx != 1) {
        void* ws = nullptr;
        hipError_t e = workspace(s, 3, (size_t)n * sizeof(T), &ws);
        if (e != hipSuccess) return e;
        xc = (T*)ws;
        vec_copy_kernel<T><<<grid1(n), 256, 0, s>>>(n, x, incx, xc, 1);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipError_t e = trsm<T>('L', uplo, trans, diag, n, 1, T(1), A, lda, xc, n, prec, sizeof(T) == 4 ? 1 : 0, s);
    if (e != hipSuccess || incx == 1) return e;
    vec_copy_kernel<T><<<grid1(n), 256, 0, s>>>(n, xc, 1, x, incx);
    return hipGetLastError();
}

namespace {

int64_t max1(int64_t v) { return v > 1 ? v : 1; }
int64_t span(int64_t len, int64_t inc) { return len > 0 ? (len - 1) * (inc < 0 ? -inc : inc) + 1 : 0; }
int finish(hipError_t e, const char* where) { return e == hipSuccess ? 0 : hip_fail(e, where); }

// gemv.mojo:76-91: trans is compared against "N", "T" and "C" only.
int check_gemv(char trans, int64_t m, int64_t n, int64_t lda, int64_t incx, int64_t incy) {
    if (trans != 'N' && trans != 'T' && trans != 'C') return 1;
    if (m < 0) return 2;
    if (n < 0) return 3;
    if (lda < max1(m)) return 6;
    if (incx == 0) return 8;
    if (incy == 0) return 11;
    return 0;
}

// trsv.mojo:57-80
int check_trsv(char uplo, char trans, char diag, int64_t n, int64_t lda, int64_t incx) {
    if (!flag_u(uplo) && !flag_l(uplo)) return 1;
    if (!flag_n(trans) && !flag_t(trans)) return 2;
    if (!flag_u(diag) && !flag_n(diag)) return 3;
    if (n < 0) return 4;
    if (lda < max1(n)) return 6;
    if (incx == 0) return 8;
    return 0;
}

template <typename T>
int gemv_dev(char trans, int64_t m, int64_t n, T alpha, const T* a, int64_t lda, const T* x, int64_t incx, T beta,
             T* y, int64_t incy, void* stream) {
    int info = check_gemv(trans, m, n, lda, incx, incy);
    if (info) return info;
    return finish(gemv<T>(trans != 'N', m, n, alpha, a, lda, x, incx, beta, y, incy, (hipStream_t)stream), "gemv");
}

// Host staging: A packed (ld = m) in slot 0, the whole strided spans of x and
// y in slots 1 and 2 (y's gaps travel both ways unchanged).
template <typename T>
int gemv_host(char trans, int64_t m, int64_t n, T alpha, const T* a, int64_t lda, const T* x, int64_t incx, T beta,
              T* y, int64_t incy) {
    int info = check_gemv(trans, m, n, lda, incx, incy);
    if (info) return info;
    if (m == 0 || n == 0 || (alpha == T(0) && beta == T(1))) return 0;
    const bool tr = trans != 'N';
    const int64_t lenx = tr ? m : n, leny = tr ? n : m, sx = span(lenx, incx), sy = span(leny, incy);
    auto lk = stage_lock();
    void *dA = nullptr, *dx = nullptr, *dy = nullptr;
    hipError_t e;
    if (alpha != T(0)) {
        if ((e = slot_get(0, (size_t)m * n * sizeof(T), &dA)) != hipSuccess) return hip_fail(e, "gemv: A");
        if ((e = put_matrix(dA, a, m, n, lda, sizeof(T))) != hipSuccess) return hip_fail(e, "gemv: copy A");
        if ((e = slot_get(1, (size_t)sx * sizeof(T), &dx)) != hipSuccess) return hip_fail(e, "gemv: x");
        if ((e = put_matrix(dx, x, sx, 1, sx, sizeof(T))) != hipSuccess) return hip_fail(e, "gemv: copy x");
    }
    if ((e = slot_get(2, (size_t)sy * sizeof(T), &dy)) != hipSuccess) return hip_fail(e, "gemv: y");
    if ((e = put_matrix(dy, y, sy, 1, sy, sizeof(T))) != hipSuccess) return hip_fail(e, "gemv: copy y");
    if ((e = gemv<T>(tr, m, n, alpha, (const T*)dA, max1(m), (const T*)dx, incx, beta, (T*)dy, incy, nullptr)) !=
        hipSuccess)
        return hip_fail(e, "gemv");
    return finish(get_matrix(y, dy, sy, 1, sy, sizeof(T)), "gemv: copy y back");
}

template <typename T>
int trsv_dev(char uplo, char trans, char diag, int64_t n, const T* a, int64_t lda, T* x, int64_t incx, int prec,
             void* stream) {
    int info = check_trsv(uplo, trans, diag, n, lda, incx);
    if (info) return info;
    if (sizeof(T) == 4 && !gemm_precision_ok(prec)) {
        set_error("trsv: invalid precision %d", prec);
        return MBH_E_PRECISION;
    }
    return finish(trsv<T>(uplo, trans, diag, n, a, lda, x, incx, prec, (hipStream_t)stream), "trsv");
}

template <typename T>
int trsv_host(char uplo, char trans, char diag, int64_t n, const T* a, int64_t lda, T* x, int64_t incx) {
    int info = check_trsv(uplo, trans, diag, n, lda, incx);
    if (info) return info;
    if (n == 0) return 0;
    const int64_t sx = span(n, incx);
    auto lk = stage_lock();
    void *dA = nullptr, *dx = nullptr;
    hipError_t e;
    if ((e = slot_get(0, (size_t)n * n * sizeof(T), &dA)) != hipSuccess) return hip_fail(e, "trsv: A");
    if ((e = put_matrix(dA, a, n, n, lda, sizeof(T))) != hipSuccess) return hip_fail(e, "trsv: copy A");
    if ((e = slot_get(1, (size_t)sx * sizeof(T), &dx)) != hipSuccess) return hip_fail(e, "trsv: x");
    if ((e = put_matrix(dx, x, sx, 1, sx, sizeof(T))) != hipSuccess) return hip_fail(e, "trsv: copy x");
    if ((e = trsv<T>(uplo, trans, diag, n, (const T*)dA, max1(n), (T*)dx, incx, default_precision(), nullptr)) !=
        hipSuccess)
        return hip_fail(e, "trsv");
    return finish(get_matrix(x, dx, sx, 1, sx, sizeof(T)), "trsv: copy x back");
}

}  // namespace
}  // namespace mbh

using namespace mbh;

extern "C" {

int mbh_sgemv_host(char trans, int64_t m, int64_t n, float alpha, const float* a, int64_t lda, const float* x,
                   int64_t incx, float beta, float* y, int64_t incy) {
    return gemv_host<float>(trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
}
int mbh_dgemv_host(char trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda, const double* x,
                   int64_t incx, double beta, double* y, int64_t incy) {
    return gemv_host<double>(trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
}
int mbh_sgemv_dev(char trans, int64_t m, int64_t n, float alpha, const float* a, int64_t lda, const float* x,
                  int64_t incx, float beta, float* y, int64_t incy, void* stream) {
    return gemv_dev<float>(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, stream);
}
int mbh_dgemv_dev(char trans, int64_t m, int64_t n, double alpha, const double* a, int64_t lda, const double* x,
                  int64_t incx, double beta, double* y, int64_t incy, void* stream) {
    return gemv_dev<double>(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, stream);
}
int mbh_strsv_host(char uplo, char trans, char diag, int64_t n, const float* a, int64_t lda, float* x, int64_t incx) {
    return trsv_host<float>(uplo, trans, diag, n, a, lda, x, incx);
}
int mbh_dtrsv_host(char uplo, char trans, char diag, int64_t n, const double* a, int64_t lda, double* x,
                   int64_t incx) {
    return trsv_host<double>(uplo, trans, diag, n, a, lda, x, incx);
}
int mbh_strsv_dev(char uplo, char trans, char diag, int64_t n, const float* a, int64_t lda, float* x, int64_t incx,
                  int precision, void* stream) {
    return trsv_dev<float>(uplo, trans, diag, n, a, lda, x, incx, precision, stream);
}
int mbh_dtrsv_dev(char uplo, char trans, char diag, int64_t n, const double* a, int64_t lda, double* x, int64_t incx,
                  void* stream) {
    return trsv_dev<double>(uplo, trans, diag, n, a, lda, x, incx, 0, stream);
}

}  // extern "C"
