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
//   trsv      x := op(A)^-1 x in one kernel: blocked substitution with
//             decoupled look-back.  One workgroup per 128-row block, in solve
//             order by an atomic ticket; it streams its off-diagonal blocks of
//             op(A) while earlier blocks are solved, waits on each earlier
//             block's flag, accumulates the residual in fp64, applies the fp64
//             inverse of its diagonal block and publishes its part of x and its
//             flag.  On a contiguous copy of x when incx != 1.
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

// 'T' with 16-byte loads (A 16-byte aligned, lda a multiple of V): each
// lane reads V consecutive rows of its wave's four columns per load, four
// such groups in flight; x is read with 16-byte loads too when it is
// contiguous and aligned (xvec).  Rows past the last whole 256 V block take
// single loads.
template <typename T>
__global__ __launch_bounds__(256) void gemv_tv_kernel(int64_t m, int64_t n, const T* __restrict__ A, int64_t lda,
                                                      const T* __restrict__ x, int64_t incx, int64_t chunk, T alpha,
                                                      T beta, T* y, int64_t incy, T* part, int xvec) {
    constexpr int V = 16 / sizeof(T);
    typedef T vT __attribute__((ext_vector_type(V)));
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
    vT accv[GT_COLS];
#pragma unroll
    for (int c = 0; c < GT_COLS; ++c) accv[c] = vT(T(0));
    int64_t base = i0;
    for (; base + 256 * V <= i1; base += 256 * V) {  // four groups of V rows per lane in flight
        vT xv[4], av[4][GT_COLS];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t i = base + V * lane + 64 * V * u;
            if (xvec) {
                xv[u] = *(const MBH_GLOBAL vT*)(xg + i);
            } else {
#pragma unroll
                for (int e = 0; e < V; ++e) xv[u][e] = xg[vidx(i + e, m, incx)];
            }
#pragma unroll
            for (int c = 0; c < GT_COLS; ++c) av[u][c] = *(const MBH_GLOBAL vT*)(a + coff[c] + i);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < GT_COLS; ++c) accv[c] += av[u][c] * xv[u];
    }
    T acc[GT_COLS];
#pragma unroll
    for (int c = 0; c < GT_COLS; ++c) {
        acc[c] = T(0);
#pragma unroll
        for (int e = 0; e < V; ++e) acc[c] += accv[c][e];
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

// ---- trsv ---------------------------------------------------------------------
// A workgroup waits only on blocks with smaller tickets, which were taken by
// workgroups that had already started, so the waits cannot deadlock.  Blocks
// of x and the flags are written and read with device-scope (agent) atomics:
// the XCDs' L2 caches are not coherent with each other for plain accesses.
// The wait is bounded (TV_SPIN polls, and no longer once any workgroup has
// given up); a workgroup that gives up sets *fail and finishes with what it
// has, so the kernel always terminates.
constexpr int TV_NB = TRI_NB;  // rows per block (the fp64 inverses' blocks)
constexpr int TV_NT = 256;     // thread t: row t % 128, column half t / 128
constexpr int TV_H = TV_NB / 2;
constexpr int TV_SPIN = 1 << 22;  // about 0.2 s of 128-cycle sleeps

__global__ void trsv_reset_kernel(int* w, int count) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) w[i] = 0;
}

template <typename T>
__global__ __launch_bounds__(TV_NT) void trsv_lookback_kernel(int64_t n, const T* __restrict__ A, int64_t lda,
                                                              int eff_lower, int trans, int avec,
                                                              const double* __restrict__ inv, T* x, int* flags,
                                                              int* ticket, int nblk, int* fail) {
    constexpr int VEC = 16 / (int)sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VEC)));
    __shared__ double xs[TV_NB];
    __shared__ double red[2][TV_NB];
    __shared__ int s_sh;
    const int tid = threadIdx.x, row = tid & (TV_NB - 1), h = tid >> 7;
    if (tid == 0) s_sh = atomicAdd(ticket, 1);
    __syncthreads();
    const int s = s_sh;  // solve-order index (grid = nblk, so s < nblk)
    const int b = eff_lower ? s : nblk - 1 - s;
    const int64_t r0 = (int64_t)b * TV_NB;
    const int nb = (n - r0 < TV_NB) ? (int)(n - r0) : TV_NB;
    const bool rok = row < nb;
    // This thread's row of the diagonal block's inverse, its half of the k's.
    const double* ib = inv + (size_t)b * TV_NB * TV_NB;
    double iv[TV_H];
#pragma unroll
    for (int q = 0; q < TV_H; ++q) iv[q] = ib[row + (size_t)(h * TV_H + q) * TV_NB];
    const MBH_GLOBAL T* Ag = (const MBH_GLOBAL T*)A;
    double acc = 0.0;
    for (int sp = 0; sp < s; ++sp) {
        const int c = eff_lower ? sp : nblk - 1 - sp;
        const int64_t c0 = (int64_t)c * TV_NB;
        const int nc = (n - c0 < TV_NB) ? (int)(n - c0) : TV_NB;
        // av[q] = op(A)(r0 + row, c0 + h*64 + q), masked loads without
        // branches, issued before the wait so that they overlap it
        T av[TV_H];
        if (!trans) {  // A[(r0+row) + (c0+j)*lda]: the lanes read consecutive rows
#pragma unroll
            for (int q = 0; q < TV_H; ++q) {
                const int j = h * TV_H + q;
                const bool ok = rok && j < nc;
                av[q] = Ag[ok ? (r0 + row) + (c0 + j) * lda : 0];
                av[q] = ok ? av[q] : T(0);
            }
        } else if (avec && nc == TV_NB) {  // A[(c0+j) + (r0+row)*lda]: a lane reads its row in 16-byte vectors
#pragma unroll
            for (int q = 0; q < TV_H; q += VEC) {
                const vec_t v = *(const MBH_GLOBAL vec_t*)(Ag + (rok ? (c0 + h * TV_H + q) + (r0 + row) * lda : 0));
#pragma unroll
                for (int e = 0; e < VEC; ++e) av[q + e] = rok ? v[e] : T(0);
            }
        } else {
#pragma unroll
            for (int q = 0; q < TV_H; ++q) {
                const int j = h * TV_H + q;
                const bool ok = rok && j < nc;
                av[q] = Ag[ok ? (c0 + j) + (r0 + row) * lda : 0];
                av[q] = ok ? av[q] : T(0);
            }
        }
        if (tid == 0) {
            int polls = 0;
            while (__hip_atomic_load(flags + c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                // give up past the limit, or at once when another block has
                if (++polls > TV_SPIN ||
                    ((polls & 255) == 0 && __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                    atomicOr(fail, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __syncthreads();
        if (tid < TV_NB)
            xs[tid] = tid < nc ? (double)__hip_atomic_load(x + c0 + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : 0.0;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < TV_H; ++q) acc += (double)av[q] * xs[h * TV_H + q];
        __syncthreads();  // xs is refilled by the next predecessor
    }
    red[h][row] = acc;
    __syncthreads();
    // residual of the block (its right-hand side is still in x: only this
    // workgroup writes these entries), then x_b = inv_b * residual
    if (tid < TV_NB) xs[tid] = tid < nb ? (double)x[r0 + tid] - (red[0][tid] + red[1][tid]) : 0.0;
    __syncthreads();
    double yv = 0.0;
#pragma unroll
    for (int q = 0; q < TV_H; ++q) yv += iv[q] * xs[h * TV_H + q];
    red[h][row] = yv;
    __syncthreads();
    if (tid < nb)
        __hip_atomic_store(x + r0 + tid, (T)(red[0][tid] + red[1][tid]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(flags + b, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
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
    // 'T': 16-byte loads of A when its columns start 16-byte aligned ('N'
    // keeps single loads along its rows: measured faster there).
    constexpr int V = 16 / sizeof(T);
    const bool avec = trans && ((uintptr_t)A % 16 == 0) && (lda % V == 0);
    const int rv = avec ? V : 1;  // rows per load
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
        chunk = (chunk + 256 * rv - 1) / (256 * rv) * (256 * rv);
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
    if (!trans) {
        gemv_n_kernel<T><<<grid, GV_NT, 0, s>>>(m, n, A, lda, x, incx, chunk, alpha, beta, y, incy, part);
    } else {
        const int xvec = incx == 1 && (uintptr_t)x % 16 == 0;
        if (avec)
            gemv_tv_kernel<T><<<grid, 256, 0, s>>>(m, n, A, lda, x, incx, chunk, alpha, beta, y, incy, part, xvec);
        else
            gemv_t_kernel<T><<<grid, 256, 0, s>>>(m, n, A, lda, x, incx, chunk, alpha, beta, y, incy, part);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || splits == 1) return e;
    gemv_reduce_kernel<T><<<grid1(leny), 256, 0, s>>>(leny, (int)splits, part, alpha, beta, y, incy);
    return hipGetLastError();
}

// Workspace (slot 3): fp64 block inverses | flags, ticket, fail | contiguous x.
template <typename T>
hipError_t trsv(char uplo, char trans, char diag, int64_t n, const T* A, int64_t lda, T* x, int64_t incx,
                hipStream_t s, int** fail_out = nullptr) {
    if (n <= 0) return hipSuccess;
    const bool upper = flag_u(uplo), tr = !flag_n(trans), unit = flag_u(diag);
    const bool eff_lower = (!upper) != tr;
    const int64_t nblk = (n + TV_NB - 1) / TV_NB;
    if (nblk > INT_MAX - 3) return hipErrorInvalidValue;
    auto al = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t inv_b = (size_t)nblk * TV_NB * TV_NB * sizeof(double), fl_b = (size_t)(nblk + 3) * sizeof(int);
    const size_t xc_b = incx != 1 ? (size_t)n * sizeof(T) : 0;
    void* ws = nullptr;
    hipError_t e = workspace(s, 3, al(inv_b) + al(fl_b) + al(xc_b), &ws);
    if (e != hipSuccess) return e;
    double* inv = (double*)ws;
    int* flags = (int*)((char*)ws + al(inv_b));
    int* ticket = flags + nblk;
    int* fail = ticket + 1;
    T* xc = incx != 1 ? (T*)((char*)ws + al(inv_b) + al(fl_b)) : x;
    trsv_reset_kernel<<<1, 256, 0, s>>>(flags, (int)nblk + 3);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = tri_block_inverses<T>(n, A, lda, eff_lower, tr, unit, inv, s)) != hipSuccess) return e;
    if (incx != 1) {
        vec_copy_kernel<T><<<grid1(n), 256, 0, s>>>(n, x, incx, xc, 1);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const int avec = ((uintptr_t)A % 16 == 0) && (lda % (16 / (int64_t)sizeof(T)) == 0);
    trsv_lookback_kernel<T><<<dim3((unsigned)nblk), TV_NT, 0, s>>>(n, A, lda, eff_lower ? 1 : 0, tr ? 1 : 0, avec, inv,
                                                                  xc, flags, ticket, (int)nblk, fail);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    workspace_note_trsv(s, fail);
    if (fail_out) *fail_out = fail;
    if (incx == 1) return hipSuccess;
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
    StreamLease lease((hipStream_t)stream);
    if (lease.error() != hipSuccess) return hip_fail(lease.error(), "gemv");
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
    HostStage st;
    if (st.error() != hipSuccess) return hip_fail(st.error(), "gemv: host staging");
    const hipStream_t hs = st.stream();
    void *dA = nullptr, *dx = nullptr, *dy = nullptr;
    hipError_t e;
    if (alpha != T(0)) {
        if ((e = slot_get(0, (size_t)m * n * sizeof(T), &dA)) != hipSuccess) return hip_fail(e, "gemv: A");
        if ((e = put_matrix(dA, a, m, n, lda, sizeof(T), hs)) != hipSuccess) return hip_fail(e, "gemv: copy A");
        if ((e = slot_get(1, (size_t)sx * sizeof(T), &dx)) != hipSuccess) return hip_fail(e, "gemv: x");
        if ((e = put_matrix(dx, x, sx, 1, sx, sizeof(T), hs)) != hipSuccess) return hip_fail(e, "gemv: copy x");
    }
    if ((e = slot_get(2, (size_t)sy * sizeof(T), &dy)) != hipSuccess) return hip_fail(e, "gemv: y");
    if ((e = put_matrix(dy, y, sy, 1, sy, sizeof(T), hs)) != hipSuccess) return hip_fail(e, "gemv: copy y");
    if ((e = gemv<T>(tr, m, n, alpha, (const T*)dA, max1(m), (const T*)dx, incx, beta, (T*)dy, incy, hs)) !=
        hipSuccess)
        return hip_fail(e, "gemv");
    return finish(get_matrix(y, dy, sy, 1, sy, sizeof(T), hs), "gemv: copy y back");
}

template <typename T>
int trsv_dev(char uplo, char trans, char diag, int64_t n, const T* a, int64_t lda, T* x, int64_t incx, void* stream) {
    int info = check_trsv(uplo, trans, diag, n, lda, incx);
    if (info) return info;
    StreamLease lease((hipStream_t)stream);
    if (lease.error() != hipSuccess) return hip_fail(lease.error(), "trsv");
    return finish(trsv<T>(uplo, trans, diag, n, a, lda, x, incx, (hipStream_t)stream), "trsv");
}

template <typename T>
int trsv_host(char uplo, char trans, char diag, int64_t n, const T* a, int64_t lda, T* x, int64_t incx) {
    int info = check_trsv(uplo, trans, diag, n, lda, incx);
    if (info) return info;
    if (n == 0) return 0;
    const int64_t sx = span(n, incx);
    HostStage st;
    if (st.error() != hipSuccess) return hip_fail(st.error(), "trsv: host staging");
    const hipStream_t hs = st.stream();
    void *dA = nullptr, *dx = nullptr;
    hipError_t e;
    if ((e = slot_get(0, (size_t)n * n * sizeof(T), &dA)) != hipSuccess) return hip_fail(e, "trsv: A");
    if ((e = put_matrix(dA, a, n, n, lda, sizeof(T), hs)) != hipSuccess) return hip_fail(e, "trsv: copy A");
    if ((e = slot_get(1, (size_t)sx * sizeof(T), &dx)) != hipSuccess) return hip_fail(e, "trsv: x");
    if ((e = put_matrix(dx, x, sx, 1, sx, sizeof(T), hs)) != hipSuccess) return hip_fail(e, "trsv: copy x");
    int* fail = nullptr;
    if ((e = trsv<T>(uplo, trans, diag, n, (const T*)dA, max1(n), (T*)dx, incx, hs, &fail)) != hipSuccess)
        return hip_fail(e, "trsv");
    if ((e = get_matrix(x, dx, sx, 1, sx, sizeof(T), hs)) != hipSuccess) return hip_fail(e, "trsv: copy x back");
    int failed = 0;
    if ((e = hipMemcpy(&failed, fail, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "trsv: status");
    if (failed) {
        set_error("trsv: a block waited past its limit for an earlier block");
        return MBH_E_HIP;
    }
    return 0;
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
                  void* stream) {
    return trsv_dev<float>(uplo, trans, diag, n, a, lda, x, incx, stream);
}
int mbh_dtrsv_dev(char uplo, char trans, char diag, int64_t n, const double* a, int64_t lda, double* x, int64_t incx,
                  void* stream) {
    return trsv_dev<double>(uplo, trans, diag, n, a, lda, x, incx, stream);
}
int mbh_trsv_status(void* stream, int* failed) {
    if (!failed) {
        set_error("mbh_trsv_status: null output pointer");
        return MBH_E_ARG;
    }
    const hipError_t e = workspace_trsv_status((hipStream_t)stream, failed);
    return e == hipSuccess ? 0 : hip_fail(e, "mbh_trsv_status");
}

}  // extern "C"
