// Level-3 routines recast onto the MFMA gemm kernels.
//
//   trsm  recursive blocked solve: the triangular dimension is halved (at
//         multiples of NB = 128) until a leaf is one 128x128 diagonal block;
//         a leaf applies the fp64 inverse of that block with an fp64-MFMA
//         product, every off-diagonal update is one large gemm on the fp32
//         (or BF16x3) matrix cores; left solves of up to 4 blocks run as one
//         panel kernel.  For fp32 data with refinement asked for, the same
//         solve runs on fp64 copies of alpha*B and of A's triangle with fp64
//         products throughout, and X is rounded to fp32 once.
//   trmm  op(A) materialized as a full triangle (zeros, unit diagonal), one
//         gemm whose tiles visit only the k-tiles where op(A) is nonzero, into
//         scratch, then copied back (trmm is in place).
//   syrk  gemm(A, A^T) over the lower/upper block-triangle tile grid with a
//         triangle-masked store.  syr2k: two such gemms.
//   symm  A symmetrized into scratch, then one gemm.
// Semantics follow the reference's docstrings (BLAS): src/level3/trsm.mojo,
// trmm.mojo, syrk.mojo, syr2k.mojo, symm.mojo.
#include <climits>

#include "common.hpp"
#include "dispatch.hpp"
#include "level3.hpp"

namespace mbh {

namespace {

constexpr int NB = 128;  // diagonal block of the trsm leaves

inline dim3 grid_for_elems(int64_t count) {
    int64_t b = (count + 255) / 256;
    if (b > 16384) b = 16384;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

// Square-matrix copies in 64 x 64 tiles, one tile per workgroup of 256
// threads (grid ceil(n/64) x ceil(n/64)); thread (tx, ty) = (tid % 64,
// tid / 64) handles rows i0 + tx of columns j0 + ty + 4q, so every global
// access of a wave covers 64 consecutive rows of one column.
constexpr int CT = 64;

inline dim3 tile_grid(int64_t n) { return dim3((unsigned)((n + CT - 1) / CT), (unsigned)((n + CT - 1) / CT)); }

// out(i,j) = element (i,j) of the `upper`/lower triangle of A (zeros outside,
// 1 on the diagonal when unit), n x n.  Tiles outside the triangle are
// written without reading A.
template <typename Tin, typename Tout>
__global__ __launch_bounds__(256) void tri_copy_kernel(int64_t n, const Tin* A, int64_t lda, int upper, int unit,
                                                       Tout* out, int64_t ldo) {
    const int tx = threadIdx.x % CT, ty = threadIdx.x / CT;
    const int64_t i0 = (int64_t)blockIdx.x * CT, j0 = (int64_t)blockIdx.y * CT, i = i0 + tx;
    if (i >= n) return;
    const bool inside = upper ? i0 < j0 : i0 > j0, outside = upper ? i0 > j0 : i0 < j0;  // off-diagonal tiles
#pragma unroll 4
    for (int q = 0; q < CT / 4; ++q) {
        const int64_t j = j0 + ty + 4 * q;
        if (j >= n) break;
        Tout v = Tout(0);
        if (inside) v = (Tout)A[i + j * lda];
        else if (!outside) {
            if (i == j) v = unit ? Tout(1) : (Tout)A[i + j * lda];
            else if (upper ? i < j : i > j) v = (Tout)A[i + j * lda];
        }
        out[i + j * ldo] = v;
    }
}

// out = A with the unreferenced triangle mirrored from the stored one.  A
// mirrored tile is read as the transposed tile of A (coalesced) and turned
// through LDS; a diagonal tile is staged whole and each element taken from
// the stored side.
template <typename T>
__global__ __launch_bounds__(256) void sym_copy_kernel(int64_t n, const T* A, int64_t lda, int upper, T* out,
                                                       int64_t ldo) {
    __shared__ T t[CT][CT + 1];
    const int tx = threadIdx.x % CT, ty = threadIdx.x / CT;
    const int64_t i0 = (int64_t)blockIdx.x * CT, j0 = (int64_t)blockIdx.y * CT;
    const bool diag = i0 == j0, stored = upper ? i0 < j0 : i0 > j0;
    if (!diag && stored) {
        const int64_t i = i0 + tx;
        if (i >= n) return;
        for (int q = 0; q < CT / 4; ++q) {
            const int64_t j = j0 + ty + 4 * q;
            if (j >= n) break;
            out[i + j * ldo] = A[i + j * lda];
        }
        return;
    }
    // t[a][b] = A(r0 + b, c0 + a) with (r0, c0) = (j0, i0) for a mirrored
    // tile (so t[a][b] = out(i0 + a, j0 + b)), (i0, j0) on the diagonal.
    const int64_t r0 = diag ? i0 : j0, c0 = diag ? j0 : i0;
    for (int q = 0; q < CT / 4; ++q) {
        const int a = ty + 4 * q;
        if (r0 + tx < n && c0 + a < n) t[a][tx] = A[(r0 + tx) + (c0 + a) * lda];
    }
    __syncthreads();
    const int64_t i = i0 + tx;
    if (i >= n) return;
    for (int q = 0; q < CT / 4; ++q) {
        const int b = ty + 4 * q;
        const int64_t j = j0 + b;
        if (j >= n) break;
        T v;
        if (!diag) v = t[tx][b];
        else v = (upper ? i <= j : i >= j) ? t[b][tx] : t[tx][b];  // A(i, j) : A(j, i)
        out[i + j * ldo] = v;
    }
}

// out(i,j) = (Tout)(alpha * in(i,j)), m x n.
template <typename Tin, typename Tout>
__global__ void copy_kernel(int64_t m, int64_t n, const Tin* in, int64_t ldi, Tout* out, int64_t ldo,
                            double alpha) {
    const int64_t total = m * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % m, j = idx / m;
        const Tin x = in[i + j * ldi];
        out[i + j * ldo] = (alpha == 1.0) ? (Tout)x : (Tout)(alpha * (double)x);
    }
}

// out(i,j) += d(i,j)
template <typename T>
__global__ void add_kernel(int64_t m, int64_t n, const T* d, int64_t ldd, T* out, int64_t ldo) {
    const int64_t total = m * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % m, j = idx / m;
        out[i + j * ldo] += d[i + j * ldd];
    }
}

// Inverses of the NB x NB diagonal blocks of op(A) (op(A) is eff_lower /
// upper triangular, trans = op is A^T), computed in fp64.  Output: inv +
// b*NB*NB, column-major, ld NB, zero-padded for a short last block.
//
// The columns of a triangular inverse are independent, so a block is split
// over TI_GROUPS workgroups of TI_COLS inverse columns each (grid nblk x
// TI_GROUPS).  Work happens in "solve space": indices are reversed for an
// upper block (s(i) = nb-1-i), which turns it into a lower one, L'(i,k) =
// op(A)_bb(s(i), s(k)) and inv(op(A)_bb)(s(i), s(j)) = inv(L')(i, j).  A
// workgroup stages the lower triangle of L' (rows and columns from its first
// column on) in LDS with all its global loads in flight at once, then walks
// the rows in chunks of TI_CH: one parallel product of the chunk rows with
// the already solved rows, then a TI_CH x TI_CH substitution per column in
// registers, scaled by precomputed reciprocals of the diagonal.
constexpr int TI_COLS = 32, TI_GROUPS = NB / TI_COLS, TI_CH = 16, TI_NT = 256;
constexpr int TI_TRI = NB * (NB + 1) / 2;  // packed lower triangle of L'

__device__ inline int ti_row(int i) { return i * (i + 1) / 2; }

template <typename Tin>
__global__ __launch_bounds__(TI_NT) void tri_inv_kernel(int64_t m, const Tin* A, int64_t lda, int eff_lower,
                                                        int trans, int unit, double* inv) {
    __shared__ double tri[TI_TRI];             // tri[ti_row(i) + k] = L'(i, k), k <= i
    __shared__ double xs[NB][TI_COLS + 1];     // xs[i][c] = inv(L')(i, c0 + c)
    __shared__ double rdiag[NB];               // 1 / L'(i, i)
    const int b = blockIdx.x, g = blockIdx.y, tid = threadIdx.x;
    const int64_t r0 = (int64_t)b * NB;
    const int nb = (int)((m - r0 < NB) ? (m - r0) : NB);
    const int c0 = g * TI_COLS;
    auto sp = [&](int i) { return eff_lower ? i : nb - 1 - i; };

    // Stage L' rows/columns >= c0 (identity outside the nb x nb block).  The
    // flat index runs along whichever of (i, k) is contiguous in memory.
    constexpr int PER = NB * NB / TI_NT;
    double v[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int idx = tid + e * TI_NT, fast = idx % NB, slow = idx / NB;
        const int i = trans ? slow : fast, k = trans ? fast : slow;
        double x = (i == k) ? 1.0 : 0.0;
        if (k <= i && k >= c0 && i < nb && !(unit && i == k)) {
            const int64_t gi = r0 + sp(i), gk = r0 + sp(k);
            x = trans ? (double)A[gk + gi * lda] : (double)A[gi + gk * lda];
        }
        v[e] = x;
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int idx = tid + e * TI_NT, fast = idx % NB, slow = idx / NB;
        const int i = trans ? slow : fast, k = trans ? fast : slow;
        if (k <= i && k >= c0) tri[ti_row(i) + k] = v[e];
    }
    for (int e = tid; e < NB * (TI_COLS + 1); e += TI_NT) (&xs[0][0])[e] = 0.0;
    __syncthreads();
    if (tid < NB && tid >= c0) rdiag[tid] = 1.0 / tri[ti_row(tid) + tid];

    const int col = tid % TI_COLS, r = tid / TI_COLS;  // product: rows r and r + 8 of the chunk
    constexpr int RPT = TI_CH / (TI_NT / TI_COLS);       // rows per thread in the product
    for (int ch = c0 / TI_CH; ch < NB / TI_CH; ++ch) {
        const int i0 = ch * TI_CH;
        // rhs(i, col) = [i == c0 + col] - sum_{k in [c0, i0)} L'(i, k) xs[k][col]
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int i = i0 + r + q * (TI_NT / TI_COLS);
            const double* li = tri + ti_row(i);
            double acc = 0.0;
            for (int k = c0; k < i0; ++k) acc += li[k] * xs[k][col];
            xs[i][col] = ((i == c0 + col) ? 1.0 : 0.0) - acc;
        }
        __syncthreads();
        if (tid < TI_COLS) {
            double x[TI_CH];
#pragma unroll
            for (int a = 0; a < TI_CH; ++a) x[a] = xs[i0 + a][tid];
#pragma unroll
            for (int a = 0; a < TI_CH; ++a) {
                const double* la = tri + ti_row(i0 + a) + i0;
                double t = x[a];
#pragma unroll
                for (int q = 0; q < a; ++q) t -= la[q] * x[q];
                x[a] = t * rdiag[i0 + a];
            }
#pragma unroll
            for (int a = 0; a < TI_CH; ++a) xs[i0 + a][tid] = x[a];
        }
        __syncthreads();
    }

    // Solve-space column c0 + c goes to output column s(c0 + c), rows s(i);
    // output columns of this group's range that lie past nb are zero.
    double* out = inv + (size_t)b * NB * NB;
    for (int e = tid; e < NB * TI_COLS; e += TI_NT) {
        const int pi = e % NB, c = e / NB, j = c0 + c;
        if (j < nb) out[pi + (size_t)sp(j) * NB] = (pi < nb) ? xs[sp(pi)][c] : 0.0;
        if (j >= nb) out[pi + (size_t)j * NB] = 0.0;
    }
}

// Leaf of the triangular solve: X_b := inv_b * B_b (left, B_b = len rows x
// n) or X_b := B_b * inv_b (right, B_b = m rows x len columns), in place, in
// fp64 (inv_b: NB x NB column-major, ld NB, zero-padded past len).
//
// A workgroup (4 waves) owns LEAF_W columns (left) or rows (right) of B_b
// and reads all of them into LDS (as fp64) before it writes any, so in place
// is safe and n / 16 (m / 16) workgroups run at once.  Wave w produces the
// 32 x 16 (left) or 16 x 32 (right) block at offset 32w along the block,
// with v_mfma_f64_16x16x4 (operand fragments as in the tile gemm: the "row"
// operand indexes C's rows by lane % 16, the "column" operand C's columns;
// register r of lane l holds C(l % 16, 4r + l / 16) of a 16 x 16 block).
// Its 2 x 32 fragments of inv_b come straight from global memory (one
// NB x NB block, shared by all workgroups through L2), all issued up front.
constexpr int LEAF_W = 16;

template <typename T, bool LEFT>
__global__ __launch_bounds__(256) void trsm_leaf_kernel(int64_t len, int64_t other, const double* __restrict__ inv,
                                                        T* B, int64_t ldb) {
    __shared__ double bs[NB][LEAF_W + 1];  // bs[k][c]: left B_b(k, c0 + c); right B_b(r0 + c, k)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lg = lane >> 4;
    const int64_t c0 = (int64_t)blockIdx.x * LEAF_W;
    // inv fragments: left, "row" operand inv(32w + 16i + l16, 4kk + lg);
    // right, "column" operand inv(4kk + lg, 32w + 16i + l16).
    double af[2][NB / 4];
#pragma unroll
    for (int kk = 0; kk < NB / 4; ++kk)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int q = 32 * wave + 16 * i + l16, k = 4 * kk + lg;
            af[i][kk] = LEFT ? inv[q + (size_t)k * NB] : inv[k + (size_t)q * NB];
        }
    // B_b into LDS, every load issued before the first is used (indices
    // clamped into B, zero outside it).
    constexpr int PER = NB * LEAF_W / 256;
    T bv_[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = tid + q * 256;
        int k, c;
        if (LEFT) {  // consecutive threads: consecutive rows k of one column
            k = e % NB;
            c = e / NB;
        } else {     // consecutive threads: consecutive rows c of one column k
            c = e % LEAF_W;
            k = e / LEAF_W;
        }
        const int64_t kk = min((int64_t)k, len - 1), cc = min(c0 + c, other - 1);
        bv_[q] = LEFT ? B[kk + cc * ldb] : B[cc + kk * ldb];
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = tid + q * 256;
        const int k = LEFT ? e % NB : e / LEAF_W, c = LEFT ? e / NB : e % LEAF_W;
        bs[k][c] = k < len && c0 + c < other ? (double)bv_[q] : 0.0;
    }
    __syncthreads();
    f64x4 acc[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) acc[i] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < NB / 4; ++kk) {
        const double bv = bs[4 * kk + lg][l16];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (LEFT)  // C^T block: column operand B_b(k, c0 + l16), row operand inv
                acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, af[i][kk], acc[i], 0, 0, 0);
            else       // column operand inv(k, col), row operand B_b(r0 + l16, k)
                acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][kk], bv, acc[i], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // left: C(row = 32w + 16i + l16, col = c0 + 4r + lg); right: C(row = c0 + l16, col = 32w + 16i + 4r + lg)
            if (LEFT) {
                const int64_t row = 32 * wave + 16 * i + l16, col = c0 + 4 * r + lg;
                if (row < len && col < other) B[row + col * ldb] = (T)acc[i][r];
            } else {
                const int64_t row = c0 + l16, col = 32 * wave + 16 * i + 4 * r + lg;
                if (row < other && col < len) B[row + col * ldb] = (T)acc[i][r];
            }
        }
}

// Left-side panel of up to PANEL_BLK diagonal blocks (PANEL_BLK * NB rows)
// solved in one kernel: for each block j in solve order, X_j = inv_j * B_j,
// then B_i -= op(A)_ij * X_j for the panel's blocks i still to come, all in
// fp64 with the workgroup's strip of B (LEAF_W columns, every panel row)
// resident in LDS.  This replaces PANEL_BLK leaves and the PANEL_BLK - 1
// small trailing products between them (k = 128 and 256, the least
// efficient gemms of the recursion) by one launch.
//
// 8 waves; wave w owns rows 16w .. 16w + 15 of a block and the strip's 16
// columns (one 16x16 v_mfma_f64_16x16x4f64 block: register r of lane l is
// X(16w + l % 16, 4r + l / 16)).  Operand fragments (inv_j rows, or op(A)_ij
// rows) come from global memory: lane l holds row 16w + l % 16 at
// k = 4kk + l / 16.  The fragments of the next product are loaded while the
// current one runs.  forward: blocks 0, 1, ... (op(A) lower), else from the
// last block down.  Rows past len are zero in LDS; inv is zero-padded past
// a short last block and op(A) is only read inside the panel.
constexpr int PANEL_BLK = 4, PANEL_NT = 512;

template <typename T>
__global__ __launch_bounds__(PANEL_NT) void trsm_panel_kernel(int64_t len, int64_t other, const double* __restrict__ inv,
                                                            const T* __restrict__ A, int64_t lda, int trans, int forward,
                                                            T* B, int64_t ldb) {
    __shared__ double bs[PANEL_BLK * NB][LEAF_W + 1];  // bs[k][c] = B(k, c0 + c), then X
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int l16 = lane & 15, lg = lane >> 4;
    const int64_t c0 = (int64_t)blockIdx.x * LEAF_W;
    const int nbk = (int)((len + NB - 1) / NB);
    {  // the strip into LDS: every load issued before the first is used
        constexpr int PER = PANEL_BLK * NB * LEAF_W / PANEL_NT;
        T v[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid + q * PANEL_NT, k = e % (PANEL_BLK * NB), c = e / (PANEL_BLK * NB);
            const int64_t kk = min((int64_t)k, len - 1), cc = min(c0 + c, other - 1);  // consecutive threads: rows
            v[q] = B[kk + cc * ldb];
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid + q * PANEL_NT, k = e % (PANEL_BLK * NB), c = e / (PANEL_BLK * NB);
            bs[k][c] = k < len && c0 + c < other ? (double)v[q] : 0.0;
        }
    }
    // Product p of the panel's sequence: step j = p's block in solve order;
    // p is the inverse (i = j) or the update of block i by X_j.
    const int row = 16 * wave + l16;
    // Fragments of an update hold op(A)'s element (fp32 data: its bits in
    // the low word), every load unconditional with indices clamped into the
    // panel; converted and masked where the MFMA uses them, so no load waits
    // where it is issued.
    auto load = [&](double (&f)[NB / 4], int i, int j, int k0, int k1) {  // k-steps [k0, k1)
        if (i == j) {
            const double* iv = inv + (size_t)j * NB * NB;
#pragma unroll
            for (int kk = k0; kk < k1; ++kk) f[kk] = iv[row + (size_t)(4 * kk + lg) * NB];
        } else {
            const int64_t gi = min((int64_t)i * NB + row, len - 1);
#pragma unroll
            for (int kk = k0; kk < k1; ++kk) {
                const int64_t gk = min((int64_t)j * NB + 4 * kk + lg, len - 1);
                const T x = trans ? A[gk + gi * lda] : A[gi + gk * lda];
                if constexpr (sizeof(T) == 4)
                    f[kk] = __builtin_bit_cast(double, (unsigned long long)__builtin_bit_cast(unsigned, x));
                else
                    f[kk] = x;
            }
        }
    };
    auto blk = [&](int step) { return forward ? step : nbk - 1 - step; };
    // The ordered list of products: for step s, (j, j), then (i, j) for the
    // blocks i after j in solve order.
    auto next_prod = [&](int& s, int& t) {  // t = 0: inverse; t > 0: t-th update of step s
        if (s + t + 1 < nbk) {
            ++t;
        } else {
            ++s;
            t = 0;
        }
    };
    double fa[NB / 4], fb[NB / 4];
    int s = 0, t = 0;
    load(fa, blk(0), blk(0), 0, NB / 4);
    __syncthreads();
    // One product with fragments fu (loaded earlier); loads the next
    // product's into fn.  Returns false after the last product.
    auto product = [&](double (&fu)[NB / 4], double (&fn)[NB / 4]) {
        const int j = blk(s), i = blk(s + t);
        int s2 = s, t2 = t;
        next_prod(s2, t2);
        // The next product's fragments (after the last product, a harmless
        // reload, so that every product issues the same loads), issued a
        // quarter at a time between the MFMAs of this one so that the load
        // unit's queue drains while the matrix cores run.
        const int ni = s2 < nbk ? blk(s2 + t2) : j, nj = s2 < nbk ? blk(s2) : j;
        const int bj = j * NB;
        const bool rok = i * NB + row < len;
        auto mfma = [&](double b, double a, f64x4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, c, 0, 0, 0); };
        // two accumulators (even and odd k-steps): two independent chains
        f64x4 acc0 = f64x4{0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
        constexpr int Q = NB / 16;  // k-steps per quarter
#pragma unroll
        for (int h = 0; h < NB / 4; h += Q) {
            load(fn, ni, nj, h, h + Q);
            double bv[Q];
#pragma unroll
            for (int kk = 0; kk < Q; ++kk) bv[kk] = bs[bj + 4 * (h + kk) + lg][l16];
            if (t != 0) {  // op(A) element (converted in place), zero outside the panel
#pragma unroll
                for (int kk = h; kk < h + Q; ++kk) {
                    double x = fu[kk];
                    if constexpr (sizeof(T) == 4)
                        x = (double)__builtin_bit_cast(float, (unsigned)__builtin_bit_cast(unsigned long long, x));
                    fu[kk] = rok && bj + 4 * kk + lg < len ? x : 0.0;
                }
            }
#pragma unroll
            for (int kk = 0; kk < Q; kk += 2) {
                acc0 = mfma(bv[kk], fu[h + kk], acc0);
                acc1 = mfma(bv[kk + 1], fu[h + kk + 1], acc1);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        const f64x4 acc = acc0 + acc1;
        const int bi = i * NB + row;  // this lane's row of the result
        auto col = [&](int r) { return 4 * r + lg; };
        if (t == 0) {
            __syncthreads();  // every wave has read B_j
#pragma unroll
            for (int r = 0; r < 4; ++r) bs[bi][col(r)] = acc[r];
            __syncthreads();  // X_j complete
        } else {
#pragma unroll
            for (int r = 0; r < 4; ++r) bs[bi][col(r)] -= acc[r];
            if (s2 != s) __syncthreads();  // every update of step s done
        }
        s = s2;
        t = t2;
        return s < nbk;
    };
    while (product(fa, fb) && product(fb, fa)) {
    }
    for (int e = tid; e < PANEL_BLK * NB * LEAF_W; e += PANEL_NT) {
        const int k = e % (PANEL_BLK * NB), c = e / (PANEL_BLK * NB);
        if (k < len && c0 + c < other) B[k + (c0 + c) * ldb] = (T)bs[k][c];
    }
}

template <typename T>
hipError_t xgemm(bool at, bool bt, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B,
                 int64_t ldb, T beta, T* C, int64_t ldc, int prec, hipStream_t s, int tri = 0, int tri_grid = 0,
                 int krange = 0);
template <>
hipError_t xgemm<float>(bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                        const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int prec, hipStream_t s,
                        int tri, int tri_grid, int krange) {
    return sgemm(at, bt, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec, s, tri, tri_grid, krange);
}
template <>
hipError_t xgemm<double>(bool at, bool bt, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                         int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int, hipStream_t s,
                         int tri, int tri_grid, int krange) {
    return dgemm(at, bt, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, tri, tri_grid, krange);
}

#define MBH_TRY(x)                          \
    do {                                    \
        hipError_t e__ = (x);               \
        if (e__ != hipSuccess) return e__;  \
    } while (0)

// In-place solve of op(A) X = B (left) or X op(A) = B (right) for B already
// scaled by alpha; `inv` holds the fp64 inverses of op(A)'s diagonal blocks.
template <typename T>
struct TriSolver {
    bool left, eff_lower, trans;
    int64_t m, n;
    const T* A;
    int64_t lda;
    const double* inv;
    int prec;
    hipStream_t s;

    // Pointer to op(A)(i, k) for a gemm operand with trans flag `trans`.
    const T* opA(int64_t i, int64_t k) const { return trans ? A + k + i * lda : A + i + k * lda; }

    static int64_t split(int64_t len) {
        int64_t h = (len / 2 + NB - 1) / NB * NB;
        return h >= len ? len - NB : h;
    }

    hipError_t leaf(T* B, int64_t ldb, int64_t r0, int64_t len) const {
        const double* iv = inv + (size_t)(r0 / NB) * NB * NB;
        if ((left ? n : m) / LEAF_W >= INT_MAX) return hipErrorInvalidValue;
        if (left) {  // X_b = inv_b * B_b  (len x n)
            trsm_leaf_kernel<T, true><<<dim3((unsigned)((n + LEAF_W - 1) / LEAF_W)), 256, 0, s>>>(len, n, iv, B + r0,
                                                                                              ldb);
        } else {     // X_b = B_b * inv_b  (m x len)
            trsm_leaf_kernel<T, false><<<dim3((unsigned)((m + LEAF_W - 1) / LEAF_W)), 256, 0, s>>>(
                len, m, iv, B + r0 * ldb, ldb);
        }
        return hipGetLastError();
    }

    // Left solves of up to PANEL_BLK diagonal blocks run as one panel kernel.
    hipError_t panel(T* B, int64_t ldb, int64_t r0, int64_t len) const {
        const double* iv = inv + (size_t)(r0 / NB) * NB * NB;
        if (n / LEAF_W >= INT_MAX) return hipErrorInvalidValue;
        trsm_panel_kernel<T><<<dim3((unsigned)((n + LEAF_W - 1) / LEAF_W)), PANEL_NT, 0, s>>>(
            len, n, iv, opA(r0, r0), lda, trans ? 1 : 0, eff_lower ? 1 : 0, B + r0, ldb);
        return hipGetLastError();
    }

    hipError_t solve(T* B, int64_t ldb, int64_t r0, int64_t len) const {
        if (left && len > NB && len <= PANEL_BLK * NB) return panel(B, ldb, r0, len);
        if (len <= NB) return leaf(B, ldb, r0, len);
        const int64_t h = split(len), r1 = r0 + h, len2 = len - h;
        if (left) {
            if (eff_lower) {  // X1 first; B2 -= op(A)21 X1
                MBH_TRY(solve(B, ldb, r0, h));
                MBH_TRY(xgemm<T>(trans, false, len2, n, h, T(-1), opA(r1, r0), lda, B + r0, ldb, T(1), B + r1, ldb,
                                 prec, s));
                return solve(B, ldb, r1, len2);
            }
            MBH_TRY(solve(B, ldb, r1, len2));  // X2 first; B1 -= op(A)12 X2
            MBH_TRY(xgemm<T>(trans, false, h, n, len2, T(-1), opA(r0, r1), lda, B + r1, ldb, T(1), B + r0, ldb, prec,
                             s));
            return solve(B, ldb, r0, h);
        }
        if (!eff_lower) {  // X op(A), op(A) upper: X1 first; B2 -= X1 op(A)12
            MBH_TRY(solve(B, ldb, r0, h));
            MBH_TRY(xgemm<T>(false, trans, m, len2, h, T(-1), B + r0 * ldb, ldb, opA(r0, r1), lda, T(1),
                             B + r1 * ldb, ldb, prec, s));
            return solve(B, ldb, r1, len2);
        }
        MBH_TRY(solve(B, ldb, r1, len2));  // op(A) lower: X2 first; B1 -= X2 op(A)21
        MBH_TRY(xgemm<T>(false, trans, m, h, len2, T(-1), B + r1 * ldb, ldb, opA(r1, r0), lda, T(1), B + r0 * ldb,
                         ldb, prec, s));
        return solve(B, ldb, r0, h);
    }
};

template <typename T>
hipError_t launch_scale_tri(int64_t m, int64_t n, T beta, T* C, int64_t ldc, hipStream_t s, int tri) {
    return scale<T>(m, n, beta, C, ldc, s, tri);
}

}  // namespace

template <typename T>
hipError_t tri_block_inverses(int64_t na, const T* A, int64_t lda, bool eff_lower, bool trans, bool unit, double* inv,
                              hipStream_t s) {
    static_assert(TRI_NB == NB, "one diagonal block size");
    const int64_t nblk = (na + NB - 1) / NB;
    if (nblk <= 0) return hipSuccess;
    if (nblk > INT_MAX) return hipErrorInvalidValue;
    tri_inv_kernel<T><<<dim3((unsigned)nblk, TI_GROUPS), dim3(TI_NT), 0, s>>>(na, A, lda, eff_lower ? 1 : 0,
                                                                             trans ? 1 : 0, unit ? 1 : 0, inv);
    return hipGetLastError();
}
template hipError_t tri_block_inverses<float>(int64_t, const float*, int64_t, bool, bool, bool, double*, hipStream_t);
template hipError_t tri_block_inverses<double>(int64_t, const double*, int64_t, bool, bool, bool, double*,
                                               hipStream_t);

// ---------------------------------------------------------------------------
// trsm (trsm.mojo:20-275, BLAS semantics)

template <typename T>
hipError_t trsm(char side, char uplo, char transa, char diag, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                T* B, int64_t ldb, int prec, int refine, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (alpha == T(0)) return scale<T>(m, n, T(0), B, ldb, s);
    const bool left = (side == 'L' || side == 'l');
    const bool upper = (uplo == 'U' || uplo == 'u');
    const bool trans = !(transa == 'N' || transa == 'n');
    const bool unit = (diag == 'U' || diag == 'u');
    const bool eff_lower = (!upper) != trans;
    const int64_t na = left ? m : n;
    const int64_t nblk = (na + NB - 1) / NB;
    const bool do_ir = (sizeof(T) == 4) && refine > 0;

    // scratch: inverses | (refine) fp64 X | fp64 triangle
    size_t inv_b = (size_t)nblk * NB * NB * sizeof(double);
    size_t x_b = do_ir ? (size_t)m * n * sizeof(double) : 0;
    size_t tri_b = do_ir ? (size_t)na * na * sizeof(double) : 0;
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    void* ws = nullptr;
    MBH_TRY(workspace(s, 1, al(inv_b) + al(x_b) + al(tri_b), &ws));
    char* base = (char*)ws;
    double* inv = (double*)base;
    double* x64 = (double*)(base + al(inv_b));
    double* tri64 = (double*)(base + al(inv_b) + al(x_b));

    MBH_TRY(tri_block_inverses<T>(na, A, lda, eff_lower, trans, unit, inv, s));
    if (do_ir) {
        // fp32 data to fp32 accuracy: the whole solve in fp64 (fp64 copies
        // of alpha*B and of A's triangle, the fp64 inverses, fp64 panels and
        // fp64 products), X rounded to fp32 once.  The same products in fp32
        // plus an fp64 residual (m^2 n flop, as many as the fp64 solve's
        // products) and a correction solve cost more and round more often.
        copy_kernel<T, double><<<grid_for_elems(m * n), 256, 0, s>>>(m, n, B, ldb, x64, m, (double)alpha);
        tri_copy_kernel<T, double><<<tile_grid(na), 256, 0, s>>>(na, A, lda, upper ? 1 : 0, unit ? 1 : 0,
                                                                           tri64, na);
        MBH_TRY(hipGetLastError());
        TriSolver<double> t64{left, eff_lower, trans, m, n, tri64, na, inv, prec, s};
        MBH_TRY(t64.solve(x64, m, 0, na));
        copy_kernel<double, T><<<grid_for_elems(m * n), 256, 0, s>>>(m, n, x64, m, B, ldb, 1.0);
        return hipGetLastError();
    }
    if (alpha != T(1)) MBH_TRY(scale<T>(m, n, alpha, B, ldb, s));
    TriSolver<T> ts{left, eff_lower, trans, m, n, A, lda, inv, prec, s};
    return ts.solve(B, ldb, 0, na);
}

// ---------------------------------------------------------------------------
// trmm (trmm.mojo:20-305, BLAS semantics): B := alpha*op(A)*B or alpha*B*op(A)

template <typename T>
hipError_t trmm(char side, char uplo, char transa, char diag, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                T* B, int64_t ldb, int prec, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (alpha == T(0)) return scale<T>(m, n, T(0), B, ldb, s);
    const bool left = (side == 'L' || side == 'l');
    const bool upper = (uplo == 'U' || uplo == 'u');
    const bool trans = !(transa == 'N' || transa == 'n');
    const bool unit = (diag == 'U' || diag == 'u');
    const bool eff_lower = (!upper) != trans;
    const int64_t na = left ? m : n;
    if constexpr (sizeof(T) == 4) {
        // Split-precision modes: the pack reads op(A) as a triangle straight
        // from A and the product is written into B, which by then lives only
        // in its packed planes.  Views are in the packed operand's (r, k):
        // left, op(A)(r, k); right, op(A)(k, r).
        if (sgemm_has_views(prec)) {
            const int view = (eff_lower == left ? PV_TRI_LE : PV_TRI_GE) | (unit ? PV_UNIT : 0);
            if (left)
                return sgemm_view(trans, false, m, n, m, alpha, A, lda, view, B, ldb, PV_PLAIN, 0.0f, B, ldb, prec,
                                  s, eff_lower ? KR_LE_ROW : KR_GE_ROW);
            return sgemm_view(false, trans, m, n, n, alpha, B, ldb, PV_PLAIN, A, lda, view, 0.0f, B, ldb, prec, s,
                              eff_lower ? KR_GE_COL : KR_LE_COL);
        }
    }
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t tri_b = (size_t)na * na * sizeof(T), w_b = (size_t)m * n * sizeof(T);
    void* ws = nullptr;
    MBH_TRY(workspace(s, 1, al(tri_b) + al(w_b), &ws));
    T* tri = (T*)ws;
    T* w = (T*)((char*)ws + al(tri_b));
    tri_copy_kernel<T, T><<<tile_grid(na), 256, 0, s>>>(na, A, lda, upper ? 1 : 0, unit ? 1 : 0, tri, na);
    MBH_TRY(hipGetLastError());
    if (left)
        MBH_TRY(xgemm<T>(trans, false, m, n, m, alpha, tri, na, B, ldb, T(0), w, m, prec, s, 0, 0,
                         eff_lower ? KR_LE_ROW : KR_GE_ROW));
    else
        MBH_TRY(xgemm<T>(false, trans, m, n, n, alpha, B, ldb, tri, na, T(0), w, m, prec, s, 0, 0,
                         eff_lower ? KR_GE_COL : KR_LE_COL));
    copy_kernel<T, T><<<grid_for_elems(m * n), 256, 0, s>>>(m, n, w, m, B, ldb, 1.0);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// syrk (syrk.mojo:20-200) and syr2k (syr2k.mojo:20-190)

template <typename T>
hipError_t syrk(char uplo, char trans, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, T beta, T* C,
                int64_t ldc, int prec, hipStream_t s) {
    if (n <= 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return hipSuccess;
    const bool upper = (uplo == 'U' || uplo == 'u');
    const bool tr = !(trans == 'N' || trans == 'n');
    const int tri = upper ? TRI_UPPER : TRI_LOWER;
    // N: C = A A^T (A n x k); T: C = A^T A (A k x n)
    return xgemm<T>(tr, !tr, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, prec, s, tri, upper ? 2 : 1);
}

template <typename T>
hipError_t syr2k(char uplo, char trans, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B,
                 int64_t ldb, T beta, T* C, int64_t ldc, int prec, hipStream_t s) {
    if (n <= 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return hipSuccess;
    const bool upper = (uplo == 'U' || uplo == 'u');
    const bool tr = !(trans == 'N' || trans == 'n');
    const int tri = upper ? TRI_UPPER : TRI_LOWER, grid = upper ? 2 : 1;
    if (alpha == T(0) || k == 0) return scale<T>(n, n, beta, C, ldc, s, tri);
    if constexpr (sizeof(T) == 4)
        return sgemm_syr2k(tr, !tr, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec, s, tri, grid);
    MBH_TRY(xgemm<T>(tr, !tr, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec, s, tri, grid));
    return xgemm<T>(tr, !tr, n, n, k, alpha, B, ldb, A, lda, T(1), C, ldc, prec, s, tri, grid);
}

// ---------------------------------------------------------------------------
// symm (symm.mojo:20-241)

template <typename T>
hipError_t symm(char side, char uplo, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, const T* B,
                int64_t ldb, T beta, T* C, int64_t ldc, int prec, hipStream_t s) {
    if (m <= 0 || n <= 0 || (alpha == T(0) && beta == T(1))) return hipSuccess;
    if (alpha == T(0)) return scale<T>(m, n, beta, C, ldc, s);
    const bool left = (side == 'L' || side == 'l');
    const bool upper = (uplo == 'U' || uplo == 'u');
    const int64_t na = left ? m : n;
    if constexpr (sizeof(T) == 4) {
        // Split-precision modes: the pack reads A's stored half and mirrors
        // the other (left, (r, k) = A(r, k); right, (r, k) = A(k, r)).
        if (sgemm_has_views(prec)) {
            const int view = (upper != left) ? PV_SYM_LE : PV_SYM_GE;
            if (left)
                return sgemm_view(false, false, m, n, m, alpha, A, lda, view, B, ldb, PV_PLAIN, beta, C, ldc, prec, s);
            return sgemm_view(false, false, m, n, n, alpha, B, ldb, PV_PLAIN, A, lda, view, beta, C, ldc, prec, s);
        }
    }
    void* ws = nullptr;
    MBH_TRY(workspace(s, 1, (size_t)na * na * sizeof(T), &ws));
    T* full = (T*)ws;
    sym_copy_kernel<T><<<tile_grid(na), 256, 0, s>>>(na, A, lda, upper ? 1 : 0, full, na);
    MBH_TRY(hipGetLastError());
    if (left) return xgemm<T>(false, false, m, n, m, alpha, full, na, B, ldb, beta, C, ldc, prec, s);
    return xgemm<T>(false, false, m, n, n, alpha, B, ldb, full, na, beta, C, ldc, prec, s);
}

#define MBH_L3(T)                                                                                                 \
    template hipError_t trsm<T>(char, char, char, char, int64_t, int64_t, T, const T*, int64_t, T*, int64_t, int, \
                                int, hipStream_t);                                                                \
    template hipError_t trmm<T>(char, char, char, char, int64_t, int64_t, T, const T*, int64_t, T*, int64_t, int, \
                                hipStream_t);                                                                     \
    template hipError_t syrk<T>(char, char, int64_t, int64_t, T, const T*, int64_t, T, T*, int64_t, int,          \
                                hipStream_t);                                                                     \
    template hipError_t syr2k<T>(char, char, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, T, T*,    \
                                 int64_t, int, hipStream_t);                                                      \
    template hipError_t symm<T>(char, char, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, T, T*,     \
                                int64_t, int, hipStream_t);
MBH_L3(float)
MBH_L3(double)
#undef MBH_L3

}  // namespace mbh
