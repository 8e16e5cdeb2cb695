// Shared device-side definitions for the gfx950 level-3 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbh {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short i16x4 __attribute__((ext_vector_type(4)));
typedef short i16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));  // 16-byte chunk (builtin vector: usable in LDS)

// Pointer qualifier for LDS.  Tile kernels select between double buffers at
// run time; through plain (generic) pointers those accesses compile to flat
// instructions, which share the vector-memory counter with the global
// prefetch loads, so every fragment read would also wait for the next tile's
// HBM loads.  LDS-qualified pointers keep them ds_read / ds_write.
#define MBH_LDS __attribute__((address_space(3)))
// Pointer qualifier for global memory, for operand pointers the compiler
// cannot trace to a kernel argument (the batched kernels load them from a
// plan in HBM): keeps their accesses global_load / global_store instead of
// flat ones.
#define MBH_GLOBAL __attribute__((address_space(1)))

enum Prec { PREC_FP32 = 0, PREC_BF16X3 = 1, PREC_FP8 = 2, PREC_BF16 = 3 };

// Epilogue store masks (syrk/syr2k write one triangle of C; the trsm
// adjoint keeps tril of dB*X^T).
enum TriMask { TRI_NONE = 0, TRI_LOWER = 1, TRI_UPPER = 2 };

// Arguments of the register/LDS-staged MFMA gemm.  Operands may have their
// own storage types (TA, TB, TC) while the MFMA runs in TS: the trsm solver
// multiplies fp32 data by fp64 block inverses and forms fp64 residuals of
// fp32 operands this way.
template <typename TA, typename TB, typename TC, typename TS>
struct GemmArgs {
    int64_t m, n, k;
    const TA* A;
    int64_t lda;
    const TB* B;
    int64_t ldb;
    TC* C;
    int64_t ldc;
    TS alpha, beta;
    int tiles_m, tiles_n;
    int tri;        // TriMask applied to (row, col) of C
    int a_vec;      // 16-byte global loads of A allowed (pointer and lda aligned)
    int b_vec;      // same for B
    int tri_tiles;  // 0: full grid; 1/2: grid enumerates only lower/upper block-triangle tiles
    int krange;     // KRange
    // Partial-round split (direct-to-LDS tile): a launch with ksplit > 1
    // covers logical tiles tile0 .. tile0 + gridDim.x / ksplit - 1,
    // workgroup w taking chunk w % ksplit of tile tile0 + w / ksplit's
    // k-tiles and storing its raw sum as a 128 x 128 column-major block of
    // TS at part + w * 128 * 128; dma_splitk_reduce_kernel finishes C.
    int tile0, ksplit;
    void* part;
};
template <typename T>
using GemmParams = GemmArgs<T, T, T, T>;

// Maps the workgroup id to a C macro-tile.
//  * XCD-aware: the hardware deals workgroups round-robin to the 8 XCDs, so
//    workgroup w runs on XCD w%8.  Renumbering w -> (w%8)*(nwg/8) + w/8 gives
//    each XCD a contiguous run of tile ids, so the tiles sharing A row-panels
//    and B column-panels share one L2.
//  * Grouped order inside that run: GROUP tile-rows are walked column by
//    column, so concurrently resident tiles reuse the same A and B panels.
__device__ inline int xcd_remap(int w, int nwg) { return (nwg & 7) == 0 ? (w & 7) * (nwg >> 3) + (w >> 3) : w; }

// Grouped order of logical tile id w (GROUP tile-rows walked column by column).
__device__ inline void tile_coords_at(int w, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int GROUP = 8;
    int per_group = GROUP * tiles_n;
    int group = w / per_group;
    int first_m = group * GROUP;
    int gsize = min(tiles_m - first_m, GROUP);
    int in_group = w - group * per_group;
    tm = first_m + in_group % gsize;
    tn = in_group / gsize;
}

__device__ inline void tile_coords(int tiles_m, int tiles_n, int& tm, int& tn) {
    tile_coords_at(xcd_remap((int)blockIdx.x, (int)gridDim.x), tiles_m, tiles_n, tm, tn);
}

// Triangle tile enumeration for syrk-style launches over T tile-rows: tile
// id t -> (r, c) with c <= r over the lower block triangle, in groups of 8
// tile-rows walked column by column (the tiles before group g are rows
// 0 .. 8g-1, as in row-major order), so concurrently resident tiles share
// few row and column panels.  Inside a group of gs rows, the columns left of
// its diagonal block come first (gs tiles each), then the diagonal block
// column by column.  mode 1 returns (tm, tn) = (r, c) (lower triangle of C),
// mode 2 returns (c, r) (upper triangle).
__host__ __device__ inline void tri_tile_coords(int mode, int t, int T, int& tm, int& tn) {
    constexpr int G = 8;
    // row of t in row-major order: rows 0 .. R-1 hold R(R+1)/2 tiles
    int rr = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while ((rr + 1) * (rr + 2) / 2 <= t) ++rr;
    while (rr * (rr + 1) / 2 > t) --rr;
    const int g0 = (rr / G) * G;                // first row of t's group
    const int gs = T - g0 < G ? T - g0 : G;     // rows in the group
    int u = t - g0 * (g0 + 1) / 2;              // index inside the group
    int r, c;
    if (u < g0 * gs) {  // columns 0 .. g0-1, gs tiles each
        c = u / gs;
        r = g0 + u % gs;
    } else {  // diagonal block, column-major: column cc has rows cc .. gs-1
        u -= g0 * gs;
        int cc = 0;
        while (u >= gs - cc) {
            u -= gs - cc;
            ++cc;
        }
        c = g0 + cc;
        r = g0 + cc + u;
    }
    if (mode == 2) {
        tm = c;
        tn = r;
    } else {
        tm = r;
        tn = c;
    }
}

// How the split-precision pack reads an operand, in the coordinates (r, k) of
// the packed operand (r: row of op(A) or column of op(B); k: the reduction
// index): as laid out, as a triangle (trmm: zeros across the diagonal, the
// diagonal 1 with PV_UNIT), or as a symmetric matrix of which one half is
// stored (symm: the other half read at (k, r)).
enum PackView {
    PV_PLAIN = 0,
    PV_TRI_LE = 1,  // op(X)(r, k) = 0 for k > r
    PV_TRI_GE = 2,  // op(X)(r, k) = 0 for k < r
    PV_SYM_LE = 3,  // (r, k) with k <= r stored; k > r read at (k, r)
    PV_SYM_GE = 4,  // (r, k) with k >= r stored; k < r read at (k, r)
    PV_UNIT = 8,    // with PV_TRI_*: op(X)(r, r) = 1
};

// k-range restriction for triangular operands (trmm): a tile
// only visits the k-tiles where op(A) or op(B) can be nonzero.
enum KRange {
    KR_ALL = 0,
    KR_LE_ROW = 1,  // op(A) lower: A(i,k) = 0 for k > i      -> k < (tm+1)*BM
    KR_GE_ROW = 2,  // op(A) upper: A(i,k) = 0 for k < i      -> k >= tm*BM
    KR_GE_COL = 3,  // op(B) lower: B(k,j) = 0 for k < j      -> k >= tn*BN
    KR_LE_COL = 4   // op(B) upper: B(k,j) = 0 for k > j      -> k < (tn+1)*BN
};

// Tile order of a k-range launch: a tile's work grows with its distance
// along the triangle (KR_LE_ROW: with tm, KR_GE_ROW: with tiles_m - tm, and
// likewise along tn), so the grid is walked heaviest tile first.  The
// hardware deals consecutive workgroups round-robin to the XCDs, which then
// share the heavy tiles evenly (the XCD-contiguous order of tile_coords
// would hand one XCD all the heavy tile rows).  Consecutive workgroups keep
// one tile row (column) for A (B) panel reuse.
__device__ inline void tile_coords_kr(int krange, int tiles_m, int tiles_n, int& tm, int& tn) {
    const int w = (int)blockIdx.x;
    switch (krange) {
        case KR_LE_ROW: tm = tiles_m - 1 - w / tiles_n; tn = w % tiles_n; return;
        case KR_GE_ROW: tm = w / tiles_n; tn = w % tiles_n; return;
        case KR_GE_COL: tn = w / tiles_m; tm = w % tiles_m; return;
        case KR_LE_COL: tn = tiles_n - 1 - w / tiles_m; tm = w % tiles_m; return;
        default: tile_coords(tiles_m, tiles_n, tm, tn); return;
    }
}

// Tile of logical id w of a launch that covers only part of the tile list
// (no k-range order): grouped order, or the block triangle (tri_tiles 1/2).
__device__ inline void tile_coords_logical(int tri_tiles, int w, int tiles_m, int tiles_n, int& tm, int& tn) {
    if (tri_tiles) tri_tile_coords(tri_tiles, w, tiles_m, tm, tn);
    else tile_coords_at(w, tiles_m, tiles_n, tm, tn);
}

__device__ inline void kt_range(int mode, int tm, int tn, int BM, int BN, int BK, int nk, int& kt0, int& kt1) {
    kt0 = 0;
    kt1 = nk;
    switch (mode) {
        case KR_LE_ROW: kt1 = min(nk, ((tm + 1) * BM) / BK); break;
        case KR_GE_ROW: kt0 = (tm * BM) / BK; break;
        case KR_GE_COL: kt0 = (tn * BN) / BK; break;
        case KR_LE_COL: kt1 = min(nk, ((tn + 1) * BN) / BK); break;
        default: break;
    }
    if (kt0 > nk) kt0 = nk;
}

__device__ inline bool tri_keep(int tri, int64_t row, int64_t col) {
    return tri == TRI_NONE || (tri == TRI_LOWER ? row >= col : row <= col);
}

}  // namespace mbh
