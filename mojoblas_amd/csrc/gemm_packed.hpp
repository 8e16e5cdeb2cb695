// Macro-tile gemm over packed low-precision panels (see pack.hpp), fp32
// accumulate.
//
//   NPROD = 3: BF16x3, acc += Bh*Ah + Bl*Ah + Bh*Al   (x = hi + lo per operand)
//   NPROD = 1: plain bf16 operands, acc += Bh*Ah (also the native 16-bit
//              gemm: bf16 or f16 operands, C in that format, CK / F16 below)
//   NPROD = 8: e4m3 operands with one power-of-two scale per 128x128 tile of
//              A and of B, applied by v_mfma_scale_f32_32x32x64_f8f6f4 (E8M0
//              scale operands), so no per-slab partial sums are needed
//
// A k-tile is 64 bytes per row in every mode: 32 bf16 or 64 e4m3 values.
//
// Each k-tile of a macro-tile panel is one contiguous BM x BK (BN x BK) bf16
// run per plane, so the global loads are 16-byte, fully coalesced and carry
// no bounds checks (the pack pass zero-pads).  LDS holds [rows][BK] per plane
// (64-byte rows for BK = 32) with the 16-byte chunk index XOR-swizzled by
// swz(row) = ((row >> 1) ^ (row >> 2)) & 3: the ds_read_b128 fragment reads
// (32 consecutive rows, one chunk) and the ds_write_b128 staging writes
// (4 rows x 4 chunks per 16 lanes) then hit distinct 16-byte bank slots
// whether the LDS services 8 or 16 lanes per cycle.
//
// MFMA v_mfma_f32_32x32x16_bf16 with the B fragment as the instruction's A
// operand: the accumulator holds C^T, so each register's 32 lanes are 32
// consecutive rows of one C column (coalesced column-major stores).
//
// Pipeline: two LDS buffers, one barrier per k-tile.  The grid kernel loads
// the next k-tile straight into the other buffer (global_load_lds); in the
// bf16 modes it also software-pipelines the fragment reads (the reads of the
// next 16-deep k-step are issued between the MFMAs of the current one,
// gemm_packed_tile_pipe).  FP8 keeps the unpipelined gemm_packed_tile_dma.
// The register-staged form (gemm_packed_tile) serves the persistent batched
// kernel.
#pragma once
#include "common.hpp"
#include "gemm_fp.hpp"


namespace mbh {

struct PackedGemmParams {
    int64_t m, n;
    int ktiles;
    const uint16_t* Ah;
    const uint16_t* Al;
    const uint16_t* Bh;
    const uint16_t* Bl;
    void* C;        // fp32, or the 16-bit format of the native 16-bit gemm
    int64_t ldc;
    float alpha, beta;
    int tiles_m, tiles_n;
    int tri;
    int tri_tiles;  // 0 full grid, 1/2 lower/upper block triangle
    int krange;     // KRange (k-tiles visited per tile)
    const int* sA;  // fp8: E8M0 scale per 128x128 tile, [rows/128][kblocks]
    const int* sB;
    int kblocks;
    // Native 16-bit gemm only: a_ld / b_ld != 0 means Ah / Bh is the caller's
    // operand itself, read in place instead of from a packed plane (K a
    // multiple of 64): k-contiguous (row r at Ah + r*a_ld), or, in the
    // kernels built for an MN-form operand, row-contiguous (k-row k at
    // Ah + k*a_ld, row count a multiple of 8).
    int64_t a_ld, b_ld;
    // Partial-round split (bf16 modes, fp32 C): a launch with ksplit > 1
    // covers logical tiles tile0 .. tile0 + gridDim.x / ksplit - 1, workgroup
    // w taking chunk w % ksplit of tile tile0 + w / ksplit's k-tiles and
    // storing its raw sum as a BM x BN column-major block at part + w*BM*BN
    // (splitk_reduce_kernel finishes C).  ksplit == 1: tiles tile0 + (XCD
    // order of) blockIdx.x, stored to C.
    int tile0, ksplit;
    float* part;
};

typedef int i32x8 __attribute__((ext_vector_type(8)));

__device__ inline int lds_swz(int row) { return ((row >> 1) ^ (row >> 2)) & 3; }

// Storage format of C: CK = 0 fp32, 1 f16, 2 bf16.  The 16-bit formats are
// read into fp32, and alpha*acc + beta*c is rounded once (nearest even).
template <int CK>
struct CElem {
    typedef float T;
    static __device__ float load(float x) { return x; }
    static __device__ float store(float v) { return v; }
};
template <>
struct CElem<1> {
    typedef uint16_t T;
    static __device__ float load(uint16_t x) { return (float)__builtin_bit_cast(_Float16, x); }
    static __device__ uint16_t store(float v) { return __builtin_bit_cast(uint16_t, (_Float16)v); }
};
template <>
struct CElem<2> {
    typedef uint16_t T;
    static __device__ float load(uint16_t x) { return __uint_as_float((uint32_t)x << 16); }
    static __device__ uint16_t store(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
};

// Epilogue of a wave's TM x TN block of 32x32 accumulators holding C^T
// (register r of lane l is C(row0 + 32i + l%32, col0 + 32j + acc_col(r, l))).
// Interior blocks without a triangle mask take a branch-free path.
template <int TM, int TN, int CK = 0>
__device__ inline void epilogue_32x32(const f32x16 (&acc)[TM][TN], void* Cgen, int64_t ldc, int64_t m, int64_t n,
                                      int64_t row0, int64_t col0, int lane, float alpha, float beta, int tri) {
    typedef CElem<CK> E;
    typedef typename E::T CT;
    MBH_GLOBAL CT* const C = (MBH_GLOBAL CT*)Cgen;
    const int lrow = lane & 31;
    // interior: the whole block inside C and, under a triangle mask, inside
    // the kept triangle (every row >= every column, or <=)
    const bool inside_tri = tri == TRI_NONE || (tri == TRI_LOWER ? row0 >= col0 + 32 * TN - 1
                                                                : row0 + 32 * TM - 1 <= col0);
    const bool interior = inside_tri && row0 + 32 * TM <= m && col0 + 32 * TN <= n;
    if (interior) {
        MBH_GLOBAL CT* cb = C + (row0 + lrow) + col0 * ldc;
        if (beta == 0.0f) {
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        cb[i * 32 + (int64_t)(j * 32 + acc_col<32>(r, lane)) * ldc] = E::store(alpha * acc[i][j][r]);
        } else {
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        MBH_GLOBAL CT* cp = cb + i * 32 + (int64_t)(j * 32 + acc_col<32>(r, lane)) * ldc;
                        *cp = E::store(alpha * acc[i][j][r] + beta * E::load(*cp));
                    }
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int64_t row = row0 + i * 32 + lrow;
        if (row >= m) continue;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t col = col0 + j * 32 + acc_col<32>(r, lane);
                if (col < n && tri_keep(tri, row, col)) {
                    MBH_GLOBAL CT* cp = C + row + col * ldc;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * E::load(*cp);
                    *cp = E::store(v);
                }
            }
        }
    }
}

// LDS (in 16-byte chunks) of one double-buffered tile: 64-byte rows, one or
// two planes per operand.
template <int NPROD, int BM, int BN, int CPR = 4>
struct PackedTileLds {
    static constexpr int NPL = (NPROD == 3) ? 2 : 1;
    static constexpr int A = BM * CPR, B = BN * CPR;  // per plane per buffer (CPR chunks per row)
    static constexpr int TOTAL = 2 * NPL * (A + B);
};

// One BM x BN tile (tm, tn) over packed panels: the panel of tile row tm
// starts at chunk tm*ktiles*BM*4 of Ah/Al (likewise B), `lds` holds
// PackedTileLds::TOTAL chunks.  Shared by the grid kernel and the persistent
// batched kernel.
template <int NPROD, int BM, int BN, int WAVES_M, int WAVES_N>
__device__ inline void gemm_packed_tile(const PackedGemmParams& p, int tm, int tn, uint4* lds) {
    constexpr bool FP8 = (NPROD == 8);
    constexpr int BK = FP8 ? 64 : 32;  // k per 64-byte LDS row
    constexpr int NT = 64 * WAVES_M * WAVES_N;
    constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
    constexpr int TM = WM / 32, TN = WN / 32;
    constexpr int NPL = (NPROD == 3) ? 2 : 1;  // planes per operand
    constexpr int A_CH = BM * 4;               // 16-byte chunks per A plane tile (64-byte rows)
    constexpr int B_CH = BN * 4;
    constexpr int A_LD = A_CH / NT, B_LD = B_CH / NT;  // chunks per thread per plane
    static_assert(A_LD * NT == A_CH && B_LD * NT == B_CH, "tile not divisible by the thread block");
    typedef PackedTileLds<NPROD, BM, BN> L;
    // LDS views: As[buf][plane] / Bs[buf][plane], each rows * 4 chunks.
    MBH_LDS u32x4* const l3 = (MBH_LDS u32x4*)lds;
    MBH_LDS u32x4* As[2][NPL];
    MBH_LDS u32x4* Bs[2][NPL];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl) {
            As[b][pl] = l3 + (b * NPL + pl) * L::A;
            Bs[b][pl] = l3 + 2 * NPL * L::A + (b * NPL + pl) * L::B;
        }

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;

    // Panel offsets in 16-byte chunks: a k-tile block is BM (BN) rows of 64
    // bytes in every mode.
    const size_t a_off = (size_t)tm * p.ktiles * A_CH + tid, b_off = (size_t)tn * p.ktiles * B_CH + tid;
    typedef const MBH_GLOBAL u32x4* gptr;
    gptr __restrict__ gAh = ((gptr)p.Ah) + a_off;
    gptr __restrict__ gBh = ((gptr)p.Bh) + b_off;
    gptr __restrict__ gAl = gAh;
    gptr __restrict__ gBl = gBh;
    if (NPL == 2) {
        gAl = ((gptr)p.Al) + a_off;
        gBl = ((gptr)p.Bl) + b_off;
    }
    // This thread stages chunk tid (and tid + NT when a plane tile has 2*NT
    // chunks) of every plane; s0/s1 are their swizzled LDS offsets (in uint4),
    // the same for every k-tile.  The chunks live in named registers: arrays
    // indexed inside the k-loop were demoted to memory by the compiler.
    static_assert(A_LD >= 1 && A_LD <= 2 && B_LD >= 1 && B_LD <= 2, "staging holds one or two chunks per plane");
    const int c0 = tid, c1 = tid + NT;
    const int s0 = (c0 >> 2) * 4 + ((c0 & 3) ^ lds_swz(c0 >> 2));
    const int s1 = (c1 >> 2) * 4 + ((c1 & 3) ^ lds_swz(c1 >> 2));

    u32x4 ah0, ah1, al0, al1, bh0, bh1, bl0, bl1;
#define MBH_GLOAD(kt)                                                    \
    do {                                                                 \
        const size_t oa = (size_t)(kt) * A_CH, ob = (size_t)(kt) * B_CH; \
        ah0 = gAh[oa];                                                   \
        if (A_LD == 2) ah1 = gAh[oa + NT];                               \
        bh0 = gBh[ob];                                                   \
        if (B_LD == 2) bh1 = gBh[ob + NT];                               \
        if (NPL == 2) {                                                  \
            al0 = gAl[oa];                                               \
            if (A_LD == 2) al1 = gAl[oa + NT];                           \
            bl0 = gBl[ob];                                               \
            if (B_LD == 2) bl1 = gBl[ob + NT];                           \
        }                                                                \
    } while (0)
#define MBH_LSTORE(buf)                                                  \
    do {                                                                 \
        As[buf][0][s0] = ah0;                                            \
        if (A_LD == 2) As[buf][0][s1] = ah1;                             \
        Bs[buf][0][s0] = bh0;                                            \
        if (B_LD == 2) Bs[buf][0][s1] = bh1;                             \
        if (NPL == 2) {                                                  \
            As[buf][NPL - 1][s0] = al0;                                  \
            if (A_LD == 2) As[buf][NPL - 1][s1] = al1;                   \
            Bs[buf][NPL - 1][s0] = bl0;                                  \
            if (B_LD == 2) Bs[buf][NPL - 1][s1] = bl1;                   \
        }                                                                \
    } while (0)

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    int kt0, kt1;
    kt_range(p.krange, tm, tn, BM, BN, BK, p.ktiles, kt0, kt1);
    if (kt0 < kt1) {
        MBH_GLOAD(kt0);
        MBH_LSTORE(0);
    }
    __syncthreads();

    // Per-lane fragment addressing: row offset lane&31, chunk 2*ks + lane/32,
    // swizzle depends only on lane bits 1..3 (row offsets are multiples of 32).
    const int lrow = lane & 31, lgrp = lane >> 5, lsw = lds_swz(lrow);
    for (int kt = kt0; kt < kt1; ++kt) {
        const int cur = (kt - kt0) & 1;
        if (kt + 1 < kt1) MBH_GLOAD(kt + 1);
        if constexpr (FP8) {
            // Lane's 32 k values: chunks 2*lgrp and 2*lgrp+1 of its row.
            const int c0 = (2 * lgrp) ^ lsw, c1 = (2 * lgrp + 1) ^ lsw;
            const int sa = ((const MBH_GLOBAL int*)p.sA)[((tm * BM + wm0) >> 7) * p.kblocks + (kt >> 1)];
            const int sb = ((const MBH_GLOBAL int*)p.sB)[((tn * BN + wn0) >> 7) * p.kblocks + (kt >> 1)];
            i32x8 af[TM], bfr[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int base = (wm0 + i * 32 + lrow) * 4;
                const u32x4 x0 = As[cur][0][base + c0], x1 = As[cur][0][base + c1];
                af[i] = i32x8{(int)x0.x, (int)x0.y, (int)x0.z, (int)x0.w, (int)x1.x, (int)x1.y, (int)x1.z, (int)x1.w};
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const int base = (wn0 + j * 32 + lrow) * 4;
                const u32x4 x0 = Bs[cur][0][base + c0], x1 = Bs[cur][0][base + c1];
                bfr[j] = i32x8{(int)x0.x, (int)x0.y, (int)x0.z, (int)x0.w, (int)x1.x, (int)x1.y, (int)x1.z, (int)x1.w};
            }
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(bfr[j], af[i], acc[i][j], 0, 0, 0, sb,
                                                                                 0, sa);
        } else {
#pragma unroll
        for (int ks = 0; ks < BK / 16; ++ks) {
            const int ch = (2 * ks + lgrp) ^ lsw;
            bf16x8 ah[TM], bh[TN], al[TM], bl[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int idx = (wm0 + i * 32 + lrow) * 4 + ch;
                ah[i] = __builtin_bit_cast(bf16x8, (u32x4)As[cur][0][idx]);
                if (NPROD == 3) al[i] = __builtin_bit_cast(bf16x8, (u32x4)As[cur][NPL - 1][idx]);
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const int idx = (wn0 + j * 32 + lrow) * 4 + ch;
                bh[j] = __builtin_bit_cast(bf16x8, (u32x4)Bs[cur][0][idx]);
                if (NPROD == 3) bl[j] = __builtin_bit_cast(bf16x8, (u32x4)Bs[cur][NPL - 1][idx]);
            }
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    if (NPROD == 3) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
                    }
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
                }
        }
        }
        if (kt + 1 < kt1) MBH_LSTORE(cur ^ 1);
        __syncthreads();
    }
#undef MBH_GLOAD
#undef MBH_LSTORE

    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    epilogue_32x32<TM, TN>(acc, p.C, p.ldc, p.m, p.n, m0 + wm0, n0 + wn0, lane, p.alpha, p.beta, p.tri);
}

// MFMAs of one k-tile held in LDS: a[pl] / b[pl] are the plane bases of the
// buffer (rows * 4 swizzled chunks each), the wave's block starts at
// (wm0, wn0).  sa / sb: E8M0 scales of the k-tile (FP8 only).
template <int NPROD, int TM, int TN>
__device__ inline void packed_ktile_mfma(const MBH_LDS u32x4* a0, const MBH_LDS u32x4* a1, const MBH_LDS u32x4* b0,
                                         const MBH_LDS u32x4* b1, int wm0, int wn0, int lane, int sa, int sb,
                                         f32x16 (&acc)[TM][TN]) {
    const int lrow = lane & 31, lgrp = lane >> 5, lsw = lds_swz(lrow);
    if constexpr (NPROD == 8) {
        // Lane's 32 k values: chunks 2*lgrp and 2*lgrp+1 of its row.
        const int c0 = (2 * lgrp) ^ lsw, c1 = (2 * lgrp + 1) ^ lsw;
        i32x8 af[TM], bfr[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            const int base = (wm0 + i * 32 + lrow) * 4;
            const u32x4 x0 = a0[base + c0], x1 = a0[base + c1];
            af[i] = i32x8{(int)x0.x, (int)x0.y, (int)x0.z, (int)x0.w, (int)x1.x, (int)x1.y, (int)x1.z, (int)x1.w};
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int base = (wn0 + j * 32 + lrow) * 4;
            const u32x4 x0 = b0[base + c0], x1 = b0[base + c1];
            bfr[j] = i32x8{(int)x0.x, (int)x0.y, (int)x0.z, (int)x0.w, (int)x1.x, (int)x1.y, (int)x1.z, (int)x1.w};
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(bfr[j], af[i], acc[i][j], 0, 0, 0, sb, 0, sa);
    } else {
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int ch = (2 * ks + lgrp) ^ lsw;
            bf16x8 ah[TM], bh[TN], al[TM], bl[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int idx = (wm0 + i * 32 + lrow) * 4 + ch;
                ah[i] = __builtin_bit_cast(bf16x8, (u32x4)a0[idx]);
                if (NPROD == 3) al[i] = __builtin_bit_cast(bf16x8, (u32x4)a1[idx]);
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const int idx = (wn0 + j * 32 + lrow) * 4 + ch;
                bh[j] = __builtin_bit_cast(bf16x8, (u32x4)b0[idx]);
                if (NPROD == 3) bl[j] = __builtin_bit_cast(bf16x8, (u32x4)b1[idx]);
            }
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    if (NPROD == 3) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
                    }
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
                }
        }
    }
}

// The same tile with the global -> LDS copy done by the load unit itself
// (global_load_lds_dwordx4: lane i of a wave writes 16 bytes at LDS base +
// 16*i), so the staged k-tile never passes through registers and no
// ds_write is issued.  Each lane loads the global chunk that belongs in its
// swizzled LDS slot: a wave-load covers 16 rows x 4 chunks, and slot
// (row, c) holds global chunk (row, c ^ swz(row)).  The two LDS buffers are
// separate arrays (lds0 / lds1) and the k-loop is unrolled by two, so every
// fragment read names a buffer that no load in flight writes; the wait for
// the loads (vmcnt) comes once per k-tile, before the barrier.
template <int NPROD, int BM, int BN, int WAVES_M, int WAVES_N>
__device__ inline void gemm_packed_tile_dma(const PackedGemmParams& p, int tm, int tn, u32x4* lds0, u32x4* lds1) {
    constexpr bool FP8 = (NPROD == 8);
    constexpr int NW = WAVES_M * WAVES_N;
    constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
    constexpr int TM = WM / 32, TN = WN / 32;
    constexpr int NPL = (NPROD == 3) ? 2 : 1;
    constexpr int A_CH = BM * 4, B_CH = BN * 4;           // chunks per plane per k-tile
    constexpr int A_WL = A_CH / (64 * NW), B_WL = B_CH / (64 * NW);  // wave-loads per plane
    static_assert(A_WL * 64 * NW == A_CH && B_WL * 64 * NW == B_CH && A_WL >= 1 && B_WL >= 1,
                  "tile not divisible into wave-loads");
    typedef const MBH_GLOBAL u32x4* gptr;

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
    // Chunk this lane fetches inside its wave-load's 16 x 4 block (row bits
    // above 3 do not change the swizzle).
    const int lrow16 = lane >> 2;
    const int lchunk = lrow16 * 4 + ((lane & 3) ^ lds_swz(lrow16));
    const size_t a_off = (size_t)tm * p.ktiles * A_CH + lchunk, b_off = (size_t)tn * p.ktiles * B_CH + lchunk;
    gptr gA0 = ((gptr)p.Ah) + a_off;
    gptr gB0 = ((gptr)p.Bh) + b_off;
    gptr gA1 = gA0;
    gptr gB1 = gB0;
    if (NPL == 2) {
        gA1 = ((gptr)p.Al) + a_off;
        gB1 = ((gptr)p.Bl) + b_off;
    }
    MBH_LDS u32x4* const L0 = (MBH_LDS u32x4*)lds0;
    MBH_LDS u32x4* const L1 = (MBH_LDS u32x4*)lds1;
    // Buffer layout: A plane 0, [A plane 1], B plane 0, [B plane 1].
    constexpr int OA1 = A_CH, OB0 = NPL * A_CH, OB1 = NPL * A_CH + B_CH;

#define MBH_DMA(kt, LB)                                                                                 \
    do {                                                                                                \
        const size_t ga = (size_t)(kt) * A_CH, gb = (size_t)(kt) * B_CH;                                \
        _Pragma("unroll") for (int v = 0; v < A_WL; ++v) {                                              \
            const int blk = (wave + v * NW) * 64;                                                       \
            __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(gA0 + ga + blk),                  \
                                             (MBH_LDS void*)((LB) + blk), 16, 0, 0);                    \
            if (NPL == 2)                                                                               \
                __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(gA1 + ga + blk),              \
                                                 (MBH_LDS void*)((LB) + OA1 + blk), 16, 0, 0);          \
        }                                                                                               \
        _Pragma("unroll") for (int v = 0; v < B_WL; ++v) {                                              \
            const int blk = (wave + v * NW) * 64;                                                       \
            __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(gB0 + gb + blk),                  \
                                             (MBH_LDS void*)((LB) + OB0 + blk), 16, 0, 0);              \
            if (NPL == 2)                                                                               \
                __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(gB1 + gb + blk),              \
                                                 (MBH_LDS void*)((LB) + OB1 + blk), 16, 0, 0);          \
        }                                                                                               \
        __builtin_amdgcn_sched_barrier(0); /* keep the loads ahead of the k-tile's MFMAs */             \
    } while (0)

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    constexpr int BK = FP8 ? 64 : 32;
    int kt0, kt1;
    kt_range(p.krange, tm, tn, BM, BN, BK, p.ktiles, kt0, kt1);
#define MBH_SCALES(kt)                                                                                          \
    if (FP8) {                                                                                                  \
        sa = ((const MBH_GLOBAL int*)p.sA)[((tm * BM + wm0) >> 7) * p.kblocks + ((kt) >> 1)];                    \
        sb = ((const MBH_GLOBAL int*)p.sB)[((tn * BN + wn0) >> 7) * p.kblocks + ((kt) >> 1)];                    \
    }
    int sa = 0, sb = 0;
    if (kt0 < kt1) MBH_DMA(kt0, L0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int kt = kt0;
    for (; kt + 1 < kt1; kt += 2) {
        MBH_DMA(kt + 1, L1);
        MBH_SCALES(kt);
        packed_ktile_mfma<NPROD, TM, TN>(L0, L0 + OA1, L0 + OB0, L0 + OB1, wm0, wn0, lane, sa, sb, acc);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (kt + 2 < kt1) MBH_DMA(kt + 2, L0);
        MBH_SCALES(kt + 1);
        packed_ktile_mfma<NPROD, TM, TN>(L1, L1 + OA1, L1 + OB0, L1 + OB1, wm0, wn0, lane, sa, sb, acc);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    if (kt < kt1) {  // odd count: the last k-tile is in buffer 0
        MBH_SCALES(kt);
        packed_ktile_mfma<NPROD, TM, TN>(L0, L0 + OA1, L0 + OB0, L0 + OB1, wm0, wn0, lane, sa, sb, acc);
    }
#undef MBH_SCALES
#undef MBH_DMA

    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    epilogue_32x32<TM, TN>(acc, p.C, p.ldc, p.m, p.n, m0 + wm0, n0 + wn0, lane, p.alpha, p.beta, p.tri);
}

// Fragments of one 16-deep k-step of a wave's TM x TN block of 32x32
// outputs: hi (and, for BF16x3, lo) planes of A and B.
template <int NPL, int TM, int TN>
struct PackedFrag {
    bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
};

// LDS -> registers: chunk `ch` (swizzled) of the wave's rows in each plane
// (CPR chunks per LDS row).  RA / RB false: that operand is read elsewhere.
template <int NPL, int TM, int TN, int CPR = 4, bool RA = true, bool RB = true>
__device__ inline void packed_frag_read(PackedFrag<NPL, TM, TN>& f, const MBH_LDS u32x4* buf, int oa1, int ob0,
                                        int ob1, int wm0, int wn0, int lrow, int ch) {
#pragma unroll
    for (int i = 0; i < TM && RA; ++i) {
        const int idx = (wm0 + i * 32 + lrow) * CPR + ch;
        f.ah[i] = __builtin_bit_cast(bf16x8, (u32x4)buf[idx]);
        if (NPL == 2) f.al[i] = __builtin_bit_cast(bf16x8, (u32x4)buf[oa1 + idx]);
    }
#pragma unroll
    for (int j = 0; j < TN && RB; ++j) {
        const int idx = (wn0 + j * 32 + lrow) * CPR + ch;
        f.bh[j] = __builtin_bit_cast(bf16x8, (u32x4)buf[ob0 + idx]);
        if (NPL == 2) f.bl[j] = __builtin_bit_cast(bf16x8, (u32x4)buf[ob1 + idx]);
    }
}

// MN form of a 16-bit operand's 64-deep k-tile in LDS, for an operand whose
// rows (m of op(A), n of op(B)) are the contiguous index in HBM (A not
// transposed, B transposed), so that it can be read in place: LDS holds
// [k][BR rows] (2*BR bytes per k-row).  A wave-load writes 1 KiB, RPW whole
// k-rows; wave-load g starts at byte g*GS (GS = 1 KiB + 64).  The k-rows are
// permuted over the wave-loads: k = 4q + r sits in wave-load
// g = 4*(q / RPW) + r as its (q % RPW)-th k-row, so the four consecutive k
// of one transposing read lie in wave-loads g = 0..3 mod 4, at bank offsets
// 64 bytes apart.
template <int BR>
struct MnPanel {
    static constexpr int RPW = 512 / BR;  // k-rows per wave-load
    static constexpr int LPK = BR / 8;    // lanes (16-byte chunks) per k-row
    static constexpr int NG = 64 / RPW;   // wave-loads per k-tile
    static constexpr int GS = 1088;       // bytes per wave-load slot
    static constexpr int CH = NG * GS / 16;  // 16-byte chunks per buffer
    static_assert(BR == 64 || BR == 128 || BR == 256, "row block");
    // k-row held by k-row slot `sub` of wave-load g.
    __device__ static int k_of(int g, int sub) { return 4 * ((g >> 2) * RPW + sub) + (g & 3); }
    // Byte offset, for one lane, of the transposing read that hands it rows
    // w0 + (lane & 31) at k = 16*s + 8*(lane >> 5) + 4*h + 0..3: in each
    // group of 16 lanes, ds_read_b64_tr_b16 gives lane 4a + j element j of
    // the 8-byte reads of lanes a, a + 4, a + 8, a + 12, so lane t of group
    // G reads rows 16*(G & 1) + 4*(t & 3) .. + 3 at k = 8*(G >> 1) + (t >> 2)
    // (plus 16 s + 4 h).  This is the lane part; step and half add
    // step_off(s) + h * 2 * BR.
    __device__ static int lane_off(int lane, int w0) {
        const int G = lane >> 4, t = lane & 15, u = 2 * (G >> 1);
        return (4 * (u / RPW) + (t >> 2)) * GS + (u % RPW) * 2 * BR + (w0 + 16 * (G & 1) + 4 * (t & 3)) * 2;
    }
    static constexpr int step_off(int s) { return 4 * ((4 * s) / RPW) * GS + ((4 * s) % RPW) * 2 * BR; }
};

// Fragments of k-step S of T 32-row blocks from an MN-form panel at byte
// `base` (lane part `loff` from MnPanel::lane_off): two transposing reads
// per fragment (k 0..3 and 4..7 of the lane's half of the step).
template <int BR, int T, int S>
__device__ inline void mn_frag_read(bf16x8 (&x)[T], const MBH_LDS u32x4* buf, int base, int loff) {
    typedef short v4i16 __attribute__((ext_vector_type(4)));
    typedef MnPanel<BR> P;
    const MBH_LDS char* b = (const MBH_LDS char*)buf + base + loff + P::step_off(S);
#pragma unroll
    for (int i = 0; i < T; ++i) {
        const v4i16 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((MBH_LDS v4i16*)(b + i * 64));
        const v4i16 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((MBH_LDS v4i16*)(b + i * 64 + 2 * BR));
        x[i] = __builtin_bit_cast(bf16x8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
    }
}

// MFMAs of one k-step.  Per accumulator the products come in the order
// lo*hi, hi*lo, hi*hi (as in packed_ktile_mfma); between two products on the
// same accumulator the other TM*TN - 1 accumulators are issued.
template <int NPROD, int NPL, int TM, int TN, bool F16 = false>
__device__ inline void packed_frag_mfma(const PackedFrag<NPL, TM, TN>& f, f32x16 (&acc)[TM][TN]) {
    if constexpr (F16) {  // native f16 operands (one plane): v_mfma_f32_32x32x16_f16
        typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, f.bh[j]),
                                                                   __builtin_bit_cast(f16x8, f.ah[i]), acc[i][j], 0, 0, 0);
        return;
    }
    if (NPROD == 3) {
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.bl[j], f.ah[i], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.bh[j], f.al[i], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.bh[j], f.ah[i], acc[i][j], 0, 0, 0);
}

// BF16x3 / bf16 tile with the fragment reads software-pipelined: two register
// sets hold consecutive k-steps, and the reads of step s+1 are issued between
// the MFMAs of step s, so a wave never waits for LDS with an empty MFMA queue
// (without this the reads of a step were issued after the previous step's
// MFMAs and their full latency fell between the two).  The k-tile loads are
// the direct-to-LDS ones of gemm_packed_tile_dma; one barrier per k-tile,
// after which the next k-tile's first step is read while the current tile's
// second step (already in registers) is multiplied.
//
// CK: storage format of C (CElem); F16: the single operand plane holds f16
// (native f16 gemm), multiplied on the f16 MFMA.  STRIDED: each lane keeps
// one pointer per wave-load and advances it by a per-operand k-tile stride,
// so an operand is either a packed plane (stride BM*BKT/8 chunks) or, when
// p.a_ld / p.b_ld is set, the caller's k-contiguous matrix read in place
// (lane chunk (row, c) at X + row*ld + 8c, stride BKT/8 chunks; rows past the
// operand's last are clamped to it: they only reach C entries that are never
// stored).  AMN / BMN: that operand is the caller's row-contiguous matrix,
// loaded whole k-rows at a time into an MnPanel (lane chunk at
// X + k*ld + row0 + 8c, stride BKT*ld/8 chunks, chunks past the last row
// clamped to the last 8 rows) and read with mn_frag_read.  BKT: k per LDS row, 32 (64-byte rows, two 16-deep MFMA steps per
// k-tile) or, in the 16-bit kernel, 64 (128-byte rows, so an in-place row
// read is one whole 128-byte line, and four steps per barrier); 64-deep rows
// are swizzled over 8 chunks by swz8(row) = (row >> 1) & 7, which spreads
// the 32-row fragment reads (even and odd rows on separate halves of the
// banks) over distinct 16-byte bank slots.
template <int NPROD, int BM, int BN, int WAVES_M, int WAVES_N, int CK = 0, bool F16 = false, bool STRIDED = false,
          int BKT = 32, bool AMN = false, bool BMN = false>
__device__ inline void gemm_packed_tile_pipe(const PackedGemmParams& p, int tm, int tn, u32x4* lds0, u32x4* lds1,
                                             int chunk = 0) {
    static_assert(NPROD == 1 || NPROD == 3, "bf16 modes only");
    static_assert(!F16 || NPROD == 1, "f16 operands use one plane");
    static_assert(!STRIDED || NPROD == 1, "in-place operands hold one plane");
    static_assert(BKT == 32 || (BKT == 64 && STRIDED), "64-deep k-tiles: 16-bit kernel only");
    constexpr int NW = WAVES_M * WAVES_N;
    static_assert(BKT == 32 || NW % 2 == 0, "swz8 of a wave-load's rows assumes an even wave count");
    static_assert(!(AMN || BMN) || (STRIDED && BKT == 64), "MN-form panels: 16-bit kernel only");
    constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
    constexpr int TM = WM / 32, TN = WN / 32;
    constexpr int NPL = (NPROD == 3) ? 2 : 1;
    constexpr int CPR = BKT / 8;  // 16-byte chunks per LDS row
    constexpr int KS = BKT / 16;  // 16-deep MFMA k-steps per k-tile
    constexpr int RW = 64 / CPR;  // rows per wave-load
    constexpr int A_CH = BM * CPR, B_CH = BN * CPR;
    constexpr int A_WL = A_CH / (64 * NW), B_WL = B_CH / (64 * NW);
    static_assert(A_WL * 64 * NW == A_CH && B_WL * 64 * NW == B_CH && A_WL >= 1 && B_WL >= 1,
                  "tile not divisible into wave-loads");
    // An MN-form operand takes two (transposing) reads per fragment.
    constexpr int NREAD = NPL * ((AMN ? 2 : 1) * TM + (BMN ? 2 : 1) * TN), NMFMA = NPROD * TM * TN;
    typedef const MBH_GLOBAL u32x4* gptr;

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
    auto swz = [](int row) { return BKT == 32 ? lds_swz(row) : (row >> 1) & 7; };
    // Row of this lane inside its wave-load (RW x CPR chunks) and the
    // swizzle of that row: for BKT = 32 the wave-load's base row (a multiple
    // of 16) leaves the swizzle bits alone; for 64 its bit 3 is the parity
    // of the wave-load index, which is the wave's parity (NW even).
    const int lroww = lane / CPR;
    const int lslot = (lane % CPR) ^ swz(BKT == 32 ? lroww : (wave & 1) * 8 + lroww);
    const int lchunk = lroww * CPR + lslot;
    const size_t a_off = (size_t)tm * p.ktiles * A_CH + lchunk, b_off = (size_t)tn * p.ktiles * B_CH + lchunk;
    gptr gA0 = ((gptr)p.Ah) + a_off;
    gptr gB0 = ((gptr)p.Bh) + b_off;
    gptr gA1 = gA0;
    gptr gB1 = gB0;
    if (NPL == 2) {
        gA1 = ((gptr)p.Al) + a_off;
        gB1 = ((gptr)p.Bl) + b_off;
    }
    MBH_LDS u32x4* const L0 = (MBH_LDS u32x4*)lds0;
    MBH_LDS u32x4* const L1 = (MBH_LDS u32x4*)lds1;
    // MN-form operands: same wave-load count (BR/8 per k-tile), wave-load
    // slots of MnPanel::GS bytes.
    typedef MnPanel<AMN ? BM : 256> PA;
    typedef MnPanel<BMN ? BN : 256> PB;
    static_assert(!AMN || PA::NG == A_WL * NW, "MN panel wave-loads (A)");
    static_assert(!BMN || PB::NG == B_WL * NW, "MN panel wave-loads (B)");
    constexpr int AWS = AMN ? PA::GS / 16 : 64, BWS = BMN ? PB::GS / 16 : 64;  // chunks per wave-load slot
    constexpr int OA1 = A_CH, OB0 = AMN ? PA::CH : NPL * A_CH, OB1 = NPL * A_CH + B_CH;
    // STRIDED: per wave-load pointers and k-tile strides (in 16-byte chunks).
    gptr pa[A_WL], pb[B_WL];
    size_t sa = A_CH, sb = B_CH;
    if constexpr (STRIDED) {
        const int c = lslot;
#pragma unroll
        for (int v = 0; v < A_WL; ++v) {
            const int blk = (wave + v * NW) * 64;
            if constexpr (AMN) {  // k-row k_of(g, sub), rows from 8*(lane % LPK); chunks past m clamped
                const int g = wave + v * NW, sub = lane / PA::LPK;
                int64_t row = (int64_t)tm * BM + 8 * (lane % PA::LPK);
                row = row < p.m ? row : p.m - 8;
                pa[v] = (gptr)(p.Ah + (int64_t)PA::k_of(g, sub) * p.a_ld + row);
            } else if (p.a_ld) {
                int64_t row = (int64_t)tm * BM + (wave + v * NW) * RW + lroww;
                row = row < p.m ? row : p.m - 1;
                pa[v] = (gptr)(p.Ah + row * p.a_ld) + c;
            } else {
                pa[v] = gA0 + blk;
            }
        }
#pragma unroll
        for (int v = 0; v < B_WL; ++v) {
            const int blk = (wave + v * NW) * 64;
            if constexpr (BMN) {
                const int g = wave + v * NW, sub = lane / PB::LPK;
                int64_t row = (int64_t)tn * BN + 8 * (lane % PB::LPK);
                row = row < p.n ? row : p.n - 8;
                pb[v] = (gptr)(p.Bh + (int64_t)PB::k_of(g, sub) * p.b_ld + row);
            } else if (p.b_ld) {
                int64_t row = (int64_t)tn * BN + (wave + v * NW) * RW + lroww;
                row = row < p.n ? row : p.n - 1;
                pb[v] = (gptr)(p.Bh + row * p.b_ld) + c;
            } else {
                pb[v] = gB0 + blk;
            }
        }
        if (p.a_ld) sa = AMN ? (size_t)p.a_ld * BKT / 8 : CPR;
        if (p.b_ld) sb = BMN ? (size_t)p.b_ld * BKT / 8 : CPR;
    }

#define MBH_DMA(kt, LB)                                                                                 \
    do {                                                                                                \
        if constexpr (STRIDED) {                                                                        \
            _Pragma("unroll") for (int v = 0; v < A_WL; ++v)                                            \
                __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(pa[v] + (size_t)(kt) * sa),   \
                                                 (MBH_LDS void*)((LB) + (wave + v * NW) * AWS), 16, 0, 0); \
            _Pragma("unroll") for (int v = 0; v < B_WL; ++v)                                            \
                __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(pb[v] + (size_t)(kt) * sb),   \
                                                 (MBH_LDS void*)((LB) + OB0 + (wave + v * NW) * BWS), 16, 0, 0); \
        } else {                                                                                        \
            const size_t ga = (size_t)(kt) * A_CH, gb = (size_t)(kt) * B_CH;                            \
            _Pragma("unroll") for (int v = 0; v < A_WL; ++v) {                                          \
                const int blk = (wave + v * NW) * 64;                                                   \
                __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(gA0 + ga + blk),              \
                                                 (MBH_LDS void*)((LB) + blk), 16, 0, 0);                \
                if (NPL == 2)                                                                           \
                    __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(gA1 + ga + blk),          \
                                                     (MBH_LDS void*)((LB) + OA1 + blk), 16, 0, 0);      \
            }                                                                                           \
            _Pragma("unroll") for (int v = 0; v < B_WL; ++v) {                                          \
                const int blk = (wave + v * NW) * 64;                                                   \
                __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(gB0 + gb + blk),              \
                                                 (MBH_LDS void*)((LB) + OB0 + blk), 16, 0, 0);          \
                if (NPL == 2)                                                                           \
                    __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(gB1 + gb + blk),          \
                                                     (MBH_LDS void*)((LB) + OB1 + blk), 16, 0, 0);      \
            }                                                                                           \
        }                                                                                               \
        __builtin_amdgcn_sched_barrier(0);                                                              \
    } while (0)
#define MBH_STEP(FN, BUF, CH, FC, S)                                                                    \
    do {                                                                                                \
        packed_frag_read<NPL, TM, TN, CPR, !AMN, !BMN>(FN, BUF, OA1, OB0, OB1, wm0, wn0, lrow, CH);     \
        if constexpr (AMN) mn_frag_read<BM, TM, S>(FN.ah, BUF, 0, a_loff);                              \
        if constexpr (BMN) mn_frag_read<BN, TN, S>(FN.bh, BUF, OB0 * 16, b_loff);                       \
        packed_frag_mfma<NPROD, NPL, TM, TN, F16>(FC, acc);                                             \
        packed_interleave<NREAD, NMFMA>();                                                              \
        __builtin_amdgcn_sched_barrier(0);                                                              \
    } while (0)
#define MBH_SYNC()                                                                                      \
    do {                                                                                                \
        __builtin_amdgcn_s_waitcnt(0);                                                                  \
        __syncthreads();                                                                                \
    } while (0)

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    int kt0, kt1;
    kt_range(p.krange, tm, tn, BM, BN, BKT, p.ktiles, kt0, kt1);
    if constexpr (CK == 0) {
        if (p.ksplit > 1) {  // this workgroup's share of the tile's k-tiles
            const int len = kt1 - kt0;
            const int c0 = kt0 + (int)((int64_t)len * chunk / p.ksplit);
            kt1 = kt0 + (int)((int64_t)len * (chunk + 1) / p.ksplit);
            kt0 = c0;
        }
    }
    const int lrow = lane & 31, lgrp = lane >> 5, lsw = swz(lrow);
    // Chunks of the k-steps (16 k = two chunks; lanes 32-63 take the second).
    const int ch0 = lgrp ^ lsw, ch1 = (2 + lgrp) ^ lsw, ch2 = (4 + lgrp) ^ lsw, ch3 = (6 + lgrp) ^ lsw;
    const int a_loff = AMN ? PA::lane_off(lane, wm0) : 0, b_loff = BMN ? PB::lane_off(lane, wn0) : 0;
    PackedFrag<NPL, TM, TN> f0, f1;
    if (kt0 < kt1) {
        MBH_DMA(kt0, L0);
        MBH_SYNC();
        if (kt0 + 1 < kt1) MBH_DMA(kt0 + 1, L1);
        packed_frag_read<NPL, TM, TN, CPR, !AMN, !BMN>(f0, L0, OA1, OB0, OB1, wm0, wn0, lrow, ch0);
        if constexpr (AMN) mn_frag_read<BM, TM, 0>(f0.ah, L0, 0, a_loff);
        if constexpr (BMN) mn_frag_read<BN, TN, 0>(f0.bh, L0, OB0 * 16, b_loff);
    }
    // Straight-line body, one exit: the accumulators stay in place.  The last
    // step's read of the next k-tile may read a stale buffer (never used).
    // Inside a k-tile, step s is multiplied from f[s % 2] while step s+1 is
    // read into the other set (KS is even, so every k-tile starts in f0).
    int kt = kt0;
    for (; kt + 1 < kt1; kt += 2) {
        // L0: k-tile kt (f0 = its step 0); L1: k-tile kt+1 in flight
        MBH_STEP(f1, L0, ch1, f0, 1);
        if constexpr (KS == 4) {
            MBH_STEP(f0, L0, ch2, f1, 2);
            MBH_STEP(f1, L0, ch3, f0, 3);
        }
        MBH_SYNC();
        if (kt + 2 < kt1) MBH_DMA(kt + 2, L0);
        MBH_STEP(f0, L1, ch0, f1, 0);
        // L1: k-tile kt+1 (f0 = its step 0); L0: k-tile kt+2 in flight
        MBH_STEP(f1, L1, ch1, f0, 1);
        if constexpr (KS == 4) {
            MBH_STEP(f0, L1, ch2, f1, 2);
            MBH_STEP(f1, L1, ch3, f0, 3);
        }
        MBH_SYNC();
        if (kt + 3 < kt1) MBH_DMA(kt + 3, L1);
        MBH_STEP(f0, L0, ch0, f1, 0);
    }
    if (kt < kt1) {  // odd count: k-tile kt in L0, f0 = its step 0
        MBH_STEP(f1, L0, ch1, f0, 1);
        if constexpr (KS == 4) {
            MBH_STEP(f0, L0, ch2, f1, 2);
            MBH_STEP(f1, L0, ch3, f0, 3);
        }
        packed_frag_mfma<NPROD, NPL, TM, TN, F16>(f1, acc);
    }
#undef MBH_SYNC
#undef MBH_STEP
#undef MBH_DMA

    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    if constexpr (CK == 0) {
        if (p.ksplit > 1) {  // raw partial sum, the whole BM x BN block
            epilogue_32x32<TM, TN, 0>(acc, p.part + (size_t)blockIdx.x * BM * BN, BM, BM, BN, wm0, wn0, lane, 1.0f,
                                      0.0f, TRI_NONE);
            return;
        }
    }
    epilogue_32x32<TM, TN, CK>(acc, p.C, p.ldc, p.m, p.n, m0 + wm0, n0 + wn0, lane, p.alpha, p.beta, p.tri);
}


// Tile of this workgroup for the packed launches (see PackedGemmParams).
__device__ inline void packed_tile_of(const PackedGemmParams& p, int& tm, int& tn, int& chunk) {
    chunk = 0;
    if (p.ksplit > 1) {
        chunk = (int)blockIdx.x % p.ksplit;
        tile_coords_logical(p.tri_tiles, p.tile0 + (int)blockIdx.x / p.ksplit, p.tiles_m, p.tiles_n, tm, tn);
    } else if (p.tile0 > 0 || p.tri_tiles) {
        tile_coords_logical(p.tri_tiles, p.tile0 + xcd_remap((int)blockIdx.x, (int)gridDim.x), p.tiles_m, p.tiles_n,
                            tm, tn);
    } else {
        tile_coords_kr(p.krange, p.tiles_m, p.tiles_n, tm, tn);
    }
}

template <int NPROD, int BM, int BN, int WAVES_M, int WAVES_N>
__global__ __launch_bounds__(64 * WAVES_M * WAVES_N) void gemm_packed_kernel(PackedGemmParams p) {
    typedef PackedTileLds<NPROD, BM, BN> L;
    __shared__ u32x4 lds0[L::TOTAL / 2];
    __shared__ u32x4 lds1[L::TOTAL / 2];
    int tm, tn, chunk;
    packed_tile_of(p, tm, tn, chunk);
    if constexpr (NPROD != 8)
        gemm_packed_tile_pipe<NPROD, BM, BN, WAVES_M, WAVES_N>(p, tm, tn, lds0, lds1, chunk);
    else
        gemm_packed_tile_dma<NPROD, BM, BN, WAVES_M, WAVES_N>(p, tm, tn, lds0, lds1);
}

// Finishes the tiles of a split launch: C = alpha * (sum of the ksplit
// partial blocks, in chunk order) + beta * C on the tile's entries inside C
// and the triangle mask.  Block (t, y): tail tile t, CPB columns from
// CPB*y; each thread one column's 4 consecutive rows, read from every
// partial as one 16-byte load (rows are the contiguous index of a partial
// block).  Grid (tail tiles, BN / CPB), 256 threads.
template <int BM, int BN>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(PackedGemmParams p) {
    constexpr int TPC = BM / 4, CPB = 256 / TPC;  // threads per column, columns per block
    static_assert(TPC * CPB == 256 && BN % CPB == 0, "block shape");
    const int t = (int)blockIdx.x;
    int tm, tn;
    tile_coords_logical(p.tri_tiles, p.tile0 + t, p.tiles_m, p.tiles_n, tm, tn);
    const int i = 4 * ((int)threadIdx.x % TPC), j = CPB * (int)blockIdx.y + (int)threadIdx.x / TPC;
    const MBH_GLOBAL f32x4* part =
        (const MBH_GLOBAL f32x4*)(p.part + (size_t)t * p.ksplit * BM * BN + (size_t)j * BM + i);
    f32x4 sum = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int c = 0; c < p.ksplit; ++c) sum += part[(size_t)c * (BM * BN / 4)];
    MBH_GLOBAL float* C = (MBH_GLOBAL float*)p.C;
    const int64_t col = (int64_t)tn * BN + j;
    if (col >= p.n) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t row = (int64_t)tm * BM + i + e;
        if (row >= p.m || !tri_keep(p.tri, row, col)) continue;
        float v = p.alpha * sum[e];
        if (p.beta != 0.0f) v += p.beta * C[row + col * p.ldc];
        C[row + col * p.ldc] = v;
    }
}

// k per packed row of the native 16-bit gemm (its pack pass uses the same).
constexpr int GEMM16_BK = 64;

// Native 16-bit gemm: one plane per operand holding f16 (F16) or bf16 values
// (packed, or the caller's operand read in place: k-contiguous, or
// row-contiguous in the AMN / BMN instantiations), products on
// the MFMA of that format with fp32 accumulation, C stored in the same 16-bit
// format.
template <bool F16, int BM, int BN, int WAVES_M, int WAVES_N, bool AMN = false, bool BMN = false>
__global__ __launch_bounds__(64 * WAVES_M * WAVES_N) void gemm_packed16_kernel(PackedGemmParams p) {
    constexpr int CH = (AMN ? MnPanel<BM>::CH : BM * GEMM16_BK / 8) + (BMN ? MnPanel<BN>::CH : BN * GEMM16_BK / 8);
    __shared__ u32x4 lds0[CH];
    __shared__ u32x4 lds1[CH];
    int tm, tn;
    if (p.tri_tiles) tri_tile_coords(p.tri_tiles, xcd_remap((int)blockIdx.x, (int)gridDim.x), p.tiles_m, tm, tn);
    else tile_coords_kr(p.krange, p.tiles_m, p.tiles_n, tm, tn);
    gemm_packed_tile_pipe<1, BM, BN, WAVES_M, WAVES_N, F16 ? 1 : 2, F16, true, GEMM16_BK, AMN, BMN>(p, tm, tn, lds0,
                                                                                                 lds1);
}

}  // namespace mbh
