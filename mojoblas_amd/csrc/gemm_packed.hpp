// Macro-tile gemm over packed bf16 panels (see pack.hpp), fp32 accumulate.
//
//   NPROD = 3: BF16x3, acc += Bh*Ah + Bl*Ah + Bh*Al   (x = hi + lo per operand)
//   NPROD = 1: plain bf16 operands, acc += Bh*Ah
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
// Pipeline: the next k-tile's global loads are issued into registers before
// the current tile's MFMAs and written to the other LDS buffer after them;
// one barrier per k-tile.
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
    float* C;
    int64_t ldc;
    float alpha, beta;
    int tiles_m, tiles_n;
    int tri;
    int tri_tiles;  // 0 full grid, 1/2 lower/upper block triangle
    int krange;     // KRange (k-tiles visited per tile)
};

__device__ inline int lds_swz(int row) { return ((row >> 1) ^ (row >> 2)) & 3; }

template <int NPROD, int BM, int BN, int WAVES_M, int WAVES_N>
__global__ __launch_bounds__(64 * WAVES_M * WAVES_N) void gemm_packed_kernel(PackedGemmParams p) {
    constexpr int BK = 32;
    constexpr int NT = 64 * WAVES_M * WAVES_N;
    constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
    constexpr int TM = WM / 32, TN = WN / 32;
    constexpr int NPL = (NPROD == 3) ? 2 : 1;  // planes per operand
    constexpr int A_CH = BM * BK / 8;          // 16-byte chunks per A plane tile
    constexpr int B_CH = BN * BK / 8;
    constexpr int A_LD = A_CH / NT, B_LD = B_CH / NT;  // chunks per thread per plane
    static_assert(A_LD * NT == A_CH && B_LD * NT == B_CH, "tile not divisible by the thread block");
    // LDS: [buf][plane][rows * 64 bytes] as uint4 (16 bytes)
    __shared__ uint4 As[2][NPL][BM * 4];
    __shared__ uint4 Bs[2][NPL][BN * 4];

    int tm, tn;
    if (p.tri_tiles) tri_tile_coords(p.tri_tiles, (int)blockIdx.x, tm, tn);
    else tile_coords(p.tiles_m, p.tiles_n, tm, tn);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;

    const size_t a_tile = (size_t)BM * BK, b_tile = (size_t)BN * BK;  // bf16 elements per k-tile block
    const uint4* gA[NPL];
    const uint4* gB[NPL];
    gA[0] = reinterpret_cast<const uint4*>(p.Ah + (size_t)tm * p.ktiles * a_tile);
    gB[0] = reinterpret_cast<const uint4*>(p.Bh + (size_t)tn * p.ktiles * b_tile);
    if (NPL == 2) {
        gA[NPL - 1] = reinterpret_cast<const uint4*>(p.Al + (size_t)tm * p.ktiles * a_tile);
        gB[NPL - 1] = reinterpret_cast<const uint4*>(p.Bl + (size_t)tn * p.ktiles * b_tile);
    }

    uint4 ra[NPL][A_LD], rb[NPL][B_LD];
    auto gload = [&](int kt) {
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl) {
#pragma unroll
            for (int q = 0; q < A_LD; ++q) ra[pl][q] = gA[pl][(size_t)kt * A_CH + tid + q * NT];
#pragma unroll
            for (int q = 0; q < B_LD; ++q) rb[pl][q] = gB[pl][(size_t)kt * B_CH + tid + q * NT];
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl) {
#pragma unroll
            for (int q = 0; q < A_LD; ++q) {
                const int c = tid + q * NT, row = c >> 2, ch = c & 3;
                As[buf][pl][row * 4 + (ch ^ lds_swz(row))] = ra[pl][q];
            }
#pragma unroll
            for (int q = 0; q < B_LD; ++q) {
                const int c = tid + q * NT, row = c >> 2, ch = c & 3;
                Bs[buf][pl][row * 4 + (ch ^ lds_swz(row))] = rb[pl][q];
            }
        }
    };

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
        gload(kt0);
        lstore(0);
    }
    __syncthreads();

    // Per-lane fragment addressing: row offset lane&31, chunk 2*ks + lane/32,
    // swizzle depends only on lane bits 1..3 (row offsets are multiples of 32).
    const int lrow = lane & 31, lgrp = lane >> 5, lsw = lds_swz(lrow);
    for (int kt = kt0; kt < kt1; ++kt) {
        const int cur = (kt - kt0) & 1;
        if (kt + 1 < kt1) gload(kt + 1);
#pragma unroll
        for (int ks = 0; ks < BK / 16; ++ks) {
            const int ch = (2 * ks + lgrp) ^ lsw;
            bf16x8 ah[TM], bh[TN], al[TM], bl[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int idx = (wm0 + i * 32 + lrow) * 4 + ch;
                ah[i] = __builtin_bit_cast(bf16x8, As[cur][0][idx]);
                if (NPROD == 3) al[i] = __builtin_bit_cast(bf16x8, As[cur][NPL - 1][idx]);
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const int idx = (wn0 + j * 32 + lrow) * 4 + ch;
                bh[j] = __builtin_bit_cast(bf16x8, Bs[cur][0][idx]);
                if (NPROD == 3) bl[j] = __builtin_bit_cast(bf16x8, Bs[cur][NPL - 1][idx]);
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
        if (kt + 1 < kt1) lstore(cur ^ 1);
        __syncthreads();
    }

    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    const float alpha = p.alpha, beta = p.beta;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int64_t row = m0 + wm0 + i * 32 + lrow;
        if (row >= p.m) continue;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t col = n0 + wn0 + j * 32 + acc_col<32>(r, lane);
                if (col < p.n && tri_keep(p.tri, row, col)) {
                    float* cp = p.C + row + col * p.ldc;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * *cp;
                    *cp = v;
                }
            }
        }
    }
}

}  // namespace mbh
