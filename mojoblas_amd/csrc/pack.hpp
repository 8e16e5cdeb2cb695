// Pack pass for the split-precision gemm: op(X) (fp32, any transpose, any ld)
// -> tile-blocked bf16 planes in HBM.
//
// Output layout (per plane): block (t, kt) of PR rows x PK k is contiguous,
// row-major with k fastest, at element offset ((t*ktiles + kt)*PR + r)*PK + kk;
// rows >= R and k >= K are zero.  The gemm kernel therefore streams each
// k-tile of a macro-tile panel as one contiguous, bounds-check-free run.
//
// Split: hi = rn_bf16(x), lo = rn_bf16(x - hi); x - hi is exact in fp32.
// One workgroup moves a 64-row x 128-k fp32 block through LDS so that both
// the HBM read (along whichever index is contiguous) and the packed write (k
// fastest) are coalesced.
#pragma once
#include "common.hpp"

namespace mbh {

// Round-to-nearest-even fp32 -> bf16 bits (NaN stays NaN).
__device__ inline uint32_t bf16_rn_bits(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

// a and b rounded to bf16 (nearest, ties to even) and packed, a in the low
// half: one v_cvt_pk_bf16_f32.  Same bits as bf16_rn_bits for every
// non-NaN input (a NaN stays a quiet NaN).
__device__ inline uint32_t bf16x2_rn_bits(float a, float b) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t{a, b}), bf16x2_t));
}

__device__ inline void split_bf16(float x, uint32_t& h, uint32_t& l) {
    h = bf16_rn_bits(x);
    const float hf = __uint_as_float(h << 16);
    l = bf16_rn_bits(__builtin_isfinite(hf) ? x - hf : 0.0f);
}

struct PackParams {
    const float* X;
    int64_t ld;
    int64_t R, K;     // logical rows (m or n) and depth k of op(X)
    int ktiles;       // ceil(K / PK)
    uint16_t* hi;     // packed planes
    uint16_t* lo;     // may be null (single bf16 plane)
    int vec;          // X 16-byte aligned and ld % 4 == 0 (vector loads along k)
    int view;         // PackView: how op(X) is read (triangle, symmetric half)
};

// Classification of a 64-row x BKK-k pack block against the diagonal r = k
// of its operand view: how its elements are loaded.
enum PackBlockLoad : int {
    PBL_AS_LAID = 0,  // every element read where op(X) has it
    PBL_MIRROR = 1,   // symmetric view, every element read at (k, r)
    PBL_SELECT = 2,   // symmetric view, block crosses the diagonal: per element
    PBL_ZERO = 3,     // triangle view, block entirely outside the triangle
};

template <int BR, int BKK>
__device__ inline int pack_block_load(int view, int64_t r0, int64_t k0) {
    const bool all_le = k0 + BKK - 1 <= r0, all_ge = k0 >= r0 + BR - 1;  // k <= r / k >= r everywhere
    const bool all_lt = k0 + BKK - 1 < r0, all_gt = k0 > r0 + BR - 1;    // strictly
    switch (view & 7) {
        case PV_TRI_LE: return all_gt ? PBL_ZERO : PBL_AS_LAID;
        case PV_TRI_GE: return all_lt ? PBL_ZERO : PBL_AS_LAID;
        case PV_SYM_LE: return all_le ? PBL_AS_LAID : all_gt ? PBL_MIRROR : PBL_SELECT;
        case PV_SYM_GE: return all_ge ? PBL_AS_LAID : all_lt ? PBL_MIRROR : PBL_SELECT;
        default: return PBL_AS_LAID;
    }
}

// One 64-row x (32*NKB)-k block (r0, k0) of the split pack, staged in the
// LDS tile `s` (64 rows of 32*NKB + 1 floats).  Every thread issues all of
// its 16-byte loads (along whichever index is contiguous in HBM) before the
// first LDS store, so a workgroup keeps 32*NKB KB of reads in flight; the
// write side stores 16 bytes per plane per thread, 64 contiguous bytes per
// packed row.  The full 16-byte loads and the stores are non-temporal: the
// source is read once and the planes (537 MB for both operands at 8192^2)
// do not fit the 256 MB last-level cache anyway (pack 0.244 -> 0.236 ms at
// 8192^3, the main kernel's time unchanged).
//
// Load phase: element (r, k) of op(X) into s[r][k].  KCONTIG: (r, k) lies at
// X[k + r ld] (4 consecutive k of one row per load), else at X[r + k ld] (4
// consecutive rows at one k).  A mirrored block of a symmetric view reads
// (r, k) at the other address, which is this function with KCONTIG flipped.
template <bool KCONTIG, int NKB>
__device__ inline void pack_load_tile(const PackParams& p, int64_t r0, int64_t k0, float* s) {
    constexpr int BR = 64, BKK = 32 * NKB, S = BKK + 1, NT = 256;
    constexpr int NLD = BR * BKK / 4 / NT;  // 16-byte loads per thread
    static_assert(NLD * 4 * NT == BR * BKK, "block not divisible by the thread block");
    const int tid = threadIdx.x;
    const MBH_GLOBAL float* const X = (const MBH_GLOBAL float*)p.X;
    auto coords = [&](int q, int& r, int& kk) {
        const int idx = tid + q * NT;
        if (KCONTIG) {
            r = idx / (BKK / 4);
            kk = (idx % (BKK / 4)) * 4;
        } else {
            kk = idx / (BR / 4);
            r = (idx % (BR / 4)) * 4;
        }
    };
    f32x4 v[NLD];
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
        int r, kk;
        coords(q, r, kk);
        const int64_t gr = r0 + r, gk = k0 + kk;
        const bool full = p.vec && (KCONTIG ? (gr < p.R && gk + 4 <= p.K) : (gr + 4 <= p.R && gk < p.K));
        if (full) {
            v[q] = __builtin_nontemporal_load((const MBH_GLOBAL f32x4*)(KCONTIG ? X + gk + gr * p.ld
                                                                              : X + gr + gk * p.ld));
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t er = KCONTIG ? gr : gr + e, ek = KCONTIG ? gk + e : gk;
                v[q][e] = (er < p.R && ek < p.K) ? (KCONTIG ? X[ek + er * p.ld] : X[er + ek * p.ld]) : 0.0f;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
        int r, kk;
        coords(q, r, kk);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (KCONTIG) s[r * S + kk + e] = v[q][e];
            else s[(r + e) * S + kk] = v[q][e];
        }
    }
}

// Load phase of a block of a symmetric view that crosses the diagonal: each
// element read from the stored half (PV_SYM_LE: k <= r stored), one scalar
// load per element in the thread mapping of pack_load_tile<KCONTIG>.
template <bool KCONTIG, int NKB>
__device__ inline void pack_load_select(const PackParams& p, int64_t r0, int64_t k0, float* s) {
    constexpr int BR = 64, BKK = 32 * NKB, S = BKK + 1, NT = 256;
    constexpr int NLD = BR * BKK / 4 / NT;
    const int tid = threadIdx.x;
    const bool le = (p.view & 7) == PV_SYM_LE;
    float v[NLD][4];
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
        const int idx = tid + q * NT;
        const int r = KCONTIG ? idx / (BKK / 4) : (idx % (BR / 4)) * 4;
        const int kk = KCONTIG ? (idx % (BKK / 4)) * 4 : idx / (BR / 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t er = r0 + r + (KCONTIG ? 0 : e), ek = k0 + kk + (KCONTIG ? e : 0);
            const bool stored = le ? ek <= er : ek >= er;
            const int64_t rr = stored ? er : ek, kc = stored ? ek : er;  // read (rr, kc)
            v[q][e] = (er < p.R && ek < p.K) ? (KCONTIG ? p.X[kc + rr * p.ld] : p.X[rr + kc * p.ld]) : 0.0f;
        }
    }
#pragma unroll
    for (int q = 0; q < NLD; ++q) {
        const int idx = tid + q * NT;
        const int r = KCONTIG ? idx / (BKK / 4) : (idx % (BR / 4)) * 4;
        const int kk = KCONTIG ? (idx % (BKK / 4)) * 4 : idx / (BR / 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (KCONTIG) s[r * S + kk + e] = v[q][e];
            else s[(r + e) * S + kk] = v[q][e];
        }
    }
}

template <bool KCONTIG, int PR, int PK, int NKB>
__device__ inline void pack_split_block(const PackParams& p, int64_t r0, int64_t k0, float* s) {
    constexpr int BR = 64, BKK = 32 * NKB, S = BKK + 1;
    static_assert(PR % BR == 0 && PK == 32, "pack block must tile the gemm block");
    const int how = p.view ? pack_block_load<BR, BKK>(p.view, r0, k0) : PBL_AS_LAID;  // uniform
    if (how == PBL_AS_LAID) pack_load_tile<KCONTIG, NKB>(p, r0, k0, s);
    else if (how == PBL_MIRROR) pack_load_tile<!KCONTIG, NKB>(p, r0, k0, s);
    else if (how == PBL_SELECT) pack_load_select<KCONTIG, NKB>(p, r0, k0, s);
    if (how != PBL_ZERO) __syncthreads();
    // Each thread: 8 consecutive k of one row per 32-k sub-block.
    const int tid = threadIdx.x;
    const int r = tid / 4, kq = (tid % 4) * 8;
    const int64_t gr = r0 + r;
    const int64_t t = gr / PR, rr = gr % PR;
    const int tv = p.view & 7;
    const bool tri = tv == PV_TRI_LE || tv == PV_TRI_GE, unit = (p.view & PV_UNIT) != 0 && gr < p.R;
#pragma unroll
    for (int sb = 0; sb < NKB; ++sb) {
        const int64_t kt = (k0 + sb * 32) / PK;  // the same for every thread of the sub-block
        if (kt >= p.ktiles) break;
        const int64_t off = ((t * p.ktiles + kt) * PR + rr) * PK + kq;
        uint32_t h[8], l[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float x = how == PBL_ZERO ? 0.0f : s[r * S + sb * 32 + kq + e];
            if (tri) {  // triangle view: zeros outside, optional unit diagonal
                const int64_t gk = k0 + sb * 32 + kq + e;
                if (tv == PV_TRI_LE ? gk > gr : gk < gr) x = 0.0f;
                if (unit && gk == gr) x = 1.0f;
            }
            split_bf16(x, h[e], l[e]);
        }
        __builtin_nontemporal_store(
            u32x4{h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16)},
            (MBH_GLOBAL u32x4*)(p.hi + off));
        if (p.lo)
            __builtin_nontemporal_store(
                u32x4{l[0] | (l[1] << 16), l[2] | (l[3] << 16), l[4] | (l[5] << 16), l[6] | (l[7] << 16)},
                (MBH_GLOBAL u32x4*)(p.lo + off));
    }
}

// Both operands of one gemm in one launch: workgroups [0, nblk_a) pack op(A)
// (grid gx_a row blocks x k blocks), the rest op(B)^T.  KA / KB: k is the
// contiguous index of op(A) / op(B)^T.
constexpr int PACK_NKB = 4;  // 32-k sub-blocks per workgroup
template <bool KA, bool KB, int PR, int PK>
__global__ __launch_bounds__(256) void pack_split_kernel(PackParams a, PackParams b, int gx_a, int nblk_a, int gx_b) {
    __shared__ float s[64 * (32 * PACK_NKB + 1)];
    int w = (int)blockIdx.x;
    if (w < nblk_a) {
        pack_split_block<KA, PR, PK, PACK_NKB>(a, (int64_t)(w % gx_a) * 64, (int64_t)(w / gx_a) * 32 * PACK_NKB, s);
    } else {
        w -= nblk_a;
        pack_split_block<KB, PR, PK, PACK_NKB>(b, (int64_t)(w % gx_b) * 64, (int64_t)(w / gx_b) * 32 * PACK_NKB, s);
    }
}

// 16-bit pack (native f16 / bf16 gemm): op(X) of 16-bit elements -> the same
// tile-blocked single plane, bits copied unchanged (the products run on the
// MFMA of that format).
struct Pack16Params {
    const uint16_t* X;
    int64_t ld;
    int64_t R, K;  // logical rows (m or n) and depth k of op(X)
    int64_t Rp;    // rows of the packed plane (whole macro-tiles)
    int ktiles;    // ceil(K / PK)
    uint16_t* out;
    int vec;       // X 16-byte aligned and ld % 8 == 0
};

// One 128-row x 128-k block per workgroup, held in registers: each thread
// loads eight 16-byte runs and writes eight 16-byte packed runs (8 k of one
// row), so no LDS is needed.
//   KCONTIG ((r, k) at X[k + r ld]): v[j] = X(gr0 + 16j, gk0 .. gk0 + 7),
//     gr0 = r0 + tid/16, gk0 = k0 + 8 (tid%16): 16 lanes read one 256-byte
//     row segment;
//   else ((r, k) at X[r + k ld]): v[j] = X(gr0 .. gr0 + 7, gk0 + j),
//     gr0 = r0 + 8 (tid%16), gk0 = k0 + 8 (tid/16): 16 lanes read 128
//     consecutive rows of one column, and the 8 x 8 block is transposed in
//     registers before the store.
// Rows >= R and k >= K are written as zeros; rows >= Rp and k-tiles >=
// ktiles are not written.  PK: k per packed row (32 or 64).
template <bool KCONTIG, int PR, int PK>
__device__ inline void pack16_block(const Pack16Params& p, int64_t r0, int64_t k0) {
    const int tid = threadIdx.x;
    const MBH_GLOBAL uint16_t* const X = (const MBH_GLOBAL uint16_t*)p.X;
    const int64_t gr0 = KCONTIG ? r0 + tid / 16 : r0 + 8 * (tid % 16);
    const int64_t gk0 = KCONTIG ? k0 + 8 * (tid % 16) : k0 + 8 * (tid / 16);
    u32x4 v[8];
    if (p.vec && r0 + 128 <= p.R && k0 + 128 <= p.K) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
            v[j] = *(const MBH_GLOBAL u32x4*)(KCONTIG ? X + gk0 + (gr0 + 16 * j) * p.ld : X + gr0 + (gk0 + j) * p.ld);
    } else {
        auto elem = [&](int j, int e) -> uint32_t {
            const int64_t gr = KCONTIG ? gr0 + 16 * j : gr0 + e, gk = KCONTIG ? gk0 + e : gk0 + j;
            return (gr < p.R && gk < p.K) ? (uint32_t)X[KCONTIG ? gk + gr * p.ld : gr + gk * p.ld] : 0u;
        };
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int w = 0; w < 4; ++w) v[j][w] = elem(j, 2 * w) | (elem(j, 2 * w + 1) << 16);
    }
    // gk0 is a multiple of 8, so a thread's 8 k stay inside one k-tile.
    if (gk0 / PK >= p.ktiles) return;
    auto dst = [&](int64_t gr) {
        return (MBH_GLOBAL u32x4*)(p.out + (((gr / PR) * p.ktiles + gk0 / PK) * PR + gr % PR) * PK + gk0 % PK);
    };
    if (KCONTIG) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (gr0 + 16 * j < p.Rp) *dst(gr0 + 16 * j) = v[j];
    } else {
        // Row gr0 + e: element e of v[0..7] (16-bit half e%2 of word e/2).
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (gr0 + e >= p.Rp) continue;
            u32x4 w;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t lo = v[2 * q][e / 2], hi = v[2 * q + 1][e / 2];
                w[q] = (e & 1) ? ((lo >> 16) | (hi & 0xffff0000u)) : ((lo & 0xffffu) | (hi << 16));
            }
            *dst(gr0 + e) = w;
        }
    }
}

// Both operands in one launch: workgroups [0, nblk_a) pack op(A) (gx_a row
// blocks of 128 x k blocks of 128), the rest op(B)^T.
template <bool KA, bool KB, int PR, int PK>
__global__ __launch_bounds__(256) void pack16_kernel(Pack16Params a, Pack16Params b, int gx_a, int nblk_a, int gx_b) {
    int w = (int)blockIdx.x;
    if (w < nblk_a) {
        pack16_block<KA, PR, PK>(a, (int64_t)(w % gx_a) * 128, (int64_t)(w / gx_a) * 128);
    } else {
        w -= nblk_a;
        pack16_block<KB, PR, PK>(b, (int64_t)(w % gx_b) * 128, (int64_t)(w / gx_b) * 128);
    }
}

}  // namespace mbh

namespace mbh {

// FP8 pack: op(X) -> e4m3 (OCP, max 448) with one power-of-two scale per
// 128-row x 128-k tile.  Packed like the bf16 planes (blocks of PR rows x 64
// k, one byte per element, k fastest, zero padding), scales as E8M0 bytes
// (exponent + 127) in a [R/128][K/128] int array: x ~= 2^(e) * q.
// One workgroup holds its whole 128x128 fp32 tile in registers (64 values
// per thread: 4 consecutive k of 16 rows when k is contiguous in HBM, 16
// consecutive k of 4 rows otherwise, so that every 16-byte load of a wave
// covers whole 512-byte runs), reduces the tile's amax (wave shuffles, then
// four partials in LDS), picks e = ceil(log2(amax / 448)) and writes the
// quantized k-runs (4 or 16 bytes per row) straight from registers.
struct PackFp8Params {
    const float* X;
    int64_t ld;
    int64_t R, K;
    int ktiles;        // packed 64-deep k-tiles (even: K padded to 128)
    int kblocks;       // K / 128 (rounded up)
    uint8_t* q;        // packed e4m3
    int* scale;        // E8M0 per 128x128 tile
    int vec;           // 16-byte loads allowed
};

__device__ inline uint32_t fp8_pack4(float a, float b, float c, float d) {
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
    return (uint32_t)w;
}

// One 128-row x 128-k block (bx, by) of op(X); `red` is four floats of the
// workgroup's LDS.
template <bool KCONTIG, int PR>
__device__ inline void pack_fp8_block(const PackFp8Params& p, int64_t bx, int64_t by, float* red) {
    constexpr int T = 128, PK = 64;
    const int tid = threadIdx.x;
    const int64_t r0 = bx * T, k0 = by * T;
    const bool full = p.vec && r0 + T <= p.R && k0 + T <= p.K;  // no edge: unguarded 16-byte loads
    // KCONTIG ((r, k) at X[k + r ld]): v[j] = X(gr0 + 8j, gk0 .. gk0 + 3),
    //   16 rows from gr0 = r0 + tid/32 (stride 8) at the 4 consecutive k from
    //   gk0 = k0 + 4 (tid%32): a wave reads two whole 128-k rows per load;
    // else ((r, k) at X[r + k ld]): v[j] = X(gr0 .. gr0 + 3, gk0 + j), four
    //   rows from gr0 = r0 + 4 (tid%32) and 16 consecutive k from
    //   gk0 = k0 + 16 (tid/32).
    f32x4 v[16];
    const int64_t gr0 = KCONTIG ? r0 + tid / 32 : r0 + 4 * (tid % 32);
    const int64_t gk0 = KCONTIG ? k0 + 4 * (tid % 32) : k0 + 16 * (tid / 32);
    if (full) {
#pragma unroll
        for (int j = 0; j < 16; ++j)
            v[j] = KCONTIG ? *reinterpret_cast<const f32x4*>(p.X + gk0 + (gr0 + 8 * j) * p.ld)
                           : *reinterpret_cast<const f32x4*>(p.X + gr0 + (gk0 + j) * p.ld);
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t gr = KCONTIG ? gr0 + 8 * j : gr0 + e, gk = KCONTIG ? gk0 + e : gk0 + j;
                v[j][e] = (gr < p.R && gk < p.K) ? p.X[KCONTIG ? gk + gr * p.ld : gr + gk * p.ld] : 0.0f;
            }
    }
    float amax = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) amax = fmaxf(amax, fabsf(v[j][e]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off, 64));
    if ((tid & 63) == 0) red[tid >> 6] = amax;
    __syncthreads();
    amax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    // Smallest power of two s with amax / s <= 448 (e4m3 max); 1 for a zero tile.
    int e = 0;
    if (amax > 0.0f && __builtin_isfinite(amax)) {
        e = (int)ceilf(log2f(amax / 448.0f));
        if (ldexpf(448.0f, e) < amax) ++e;
        if (e < -127) e = -127;
        if (e > 127) e = 127;
    }
    const float inv = ldexpf(1.0f, -e);
    if (tid == 0) p.scale[(r0 / T) * p.kblocks + k0 / T] = e + 127;
    // Packed byte offset of element (gr, gk), gk a multiple of 4.
    auto dst = [&](int64_t gr, int64_t gk) {
        return p.q + (((gr / PR) * p.ktiles + gk / PK) * PR + gr % PR) * PK + gk % PK;
    };
    if (KCONTIG) {  // 4 quantized k per row: one dword, a wave writing whole 64-byte runs
#pragma unroll
        for (int j = 0; j < 16; ++j)
            *reinterpret_cast<uint32_t*>(dst(gr0 + 8 * j, gk0)) =
                fp8_pack4(v[j][0] * inv, v[j][1] * inv, v[j][2] * inv, v[j][3] * inv);
    } else {
#pragma unroll
        for (int e4 = 0; e4 < 4; ++e4) {
            uint4 w;
            w.x = fp8_pack4(v[0][e4] * inv, v[1][e4] * inv, v[2][e4] * inv, v[3][e4] * inv);
            w.y = fp8_pack4(v[4][e4] * inv, v[5][e4] * inv, v[6][e4] * inv, v[7][e4] * inv);
            w.z = fp8_pack4(v[8][e4] * inv, v[9][e4] * inv, v[10][e4] * inv, v[11][e4] * inv);
            w.w = fp8_pack4(v[12][e4] * inv, v[13][e4] * inv, v[14][e4] * inv, v[15][e4] * inv);
            *reinterpret_cast<uint4*>(dst(gr0 + e4, gk0)) = w;
        }
    }
}

// Both operands of one FP8 gemm in one launch: workgroups [0, nblk_a) pack
// op(A) (gx_a row blocks x k blocks), the rest op(B)^T.  KA / KB: k is the
// contiguous index of op(A) / op(B)^T.
template <bool KA, bool KB, int PR>
__global__ __launch_bounds__(256) void pack_fp8_kernel(PackFp8Params a, PackFp8Params b, int gx_a, int nblk_a,
                                                       int gx_b) {
    __shared__ float red[4];
    int w = (int)blockIdx.x;
    if (w < nblk_a) {
        pack_fp8_block<KA, PR>(a, w % gx_a, w / gx_a, red);
    } else {
        w -= nblk_a;
        pack_fp8_block<KB, PR>(b, w % gx_b, w / gx_b, red);
    }
}

}  // namespace mbh
