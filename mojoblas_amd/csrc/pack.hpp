// Pack pass for the split-precision gemm: op(X) (fp32, any transpose, any ld)
// -> tile-blocked bf16 planes in HBM.
//
// Output layout (per plane): block (t, kt) of PR rows x PK k is contiguous,
// row-major with k fastest, at element offset ((t*ktiles + kt)*PR + r)*PK + kk;
// rows >= R and k >= K are zero.  The gemm kernel therefore streams each
// k-tile of a macro-tile panel as one contiguous, bounds-check-free run.
//
// Split: hi = rn_bf16(x), lo = rn_bf16(x - hi); x - hi is exact in fp32.
// One workgroup moves a 64-row x 32-k fp32 block through LDS so that both the
// HBM read (along whichever index is contiguous) and the packed write (k
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
};

// One 64-row x 32-k block (r0, k0) of the split pack; `s` is a 64 x 33
// float LDS tile.  Shared by the grid kernel and the batched pack kernel.
template <bool KCONTIG, int PR, int PK>
__device__ inline void pack_split_block(const PackParams& p, int64_t r0, int64_t k0, float (*s)[33]) {
    constexpr int BR = 64, BKK = 32;  // fp32 block moved per workgroup
    static_assert(PR % BR == 0 && PK % BKK == 0, "pack block must tile the gemm block");
    const int tid = threadIdx.x;
    const MBH_GLOBAL float* const X = (const MBH_GLOBAL float*)p.X;
    if (KCONTIG) {
        // element (r, k) at X[k + r*ld]: 8 lanes cover one row's 32 k.
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
            const int r = pass * 32 + tid / 8, kq = (tid % 8) * 4;
            const int64_t gr = r0 + r;
            if (p.vec && gr < p.R && k0 + kq + 4 <= p.K) {
                const f32x4 v = *(const MBH_GLOBAL f32x4*)(X + (k0 + kq) + gr * p.ld);
#pragma unroll
                for (int e = 0; e < 4; ++e) s[r][kq + e] = v[e];
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t gk = k0 + kq + e;
                    s[r][kq + e] = (gr < p.R && gk < p.K) ? X[gk + gr * p.ld] : 0.0f;
                }
            }
        }
    } else {
        // element (r, k) at X[r + k*ld]: 64 lanes cover one k column's 64 rows.
#pragma unroll
        for (int pass = 0; pass < 8; ++pass) {
            const int kk = pass * 4 + tid / 64, r = tid % 64;
            const int64_t gr = r0 + r, gk = k0 + kk;
            s[r][kk] = (gr < p.R && gk < p.K) ? X[gr + gk * p.ld] : 0.0f;
        }
    }
    __syncthreads();
    // Each thread: 8 consecutive k of one row -> 16 bytes per plane.
    const int r = tid / 4, kq = (tid % 4) * 8;
    const int64_t gr = r0 + r;
    const int64_t t = gr / PR, rr = gr % PR;
    const int64_t gk = k0 + kq;
    const int64_t kt = gk / PK, kk = gk % PK;
    const int64_t off = ((t * p.ktiles + kt) * PR + rr) * PK + kk;
    uint32_t h[8], l[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) split_bf16(s[r][kq + e], h[e], l[e]);
    uint4 hv, lv;
    hv.x = h[0] | (h[1] << 16);
    hv.y = h[2] | (h[3] << 16);
    hv.z = h[4] | (h[5] << 16);
    hv.w = h[6] | (h[7] << 16);
    *(MBH_GLOBAL u32x4*)(p.hi + off) = u32x4{hv.x, hv.y, hv.z, hv.w};
    if (p.lo) {
        lv.x = l[0] | (l[1] << 16);
        lv.y = l[2] | (l[3] << 16);
        lv.z = l[4] | (l[5] << 16);
        lv.w = l[6] | (l[7] << 16);
        *(MBH_GLOBAL u32x4*)(p.lo + off) = u32x4{lv.x, lv.y, lv.z, lv.w};
    }
}

template <bool KCONTIG, int PR, int PK>
__global__ __launch_bounds__(256) void pack_split_kernel(PackParams p) {
    __shared__ float s[64][33];
    pack_split_block<KCONTIG, PR, PK>(p, (int64_t)blockIdx.x * 64, (int64_t)blockIdx.y * 32, s);
}

}  // namespace mbh

namespace mbh {

// FP8 pack: op(X) -> e4m3 (OCP, max 448) with one power-of-two scale per
// 128-row x 128-k tile.  Packed like the bf16 planes (blocks of PR rows x 64
// k, one byte per element, k fastest, zero padding), scales as E8M0 bytes
// (exponent + 127) in a [R/128][K/128] int array: x ~= 2^(e) * q.
// One workgroup stages its whole 128x128 fp32 tile in LDS, reduces the tile's
// amax, picks e = ceil(log2(amax / 448)) and writes 16 quantized bytes per
// store.
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

template <bool KCONTIG, int PR>
__global__ __launch_bounds__(256) void pack_fp8_kernel(PackFp8Params p) {
    constexpr int T = 128, PK = 64;
    __shared__ float s[T][T + 1];
    __shared__ float red[256];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * T, k0 = (int64_t)blockIdx.y * T;
    float amax = 0.0f;
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
        const int idx = tid + 256 * it;
        if (KCONTIG) {  // (r, k) at X[k + r*ld]: 32 float4 per row
            const int r = idx / 32, kq = (idx % 32) * 4;
            const int64_t gr = r0 + r, gk = k0 + kq;
            f32x4 v = {0.f, 0.f, 0.f, 0.f};
            if (p.vec && gr < p.R && gk + 4 <= p.K) {
                v = *reinterpret_cast<const f32x4*>(p.X + gk + gr * p.ld);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = (gr < p.R && gk + e < p.K) ? p.X[gk + e + gr * p.ld] : 0.0f;
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                s[r][kq + e] = v[e];
                amax = fmaxf(amax, fabsf(v[e]));
            }
        } else {  // (r, k) at X[r + k*ld]: 32 float4 per k column
            const int kk = idx / 32, rq = (idx % 32) * 4;
            const int64_t gr = r0 + rq, gk = k0 + kk;
            f32x4 v = {0.f, 0.f, 0.f, 0.f};
            if (p.vec && gk < p.K && gr + 4 <= p.R) {
                v = *reinterpret_cast<const f32x4*>(p.X + gr + gk * p.ld);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = (gr + e < p.R && gk < p.K) ? p.X[gr + e + gk * p.ld] : 0.0f;
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                s[rq + e][kk] = v[e];
                amax = fmaxf(amax, fabsf(v[e]));
            }
        }
    }
    red[tid] = amax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] = fmaxf(red[tid], red[tid + w]);
        __syncthreads();
    }
    amax = red[0];
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
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const int idx = tid + 256 * it;
        const int r = idx / 8, kc = (idx % 8) * 16;
        const int64_t gr = r0 + r, gk = k0 + kc;
        const int64_t t = gr / PR, rr = gr % PR, kt = gk / PK, kk = gk % PK;
        uint4 w;
        w.x = fp8_pack4(s[r][kc + 0] * inv, s[r][kc + 1] * inv, s[r][kc + 2] * inv, s[r][kc + 3] * inv);
        w.y = fp8_pack4(s[r][kc + 4] * inv, s[r][kc + 5] * inv, s[r][kc + 6] * inv, s[r][kc + 7] * inv);
        w.z = fp8_pack4(s[r][kc + 8] * inv, s[r][kc + 9] * inv, s[r][kc + 10] * inv, s[r][kc + 11] * inv);
        w.w = fp8_pack4(s[r][kc + 12] * inv, s[r][kc + 13] * inv, s[r][kc + 14] * inv, s[r][kc + 15] * inv);
        *reinterpret_cast<uint4*>(p.q + ((t * p.ktiles + kt) * PR + rr) * PK + kk) = w;
    }
}

}  // namespace mbh
