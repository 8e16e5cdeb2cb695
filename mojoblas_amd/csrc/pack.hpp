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

template <bool KCONTIG, int PR, int PK>
__global__ __launch_bounds__(256) void pack_split_kernel(PackParams p) {
    constexpr int BR = 64, BKK = 32;  // fp32 block moved per workgroup
    static_assert(PR % BR == 0 && PK % BKK == 0, "pack block must tile the gemm block");
    __shared__ float s[BR][BKK + 1];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * BR;
    const int64_t k0 = (int64_t)blockIdx.y * BKK;
    if (KCONTIG) {
        // element (r, k) at X[k + r*ld]: 8 lanes cover one row's 32 k.
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
            const int r = pass * 32 + tid / 8, kq = (tid % 8) * 4;
            const int64_t gr = r0 + r;
            if (p.vec && gr < p.R && k0 + kq + 4 <= p.K) {
                const f32x4 v = *reinterpret_cast<const f32x4*>(p.X + (k0 + kq) + gr * p.ld);
#pragma unroll
                for (int e = 0; e < 4; ++e) s[r][kq + e] = v[e];
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t gk = k0 + kq + e;
                    s[r][kq + e] = (gr < p.R && gk < p.K) ? p.X[gk + gr * p.ld] : 0.0f;
                }
            }
        }
    } else {
        // element (r, k) at X[r + k*ld]: 64 lanes cover one k column's 64 rows.
#pragma unroll
        for (int pass = 0; pass < 8; ++pass) {
            const int kk = pass * 4 + tid / 64, r = tid % 64;
            const int64_t gr = r0 + r, gk = k0 + kk;
            s[r][kk] = (gr < p.R && gk < p.K) ? p.X[gr + gk * p.ld] : 0.0f;
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
    *reinterpret_cast<uint4*>(p.hi + off) = hv;
    if (p.lo) {
        lv.x = l[0] | (l[1] << 16);
        lv.y = l[2] | (l[3] << 16);
        lv.z = l[4] | (l[5] << 16);
        lv.w = l[6] | (l[7] << 16);
        *reinterpret_cast<uint4*>(p.lo + off) = lv;
    }
}

}  // namespace mbh
