// Variable-size batched gemm: thousands of independent N,N problems
//   C_p := alpha * A_p * B_p + beta * C_p,   p = 0..count-1,
// with their own m, n, k and leading dimensions, run by one persistent
// kernel.  (The reference has no batched routine; each problem has the
// semantics of gemm("N", "N", ...), src/level3/gemm.mojo:56-124.)
//
// Plan (host, once per shape list): validate, give every problem a tile class
// from the tuning table, enumerate its C tiles as work items and copy them to
// HBM, either as one fixed list per workgroup of each class's persistent grid
// (balanced longest-processing-time first over the XCDs and their
// workgroups) or, for batches the warp-specialized tiles cannot take, as
// work queues.
// Exec (device only, no allocation, no host synchronisation, capturable in a
// hipGraph): a fill kernel writes one record per item from the caller's
// pointer arrays and the warp-specialized kernels walk their lists; or a
// one-thread kernel resets the queue counters and the work-queue kernels'
// workgroups claim items until the queues are drained.  In bf16 mode fp32
// operands are rounded to bf16 as they are staged into LDS (no conversion
// pass), and bf16-stored operands are staged as they are.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <vector>

#include "../../include/mojoblas_hip.h"
#include "abi_internal.hpp"
#include "gemm_fp.hpp"
#include "gemm_packed.hpp"
#include "pack.hpp"
#include "vbatched_tuning.hpp"

namespace mbh {

struct VbProblem {
    int64_t m, n, k, lda, ldb, ldc;
    int cls, tiles_n;
};

struct VbItem {
    int prob, tile;
};

// Tile classes: 0 = 64x64 (2x2 waves of 32x32), 1 = 128x128 (2x2 waves of 64x64).
__host__ __device__ constexpr int vb_tile(int cls) { return cls == 0 ? 64 : 128; }

namespace {

// Work queues of the lean path: VB_QUEUES per tile class, one per XCD, each
// with its own counter on its own 256-byte line (VB_CTR_STRIDE ints apart),
// so the claims of a class spread over VB_QUEUES addresses instead of
// serialising on one.  The two-class kernels use counter 0 alone.
constexpr int VB_QUEUES = 8, VB_CTR_STRIDE = 64;
constexpr int VB_COUNTERS = 2 * VB_QUEUES;

__global__ void vb_reset_kernel(int* ctr) {
    if (blockIdx.x == 0 && threadIdx.x < VB_COUNTERS) ctr[threadIdx.x * VB_CTR_STRIDE] = 0;
}

// fp32 batched tiles: the 64x64 class runs 32-deep k-tiles with prefetch
// distance two (its short k-loops are latency-bound); the 128x128 class keeps
// 16-deep k-tiles at distance one, whose register footprint leaves room for
// two resident workgroups per CU.
constexpr int VB_BK64 = 32, VB_BK128 = 16;
constexpr int FP_LDS64 = FpTileLds<float, float, float, false, false, 64, 64, VB_BK64>::BYTES;
constexpr int FP_LDS128 = FpTileLds<float, float, float, false, false, 128, 128, VB_BK128>::BYTES;
constexpr int FP_LDS = FP_LDS128 > FP_LDS64 ? FP_LDS128 : FP_LDS64;

// Work-queue walk shared by the persistent kernels: thread 0 claims item
// i+1 from the counter while the workgroup computes item i, so the atomic's
// latency hides behind the tile.  Every workgroup leaves once the counter
// passes nitems.
template <typename F>
__device__ inline void vb_walk(int* ctr, int nitems, int* slot, F&& body) {
    if (threadIdx.x == 0) *slot = atomicAdd(ctr, 1);
    __syncthreads();
    int it = *slot;
    while (it < nitems) {
        int next = 0;
        if (threadIdx.x == 0) next = atomicAdd(ctr, 1);
        body(it);
        __syncthreads();  // everyone is done with *slot and with the tile's LDS
        if (threadIdx.x == 0) *slot = next;
        __syncthreads();
        it = *slot;
    }
}

// ---- lean tiles -------------------------------------------------------------
// One T x T C tile (T = 64 or 128) of a batched problem on 256 threads
// (2 x 2 waves of T/2 x T/2), 32-deep k-tiles double-buffered in LDS through
// registers.  The batched operands are column-major with ld = rows, so
// columns start at any 4-byte offset; the loads are 16-byte loads at dword
// alignment (which the hardware serves as one vector load) at 32-bit byte
// offsets from the problem's (uniform) base.  The plan takes this path only
// when every problem's offsets fit in int32 (vb_lean_ok).
//
//   A staging: load q (< T/32) of a wave covers rows m0 + 32q .. + 31: lane
//              l holds rows m0 + 32q + 4(l % 8) + e, e < 4 (one 16-byte
//              load), at k = k0 + 8 wave + kperm(l / 8);
//   B staging: load q covers columns n0 + 32q .. + 31: lane l holds
//              k = k0 + 4(l % 8) + e, e < 4, in column n0 + 32q + 8 wave + l / 8.
// A load whose 32 rows (A) or columns (B) all lie past the problem's edge is
// not issued and its LDS rows keep stale data; so are the MFMAs and the
// stores of a wave's 32 x 32 blocks past the edge.  Stale rows and columns,
// like rows >= m of A and columns >= n of B inside an issued load (which
// hold other data of the operand), only reach C entries that are never
// stored.  k >= K is zeroed in both operands.
// Edges: a group of four that would run past the operand's last element is
// loaded from its last four instead and shifted back, so every load stays
// inside the operand.  Operands of fewer than four elements load single
// elements.
//
// LDS, per buffer (every write and fragment read is free of bank conflicts):
//   fp32: A [k][T + 4] floats (one 16-byte write of four rows per load), B
//         [n][36] floats (one 16-byte write of four k per load).  The 2-deep
//         steps of v_mfma_f32_32x32x2_f32 visit k in the order 8c + 4h + e
//         (step (c, e), lane half h): one float of A per step and one
//         16-byte read of B per four steps; A and B use the same order, so
//         the sum runs over the same products.
//   bf16: A [k][T] bf16 with 2T + 32 bytes per k-row (one 8-byte write of
//         four rows per load; the k-rows of lanes 0-7 and 8-15 of a load sit
//         two rows apart, 64 bytes apart in the banks), fragments by two
//         ds_read_b64_tr_b16 per 32-row block; B the swizzled 64-byte rows of
//         the split-precision kernels (8-byte writes of four k, one 16-byte
//         read per fragment); v_mfma_f32_32x32x16_bf16.
template <bool BF16, int T>
struct VbLean {
    static_assert(T == 64 || T == 128, "lean tiles are 64 or 128 wide");
    static constexpr int NQ = T / 32;  // loads per operand per thread
    static constexpr int TW = T / 64;  // 32x32 blocks per wave and dimension
    // fp32: row strides in floats; bf16: A k-row stride in bytes.
    static constexpr int SAF = T + 4, SBF = 36, SAH = 2 * T + 32;
    static constexpr int A_BYTES = BF16 ? 32 * SAH : 32 * SAF * 4;
    static constexpr int B_BYTES = BF16 ? T * 64 : T * SBF * 4;
    static constexpr int BUF = A_BYTES + B_BYTES;  // bytes per buffer
    static constexpr int BYTES = 2 * BUF;
    static constexpr int MINWG = T == 64 ? 4 : 2;  // resident workgroups per CU (register bound)
};

// k-row of a lane group's A load: groups 2p and 2p + 1 two rows apart.
__device__ inline int vb_kperm(int j) { return (j & 4) | ((j & 1) << 1) | ((j >> 1) & 1); }

// One work item of the lean path (a tile of one problem), read once from the
// plan; the item index is wave-uniform, so this sits in scalar registers.
struct VbLeanItem {
    const MBH_GLOBAL float* A;
    const MBH_GLOBAL float* B;
    MBH_GLOBAL float* C;
    int m, n, K, lda, ldb, ldc, m0, n0, nkt;
};

template <int T>
__device__ inline VbLeanItem vb_lean_item(const VbProblem* __restrict__ probs, const VbItem* __restrict__ items,
                                          int it, const float* const* A, const float* const* B, float* const* C) {
    const VbItem w = items[it];
    const VbProblem pr = probs[w.prob];
    VbLeanItem d;
    d.A = (const MBH_GLOBAL float*)A[w.prob];
    d.B = (const MBH_GLOBAL float*)B[w.prob];
    d.C = (MBH_GLOBAL float*)C[w.prob];
    d.m = (int)pr.m;
    d.n = (int)pr.n;
    d.K = (int)pr.k;
    d.lda = (int)pr.lda;
    d.ldb = (int)pr.ldb;
    d.ldc = (int)pr.ldc;
    d.m0 = (w.tile / pr.tiles_n) * T;
    d.n0 = (w.tile % pr.tiles_n) * T;
    d.nkt = (d.K + 31) / 32;
    return d;
}

// 16 bytes at byte offset off (a multiple of 4) from a uniform base: one
// global_load_dwordx4 with the base in scalar registers and a 32-bit offset.
__device__ inline float4 vb_load4(const MBH_GLOBAL float* base, unsigned off) {
    float4 v;
    __builtin_memcpy(&v, (const MBH_GLOBAL unsigned char*)base + off, 16);
    return v;
}

// 16 bytes to / from a dword-aligned global address (one dwordx4 access).
__device__ inline f32x4 vb_ld16(const MBH_GLOBAL float* p) {
    f32x4 v;
    __builtin_memcpy(&v, (const MBH_GLOBAL unsigned char*)p, 16);
    return v;
}
__device__ inline void vb_st16(MBH_GLOBAL float* p, const f32x4& v) {
    __builtin_memcpy((MBH_GLOBAL unsigned char*)p, &v, 16);
}

// a * b (low 32 bits) as its own v_mul_lo_u32.  Written as a multiply-add,
// the compiler may form v_mad_u64_u32 whose unused high addend is a register
// that a load in flight will write, and then wait for that load.
__device__ inline int vb_mul(int a, int b) {
    int r;
    asm("v_mul_lo_u32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// 16 / 4 bytes at byte offset off (a multiple of 4) from a uniform base.
__device__ inline u32x4 vb_ld16w(const MBH_GLOBAL unsigned char* base, unsigned off) {
    u32x4 v;
    __builtin_memcpy(&v, (const MBH_GLOBAL unsigned*)(base + off), 16);
    return v;
}
__device__ inline unsigned vb_ld4w(const MBH_GLOBAL unsigned char* base, unsigned off) {
    return *(const MBH_GLOBAL unsigned*)(base + off);
}

// Eight 16-bit elements from the five dwords d, d4 that hold them: dwords
// d..d4 shifted down by 16 bits when bit 3 of f is set (the first element in
// a dword's upper half), else d itself.
__device__ inline u32x4 vb_join8(const u32x4& d, unsigned d4, int f) {
    const unsigned sh = (f & 8) ? 16u : 0u;
    return u32x4{__builtin_amdgcn_alignbit(d.y, d.x, sh), __builtin_amdgcn_alignbit(d.z, d.y, sh),
                 __builtin_amdgcn_alignbit(d.w, d.z, sh), __builtin_amdgcn_alignbit(d4, d.w, sh)};
}

// Eight 16-bit elements shifted down by s (0 .. 7): element e of the result
// is element e + s of w for e + s <= 7 (the others are left unspecified).
__device__ inline u32x4 vb_shift8(const u32x4& w, int s) {
    const int dq = s >> 1;
    const unsigned t0 = dq == 0 ? w.x : dq == 1 ? w.y : dq == 2 ? w.z : w.w;
    const unsigned t1 = dq == 0 ? w.y : dq == 1 ? w.z : w.w;
    const unsigned t2 = dq == 0 ? w.z : w.w;
    const unsigned t3 = w.w;
    const unsigned sh = (s & 1) ? 16u : 0u;
    return u32x4{__builtin_amdgcn_alignbit(t1, t0, sh), __builtin_amdgcn_alignbit(t2, t1, sh),
                 __builtin_amdgcn_alignbit(t3, t2, sh), __builtin_amdgcn_alignbit(t3, t3, sh)};
}

// v shifted down by s elements: out[e] = v[min(e + s, 3)], as a two-stage
// shifter of selects (shift by 2, then by 1; s >= 3 shifts by 3).
__device__ inline void vb_shift4(const float4& v, int s, float* out) {
    const bool b2 = s >= 2, b1 = (s & 1) || s >= 3;
    const float t0 = b2 ? v.z : v.x, t1 = b2 ? v.w : v.y, t2 = b2 ? v.w : v.z;
    out[0] = b1 ? t1 : t0;
    out[1] = b1 ? t2 : t1;
    out[2] = b1 ? v.w : t2;
    out[3] = v.w;
}

// Persistent walk over the work queues of T x T items with the k-loop
// pipelined across items: the last k-step of an item loads the first k-tile
// of the next item (claimed one item ahead), so a workgroup moves from tile
// to tile without an exposed load.  Two slots in LDS alternate between
// items: the slot an item reads was written before the previous item's last
// barrier, and the other slot is rewritten only after every thread has
// passed a barrier since reading it (each item passes at least one).  Thread
// 0 issues the claim when an item starts and resolves and publishes it
// before the item's last barrier.
// Queue x holds items [qoff[x], qoff[x + 1]) (the plan puts all tiles of a
// problem in one queue, so its operands are fetched into one XCD's L2).  A
// workgroup serves the queue of its XCD first (blockIdx % 8: workgroups are
// dispatched to the XCDs round-robin; the mapping only affects locality),
// then the others in turn once its own is drained, and leaves when all are.
template <bool BF16, int T>
__global__ __launch_bounds__(256, (VbLean<BF16, T>::MINWG)) void vb_gemm_lean_kernel(
    const VbProblem* __restrict__ probs, const VbItem* __restrict__ items, int nitems, int* ctr,
    const int* __restrict__ qoff, float alpha, float beta, const float* const* A, const float* const* B,
    float* const* C) {
    typedef VbLean<BF16, T> L;
    constexpr int NQ = L::NQ, TW = L::TW, SAF = L::SAF, SBF = L::SBF, SAH = L::SAH;
    __shared__ __attribute__((aligned(16))) unsigned char lds[L::BYTES];
    __shared__ int slot[2];
    __shared__ int qo[VB_QUEUES + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g8 = lane & 7;                          // A: row group; B: k group
    const int ka = 8 * wave + vb_kperm(lane >> 3);    // A: k-row of the lane's loads
    const int cb = tid >> 3;                          // B: column within a 32-column load
    const int wm0 = (wave & 1) * (T / 2), wn0 = (wave >> 1) * (T / 2);
    const int lr = lane & 31, lh = lane >> 5;
    // Staged k-tile: va[q] holds the four elements of A from linear index
    // r + k lda - sa[q] (r = m0 + 32q + 4 g8, k = k0 + ka), vb[q] those of B
    // from kk + col ldb - sb[q] (kk = k0 + 4 g8, col = n0 + cb + 32q); the
    // store shifts them back by sa / sb and zeroes k >= K.  The loads only
    // issue here and are consumed by the store after the k-tile's MFMAs.
    // Loads q >= qa (A) / qb (B) are not issued (past the tile's edge).
    float4 va[NQ], vb[NQ];
    int sa[NQ], sb[NQ];
    auto nload = [&](int extent) { return min(NQ, (extent + 31) >> 5); };
    auto load = [&](const VbLeanItem& d, int k0) {
        const int qa = nload(d.m - d.m0), qb = nload(d.n - d.n0);
        const int kk = k0 + 4 * g8;
        // Last element of each operand (column-major, the final column
        // holds m (A) or k (B) elements).
        const int alast = (d.K - 1) * d.lda + d.m - 1, blast = (d.n - 1) * d.ldb + d.K - 1;
        if (alast >= 3 && blast >= 3) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (q < qa) {
                    const int ia = d.m0 + 32 * q + 4 * g8 + (k0 + ka) * d.lda, ja = min(ia, alast - 3);
                    sa[q] = ia - ja;
                    va[q] = vb_load4(d.A, 4u * (unsigned)ja);
                }
                if (q < qb) {
                    const int ib = kk + min(d.n0 + cb + 32 * q, d.n - 1) * d.ldb, jb = min(ib, blast - 3);
                    sb[q] = ib - jb;
                    vb[q] = vb_load4(d.B, 4u * (unsigned)jb);
                }
            }
        } else {
            // Operands of fewer than four elements: single elements, relaxed
            // atomic loads so that they stay apart from the 16-byte loads.
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                sa[q] = 0;
                sb[q] = 0;
                const int r = d.m0 + 32 * q + 4 * g8, k = k0 + ka, col = d.n0 + cb + 32 * q;
                float a4[4], b4[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int ia = r + e < d.m && k < d.K ? r + e + k * d.lda : 0;
                    const int ib = kk + e < d.K && col < d.n ? kk + e + col * d.ldb : 0;
                    a4[e] = __hip_atomic_load((const float*)d.A + ia, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    b4[e] = __hip_atomic_load((const float*)d.B + ib, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                va[q] = float4{a4[0], a4[1], a4[2], a4[3]};
                vb[q] = float4{b4[0], b4[1], b4[2], b4[3]};
            }
        }
    };
    // krem = K - k0 of the staged k-tile.  In a k-tile with krem > 32 every
    // k lies below K - 1, so a group of B (four k of one column) or a group
    // of A holding a row < m (four rows of one column, lda >= 3) ends at or
    // before the operand's last element: no shift and no k mask.  (With
    // lda <= 2, four rows from a valid one can run past the last element.)
    auto store = [&](MBH_LDS unsigned char* buf, const VbLeanItem& d, int krem) {
        const int qa = nload(d.m - d.m0), qb = nload(d.n - d.n0);
        float ra[4 * NQ], rb[4 * NQ];
        if (krem > 32 && d.lda >= 3) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                ra[4 * q] = va[q].x, ra[4 * q + 1] = va[q].y, ra[4 * q + 2] = va[q].z, ra[4 * q + 3] = va[q].w;
                rb[4 * q] = vb[q].x, rb[4 * q + 1] = vb[q].y, rb[4 * q + 2] = vb[q].z, rb[4 * q + 3] = vb[q].w;
            }
        } else {
            const bool kva = ka < krem;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                float a4[4], b4[4];
                vb_shift4(va[q], sa[q], a4);
                vb_shift4(vb[q], sb[q], b4);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    ra[4 * q + e] = kva ? a4[e] : 0.0f;
                    rb[4 * q + e] = 4 * g8 + e < krem ? b4[e] : 0.0f;
                }
            }
        }
        if constexpr (BF16) {
            typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
            MBH_LDS unsigned char* arow = buf + ka * SAH + 8 * g8;
            MBH_LDS uint16_t* b16 = (MBH_LDS uint16_t*)(buf + L::A_BYTES);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (q < qa)
                    *(MBH_LDS u32x2*)(arow + 64 * q) = u32x2{bf16x2_rn_bits(ra[4 * q], ra[4 * q + 1]),
                                                             bf16x2_rn_bits(ra[4 * q + 2], ra[4 * q + 3])};
                if (q < qb) {
                    const int j = cb + 32 * q, k4 = 4 * g8;
                    *(MBH_LDS u32x2*)(b16 + j * 32 + (((k4 >> 3) ^ lds_swz(j)) << 3) + (k4 & 7)) =
                        u32x2{bf16x2_rn_bits(rb[4 * q], rb[4 * q + 1]), bf16x2_rn_bits(rb[4 * q + 2], rb[4 * q + 3])};
                }
            }
        } else {
            MBH_LDS float* as = (MBH_LDS float*)buf + ka * SAF + 4 * g8;
            MBH_LDS float* bs = (MBH_LDS float*)(buf + L::A_BYTES) + cb * SBF + 4 * g8;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (q < qa)
                    *(MBH_LDS f32x4*)(as + 32 * q) = f32x4{ra[4 * q], ra[4 * q + 1], ra[4 * q + 2], ra[4 * q + 3]};
                if (q < qb)
                    *(MBH_LDS f32x4*)(bs + 32 * q * SBF) = f32x4{rb[4 * q], rb[4 * q + 1], rb[4 * q + 2], rb[4 * q + 3]};
            }
        }
    };
    f32x16 acc[TW][TW];
#pragma unroll
    for (int i = 0; i < TW; ++i)
#pragma unroll
        for (int j = 0; j < TW; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    // MFMAs of one k-tile over the wave's blocks below the tile's edge
    // (bi block-rows, bj block-columns).  nk: k values of the tile below K;
    // only the MFMA steps that hold such k run (the rows past K are zero in
    // both operands, so the sum is unchanged).  FULL: every block inside.
    // The scheduling barrier keeps the staging of the next k-tile (which
    // waits for its loads) behind this k-tile's MFMAs.
    auto compute = [&](const MBH_LDS unsigned char* buf, int nk, auto full, int bi, int bj) {
        constexpr bool FULL = decltype(full)::value;
        auto on = [&](int i, int j) { return FULL || (i < bi && j < bj); };
        if constexpr (BF16) {
            typedef short v4i16 __attribute__((ext_vector_type(4)));
            // transposing read: lane t of a 16-lane group G reads rows
            // 16(G & 1) + 4(t & 3) .. + 3 at k-row 8(G >> 1) + (t >> 2)
            const int G = lane >> 4, t = lane & 15;
            const MBH_LDS unsigned char* ab =
                buf + (8 * (G >> 1) + (t >> 2)) * SAH + 2 * (wm0 + 16 * (G & 1) + 4 * (t & 3));
            const MBH_LDS u32x4* b4 = (const MBH_LDS u32x4*)(buf + L::A_BYTES);
            const int nks = nk >= 32 ? 2 : (nk + 15) >> 4;
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                if (ks >= nks) break;
                bf16x8 ah[TW], bh[TW];
#pragma unroll
                for (int i = 0; i < TW; ++i)
                    if (FULL || i < bi) {
                        const MBH_LDS unsigned char* p = ab + 16 * ks * SAH + 64 * i;
                        const v4i16 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((MBH_LDS v4i16*)p);
                        const v4i16 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((MBH_LDS v4i16*)(p + 4 * SAH));
                        ah[i] = __builtin_bit_cast(bf16x8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
                    }
#pragma unroll
                for (int j = 0; j < TW; ++j)
                    if (FULL || j < bj) {
                        const int row = wn0 + 32 * j + lr;
                        bh[j] = __builtin_bit_cast(bf16x8, (u32x4)b4[row * 4 + ((2 * ks + lh) ^ lds_swz(row))]);
                    }
#pragma unroll
                for (int i = 0; i < TW; ++i)
#pragma unroll
                    for (int j = 0; j < TW; ++j)
                        if (on(i, j)) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
            }
        } else {
            const MBH_LDS float* as = (const MBH_LDS float*)buf + 4 * lh * SAF + wm0 + lr;
            const MBH_LDS float* bs = (const MBH_LDS float*)(buf + L::A_BYTES) + (wn0 + lr) * SBF + 4 * lh;
            auto kblock = [&](int c) {
                f32x4 bq[TW];
#pragma unroll
                for (int j = 0; j < TW; ++j)
                    if (FULL || j < bj) bq[j] = *(const MBH_LDS f32x4*)(bs + 32 * j * SBF + 8 * c);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float af[TW];
#pragma unroll
                    for (int i = 0; i < TW; ++i)
                        if (FULL || i < bi) af[i] = as[(8 * c + e) * SAF + 32 * i];
#pragma unroll
                    for (int i = 0; i < TW; ++i)
#pragma unroll
                        for (int j = 0; j < TW; ++j)
                            if (on(i, j)) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(bq[j][e], af[i], acc[i][j], 0, 0, 0);
                }
            };
            if (nk >= 32) {
#pragma unroll
                for (int c = 0; c < 4; ++c) kblock(c);
            } else {
                for (int c = 0; 8 * c < nk; ++c) kblock(c);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    // C^T accumulators: register r of block (i, j), lane l is
    // C(row0 + 32i + l%32, col0 + 32j + acc_col(r, l)).  Blocks past the
    // tile's edge store nothing; blocks inside it store without masks.
    auto epilogue = [&](const VbLeanItem& d) {
        const int mr = d.m - d.m0 - wm0, nr = d.n - d.n0 - wn0;  // rows / columns left for the wave
#pragma unroll
        for (int i = 0; i < TW; ++i) {
            if (32 * i >= mr) continue;
            const int row = d.m0 + wm0 + 32 * i + lr;
            const bool rok = lr < mr - 32 * i;
#pragma unroll
            for (int j = 0; j < TW; ++j) {
                if (32 * j >= nr) continue;
                const int col0 = d.n0 + wn0 + 32 * j;
                MBH_GLOBAL float* cp = d.C + (row + col0 * d.ldc);
                const bool inside = 32 * i + 32 <= mr && 32 * j + 32 <= nr;
                if (beta == 0.0f) {
                    if (inside) {
#pragma unroll
                        for (int r = 0; r < 16; ++r) cp[acc_col<32>(r, lane) * d.ldc] = alpha * acc[i][j][r];
                    } else {
#pragma unroll
                        for (int r = 0; r < 16; ++r) {
                            const int cc = acc_col<32>(r, lane);
                            if (rok && col0 + cc < d.n) cp[cc * d.ldc] = alpha * acc[i][j][r];
                        }
                    }
                } else {
                    // C in quarters of four, each loaded whole (lanes past
                    // the edge read C(m0, n0) instead) before its results are
                    // formed and stored: one wait per quarter, not per element
                    auto ok = [&](int r) { return inside || (rok && col0 + acc_col<32>(r, lane) < d.n); };
#pragma unroll
                    for (int h = 0; h < 16; h += 4) {
                        float cv[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            cv[r] = *(ok(h + r) ? cp + acc_col<32>(h + r, lane) * d.ldc : d.C + (d.m0 + d.n0 * d.ldc));
#pragma unroll
                        for (int r = 0; r < 4; ++r) cv[r] = alpha * acc[i][j][h + r] + beta * cv[r];
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (ok(h + r)) cp[acc_col<32>(h + r, lane) * d.ldc] = cv[r];
                    }
                }
            }
        }
    };
    MBH_LDS unsigned char* const L0 = (MBH_LDS unsigned char*)lds;
    auto lbuf = [&](int b) { return L0 + b * L::BUF; };
    // Blocks of the wave inside the current item's edge, and the k-tile's
    // MFMAs with the right form.
    auto run = [&](const VbLeanItem& d, const MBH_LDS unsigned char* buf, int nk) {
        const int bi = min(TW, max(0, (d.m - d.m0 - wm0 + 31) >> 5));
        const int bj = min(TW, max(0, (d.n - d.n0 - wn0 + 31) >> 5));
        if (bi == TW && bj == TW)
            compute(buf, nk, std::true_type(), bi, bj);
        else if (bi > 0 && bj > 0)
            compute(buf, nk, std::false_type(), bi, bj);
    };

    // Claims (thread 0): qs = queues drained so far, counting from the own.
    const int home = blockIdx.x % VB_QUEUES;
    int qs = 0;
    auto issue = [&]() { return qs < VB_QUEUES ? atomicAdd(ctr + ((home + qs) % VB_QUEUES) * VB_CTR_STRIDE, 1) : 0; };
    auto resolve = [&](int raw) {
        for (;;) {
            if (qs >= VB_QUEUES) return nitems;
            const int x = (home + qs) % VB_QUEUES;
            if (raw < qo[x + 1] - qo[x]) return qo[x] + raw;
            ++qs;
            raw = issue();
        }
    };
    if (tid <= VB_QUEUES) qo[tid] = qoff[tid];
    __syncthreads();
    if (tid == 0) {
        slot[0] = resolve(issue());
        slot[1] = resolve(issue());
    }
    __syncthreads();
    const int it0 = __builtin_amdgcn_readfirstlane(slot[0]);
    if (it0 >= nitems) return;
    VbLeanItem cur = vb_lean_item<T>(probs, items, it0, A, B, C);
    int sel = 1, buf = 0;
    if (cur.nkt > 0) {
        load(cur, 0);
        store(lbuf(0), cur, cur.K);
    }
    __syncthreads();
    for (;;) {
        // Invariant: the current item's first k-tile is in lbuf(buf).
        const int itn = __builtin_amdgcn_readfirstlane(slot[sel]);
        const bool has_next = itn < nitems;
        int claim = 0;
        if (tid == 0 && has_next) claim = issue();
        VbLeanItem nx = cur;
        if (has_next) nx = vb_lean_item<T>(probs, items, itn, A, B, C);
        const bool pre = has_next && nx.nkt > 0;
        // k-steps but the last, then the last one, which loads the next
        // item's first k-tile instead (k = 0: C := beta C, no k-steps).
        for (int kt = 0; kt + 1 < cur.nkt; ++kt) {
            load(cur, (kt + 1) * 32);
            run(cur, lbuf(buf), 32);
            store(lbuf(buf ^ 1), cur, cur.K - (kt + 1) * 32);
            __syncthreads();
            buf ^= 1;
        }
        if (pre) load(nx, 0);
        if (cur.nkt > 0) run(cur, lbuf(buf), cur.K - (cur.nkt - 1) * 32);
        const int nbuf = cur.nkt > 0 ? buf ^ 1 : buf;
        if (pre) store(lbuf(nbuf), nx, nx.K);
        if (tid == 0 && has_next) slot[sel ^ 1] = resolve(claim);
        __syncthreads();
        buf = nbuf;
        epilogue(cur);
#pragma unroll
        for (int i = 0; i < TW; ++i)
#pragma unroll
            for (int j = 0; j < TW; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
        if (!has_next) break;
        cur = nx;
        sel ^= 1;
    }
}

// ---- warp-specialized tiles -------------------------------------------------
// One record per work item (a T x T tile of one problem), written at exec
// time by vb_ws_fill_kernel from the plan's item list and the caller's
// pointer arrays, so a workgroup reads an item with one 64-byte scalar load.
struct VbWsRec {
    const float* A;
    const float* B;
    float* C;
    int m, n, K, lda, ldb, ldc, m0, n0, nkt, pad;
};
static_assert(sizeof(VbWsRec) == 64, "one 64-byte record per item");
static_assert(offsetof(VbWsRec, C) == 16 && offsetof(VbWsRec, m) == 24 && offsetof(VbWsRec, K) == 32 &&
                  offsetof(VbWsRec, ldc) == 44 && offsetof(VbWsRec, m0) == 48 && offsetof(VbWsRec, nkt) == 56,
              "vb_tile_rec reads the record by dword");

// The part of record j the MFMA waves use (the C tile and the k-steps), by
// scalar loads as in vb_rec below.
struct VbWsTile {
    float* C;
    int m, n, K, ldc, m0, n0, nkt;
};
__device__ inline VbWsTile vb_tile_rec(const VbWsRec* __restrict__ recs, int j) {
    typedef const __attribute__((address_space(4))) unsigned* cu32_t;
    const cu32_t p = (cu32_t)(recs + j);
    VbWsTile t;
    const uint64_t c = (uint64_t)p[4] | ((uint64_t)p[5] << 32);
    t.C = (float*)c;
    t.m = (int)p[6];
    t.n = (int)p[7];
    t.K = (int)p[8];
    t.ldc = (int)p[11];
    t.m0 = (int)p[12];
    t.n0 = (int)p[13];
    t.nkt = (int)p[14];
    return t;
}

// Record j (a wave-uniform index) by a scalar load: the records are written
// by the fill kernel before the tile kernel starts and only read by it, so
// they are read through the constant address space.  (Through the vector
// cache the loads would share the vmcnt counter with the operand loads, and
// moving them to scalar registers would wait for every operand load in
// flight.)
__device__ inline VbWsRec vb_rec(const VbWsRec* __restrict__ recs, int j) {
    typedef const __attribute__((address_space(4))) unsigned* cu32_t;
    const cu32_t p = (cu32_t)(recs + j);
    unsigned w[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) w[i] = p[i];
    VbWsRec r;
    __builtin_memcpy(&r, w, 64);
    return r;
}

__global__ void vb_ws_fill_kernel(const VbProblem* __restrict__ probs, const VbItem* __restrict__ items, int nitems,
                                  const float* const* A, const float* const* B, float* const* C,
                                  VbWsRec* __restrict__ recs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    const VbItem w = items[i];
    const VbProblem pr = probs[w.prob];
    const int T = vb_tile(pr.cls);
    VbWsRec r;
    r.A = A[w.prob];
    r.B = B[w.prob];
    r.C = C[w.prob];
    r.m = (int)pr.m;
    r.n = (int)pr.n;
    r.K = (int)pr.k;
    r.lda = (int)pr.lda;
    r.ldb = (int)pr.ldb;
    r.ldc = (int)pr.ldc;
    r.m0 = (w.tile / pr.tiles_n) * T;
    r.n0 = (w.tile % pr.tiles_n) * T;
    r.nkt = max(1, (int)((pr.k + 31) / 32));
    r.pad = 0;
    recs[i] = r;
}

// T x T tiles on 512 threads: waves 0-3 run the MFMAs (2 x 2 waves of
// T/2 x T/2, the lean tile's LDS layouts and fragment reads) and store C;
// waves 4-7 load the operands, two 32-deep k-tiles ahead in two register
// sets, and stage each k-tile into LDS (rounded to bf16 in BF16 mode).  Both
// roles walk the workgroup's item list from the same records, so the LDS
// k-tiles are all they pass.  The waves that compute never wait for a load,
// and a load has two k-steps to arrive.  Barriers order the two roles:
//   step s: loaders issue the loads of step s + 2 (set s % 2) and stage step
//           s + 1 (set (s + 1) % 2) into LDS buffer (s + 1) % 2; MFMA waves
//           run step s from buffer s % 2; barrier.  At an item's last
//           k-tile the MFMA waves pass a second barrier first and send the C
//           tile out through buffer s % 2 (16-byte stores).
// 128 registers per lane, four waves per SIMD: two workgroups per CU.
// Every step issues the same loads (those a tile does not need read 16 bytes
// of the record array or of A's first elements), so the loader's wait for
// set (s + 1) % 2 is a fixed count that leaves step s + 2's loads in flight.
// A workgroup walks its own list of items, recs[wsoff[b] .. wsoff[b + 1])
// for block b, the list fixed by the plan (longest-processing-time over the
// XCDs, then over each XCD's workgroups); the loaders read each record one
// item ahead.  Every item takes at least one k-step (k = 0: a zero k-tile,
// so C := beta C).
// SRC16 (BF16 only): the operands are stored as bf16 in HBM (16-bit bit
// patterns, any 2-byte alignment), so each load moves eight elements: a
// dword-aligned 16-byte load plus the dword after it, funnel-shifted by 16
// bits when the first element sits in a dword's upper half.  The LDS
// layouts, the MFMA waves and the results are those of BF16 on the fp32
// operands that round to the same bf16 values.
template <bool BF16, int T, bool SRC16 = false>
__global__ __launch_bounds__(512, 4) void vb_gemm_ws_kernel(const VbWsRec* __restrict__ recs,
                                                             const int* __restrict__ wsoff, float alpha,
                                                             float beta) {
    static_assert(BF16 || !SRC16, "16-bit operands feed the bf16 tiles");
    typedef VbLean<BF16, T> L;
    constexpr int NQ = L::NQ, TW = L::TW, SAF = L::SAF, SBF = L::SBF, SAH = L::SAH;
    // A stage holds one k-tile, or at an item's end the MFMA waves' C blocks
    // on their way out (4 KB per wave).
    constexpr int SB = L::BUF > 16384 ? L::BUF : 16384;
    __shared__ __attribute__((aligned(16))) unsigned char lds[2 * SB];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool loader = wave >= 4;
    MBH_LDS unsigned char* const L0 = (MBH_LDS unsigned char*)lds;
    auto lbuf = [&](int b) { return L0 + b * SB; };

    // ---- loader waves ----
    // A (m-contiguous columns): T/4 lanes cover a k-column's T rows, four
    // rows each, so a wave's 16-byte load reads whole 4T-byte column runs
    // (at any 4-byte alignment a run of 512 bytes touches 5 cache lines, a
    // run of 128 bytes 2); load q of a lane is at k = akg + (1024/T) q.
    // B (k-contiguous columns): lane group g8 holds k = 4 g8 .. + 3 of
    // column cb + 32 q (the k-tile's 128-byte run of each column).
    constexpr int ARG = T / 4, AKG = 1024 / T;
    const int ltid = tid & 255;
    const int arg = ltid % ARG, akg = ltid / ARG;  // A: row group, first k
    const int g8 = lane & 7;                       // B: k group
    const int cb = ltid >> 3;                      // B: column within a 32-column load
    // SRC16: eight elements per load.  A: T/8 lanes per k-column (eight rows
    // each), 2048/T k-columns per load; B: lane group g4 holds k = 8 g4 ..
    // + 7 of column cb16 + 64 q.  NQ16 loads per operand, each a 16-byte
    // load and a 4-byte one, so a set issues 2 NQ loads as in fp32.
    constexpr int NQ16 = T / 64, ARG16 = T / 8, AKG16 = 2048 / T;
    const int arg16 = ltid % ARG16, akg16 = ltid / ARG16;
    const int g4 = lane & 3, cb16 = ltid >> 2;
    struct Set {
        float4 va[NQ], vb[NQ];
        int sa[NQ], sb[NQ];
        u32x4 ha[NQ16], hb[NQ16];    // SRC16: dwords 0-3 of each 8-element load
        unsigned ha4[NQ16], hb4[NQ16];  // SRC16: dword 4 (dword 3 again when aligned)
        // whole: every k below K and no group past the operand's end (any
        // k-tile but an item's last, given lda >= 3, or >= 7 for 16-bit
        // operands), so the k-tile is staged without shifts or masks
        int qb, krem, whole, last, end, mr;
    };
    const MBH_GLOBAL float* const nullp = (const MBH_GLOBAL float*)recs;
    // s_waitcnt vmcnt(2 NQ) (expcnt and lgkmcnt at their maxima: no wait)
    constexpr int VMWAIT = ((2 * NQ) & 15) | (((2 * NQ) >> 4) << 14) | (7 << 4) | (15 << 8);
    int j = wsoff[blockIdx.x];
    const int jend = wsoff[blockIdx.x + 1];
    bool done = j >= jend;
    int kt = 0;
    VbWsRec cur{}, nxt{};
    if (loader) {
        if (!done) cur = vb_rec(recs, j);
        if (j + 1 < jend) nxt = vb_rec(recs, j + 1);
    }
    auto nload = [&](int extent) { return min(NQ, (extent + 31) >> 5); };
    // The next k-tile of the list into S: loads issued, not waited for.  S's
    // previous loads completed before S was staged; the wait for all but the
    // other set's 2 NQ loads only says so to the compiler (which would
    // otherwise wait for every load in flight before reusing S's registers).
    auto fetch = [&](Set& S) {
        __builtin_amdgcn_s_waitcnt(VMWAIT);
        if (!done && kt == cur.nkt) {
            ++j;
            if (j >= jend) {
                done = true;
            } else {
                cur = nxt;
                kt = 0;
                if (j + 1 < jend) nxt = vb_rec(recs, j + 1);
            }
        }
        const bool live = !done && cur.K > 0;
        const int k0 = 32 * kt;
        S.end = done;
        S.qb = SRC16 ? min(NQ16, (cur.n - cur.n0 + 63) >> 6) : nload(cur.n - cur.n0);
        S.krem = cur.K - k0;
        S.whole = live && S.krem > 32 && cur.lda >= (SRC16 ? 7 : 3);
        S.last = kt == cur.nkt - 1;
        S.mr = cur.m - cur.m0;
        if (!done) ++kt;
        // last element of each operand (k > 0: at least four elements each)
        const int alast = (cur.K - 1) * cur.lda + cur.m - 1, blast = (cur.n - 1) * cur.ldb + cur.K - 1;
        if constexpr (SRC16) {
            // byte offsets from the operand's dword-aligned base (lanes past
            // the edge read elements 0 .. 7, idle lanes the record array);
            // an 8-element group past the operand's end is clamped to its
            // last eight elements (every operand of a problem with k > 0
            // has at least eight).  sa / sb: element shift | odd-half flag << 3.
            const uintptr_t pa = live ? (uintptr_t)cur.A : (uintptr_t)nullp;
            const uintptr_t pb = live ? (uintptr_t)cur.B : (uintptr_t)nullp;
            const MBH_GLOBAL unsigned char* a4 = (const MBH_GLOBAL unsigned char*)(pa & ~(uintptr_t)3);
            const MBH_GLOBAL unsigned char* b4 = (const MBH_GLOBAL unsigned char*)(pb & ~(uintptr_t)3);
            const unsigned ua = (unsigned)pa & 2u, ub = (unsigned)pb & 2u;
            const bool arow_ok = live && 8 * arg16 < S.mr;
#pragma unroll
            for (int q = 0; q < NQ16; ++q) {
                const int ia = cur.m0 + 8 * arg16 + vb_mul(k0 + akg16 + AKG16 * q, cur.lda), ja = min(ia, alast - 7);
                const unsigned boa = arow_ok ? ua + 2u * (unsigned)ja : ua;
                S.sa[q] = (arow_ok ? ia - ja : 0) | (int)((boa & 2u) << 2);
                S.ha[q] = vb_ld16w(a4, boa & ~3u);
                S.ha4[q] = vb_ld4w(a4, (boa & ~3u) + ((boa & 2u) ? 16u : 12u));
                const bool okb = live && q < S.qb;
                const int ib = k0 + 8 * g4 + vb_mul(min(cur.n0 + cb16 + 64 * q, cur.n - 1), cur.ldb),
                          jb = min(ib, blast - 7);
                const unsigned bob = okb ? ub + 2u * (unsigned)jb : 0u;
                S.sb[q] = (okb ? ib - jb : 0) | (int)((bob & 2u) << 2);
                S.hb[q] = vb_ld16w(okb ? b4 : (const MBH_GLOBAL unsigned char*)nullp, bob & ~3u);
                S.hb4[q] = vb_ld4w(okb ? b4 : (const MBH_GLOBAL unsigned char*)nullp,
                                   (bob & ~3u) + ((bob & 2u) ? 16u : 12u));
            }
            return;
        }
        const int kk = k0 + 4 * g8;
        const bool arow_ok = live && 4 * arg < S.mr;  // lanes past the edge read A[0 .. 3]
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const bool oka = arow_ok, okb = live && q < S.qb;
            const int ia = cur.m0 + 4 * arg + (k0 + akg + AKG * q) * cur.lda, ja = min(ia, alast - 3);
            S.sa[q] = ia - ja;
            S.va[q] = vb_load4(live ? (const MBH_GLOBAL float*)cur.A : nullp, oka ? 4u * (unsigned)ja : 0u);
            const int ib = kk + min(cur.n0 + cb + 32 * q, cur.n - 1) * cur.ldb, jb = min(ib, blast - 3);
            S.sb[q] = ib - jb;
            S.vb[q] = vb_load4(okb ? (const MBH_GLOBAL float*)cur.B : nullp, okb ? 4u * (unsigned)jb : 0u);
        }
    };
    // S's k-tile into LDS buffer b.  S's loads are the older
    // of the two sets in flight: one wait for all but the newer set's 2 NQ
    // loads (s_waitcnt with only vmcnt set), whatever loads the staging then
    // reads, so no later reuse of S's registers waits again.
    auto stage = [&](int b, const Set& S) {
        MBH_LDS unsigned char* buf = lbuf(b);
        __builtin_amdgcn_s_waitcnt(VMWAIT);
        if (SRC16 && !S.end) {
            MBH_LDS unsigned char* arow = buf + akg16 * SAH + 16 * arg16;
            MBH_LDS uint16_t* b16 = (MBH_LDS uint16_t*)(buf + L::A_BYTES);
#pragma unroll
            for (int q = 0; q < NQ16; ++q) {
                u32x4 wa = vb_join8(S.ha[q], S.ha4[q], S.sa[q]), wb = vb_join8(S.hb[q], S.hb4[q], S.sb[q]);
                if (!S.whole) {  // a clamped group, or k past K
                    wa = vb_shift8(wa, S.sa[q] & 7);
                    wb = vb_shift8(wb, S.sb[q] & 7);
                    if (akg16 + AKG16 * q >= S.krem) wa = u32x4{0u, 0u, 0u, 0u};
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int k = 8 * g4 + 2 * i;
                        wb[i] &= (k < S.krem ? 0xffffu : 0u) | (k + 1 < S.krem ? 0xffff0000u : 0u);
                    }
                }
                *(MBH_LDS u32x4*)(arow + AKG16 * SAH * q) = wa;
                if (q < S.qb) {
                    const int jj = cb16 + 64 * q;
                    *(MBH_LDS u32x4*)(b16 + jj * 32 + ((g4 ^ lds_swz(jj)) << 3)) = wb;
                }
            }
        } else if (!S.end) {
            float ra[4 * NQ], rb[4 * NQ];
            if (S.whole) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    ra[4 * q] = S.va[q].x, ra[4 * q + 1] = S.va[q].y, ra[4 * q + 2] = S.va[q].z, ra[4 * q + 3] = S.va[q].w;
                    rb[4 * q] = S.vb[q].x, rb[4 * q + 1] = S.vb[q].y, rb[4 * q + 2] = S.vb[q].z, rb[4 * q + 3] = S.vb[q].w;
                }
            } else {
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    float a4[4], b4[4];
                    vb_shift4(S.va[q], S.sa[q], a4);
                    vb_shift4(S.vb[q], S.sb[q], b4);
                    const bool kva = akg + AKG * q < S.krem;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        ra[4 * q + e] = kva ? a4[e] : 0.0f;
                        rb[4 * q + e] = 4 * g8 + e < S.krem ? b4[e] : 0.0f;
                    }
                }
            }
            if constexpr (BF16) {
                typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
                MBH_LDS unsigned char* arow = buf + akg * SAH + 8 * arg;
                MBH_LDS uint16_t* b16 = (MBH_LDS uint16_t*)(buf + L::A_BYTES);
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    *(MBH_LDS u32x2*)(arow + AKG * SAH * q) = u32x2{bf16x2_rn_bits(ra[4 * q], ra[4 * q + 1]),
                                                                    bf16x2_rn_bits(ra[4 * q + 2], ra[4 * q + 3])};
                    if (q < S.qb) {
                        const int jj = cb + 32 * q, k4 = 4 * g8;
                        *(MBH_LDS u32x2*)(b16 + jj * 32 + (((k4 >> 3) ^ lds_swz(jj)) << 3) + (k4 & 7)) =
                            u32x2{bf16x2_rn_bits(rb[4 * q], rb[4 * q + 1]), bf16x2_rn_bits(rb[4 * q + 2], rb[4 * q + 3])};
                    }
                }
            } else {
                MBH_LDS float* as = (MBH_LDS float*)buf + akg * SAF + 4 * arg;
                MBH_LDS float* bs = (MBH_LDS float*)(buf + L::A_BYTES) + cb * SBF + 4 * g8;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    *(MBH_LDS f32x4*)(as + AKG * SAF * q) = f32x4{ra[4 * q], ra[4 * q + 1], ra[4 * q + 2], ra[4 * q + 3]};
                    if (q < S.qb)
                        *(MBH_LDS f32x4*)(bs + 32 * q * SBF) = f32x4{rb[4 * q], rb[4 * q + 1], rb[4 * q + 2], rb[4 * q + 3]};
                }
            }
        }
    };

    // ---- MFMA waves ----
    // The wave's part of the current item's C tile starts at (wm0, wn0): its
    // quadrant, or for an fp32 128 x 128 tile with at most 64 rows (columns)
    // below the edge a 32-column (32-row) strip, so that the four waves share
    // the tile's blocks evenly (they step together).
    int wm0 = (wave & 1) * (T / 2), wn0 = ((wave >> 1) & 1) * (T / 2);
    const int lr = lane & 31, lh = lane >> 5;
    // MFMAs of one k-tile over the wave's BI x BJ blocks below the tile's
    // edge (compile-time subsets, so no MFMA is predicated); only the MFMA
    // steps holding k < K run (nk: k values of the k-tile below K).
    auto compute = [&](f32x16 (&acc)[TW][TW], const MBH_LDS unsigned char* buf, int nk, auto bi_c, auto bj_c) {
        constexpr int BI = decltype(bi_c)::value, BJ = decltype(bj_c)::value;
        if constexpr (BI == 0 || BJ == 0) {
            return;
        } else if constexpr (BF16) {
            typedef short v4i16 __attribute__((ext_vector_type(4)));
            const int G = lane >> 4, t = lane & 15;
            const MBH_LDS unsigned char* ab =
                buf + (8 * (G >> 1) + (t >> 2)) * SAH + 2 * (wm0 + 16 * (G & 1) + 4 * (t & 3));
            const MBH_LDS u32x4* b4 = (const MBH_LDS u32x4*)(buf + L::A_BYTES);
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                if (ks > 0 && nk <= 16) break;
                bf16x8 ah[BI], bh[BJ];
#pragma unroll
                for (int i = 0; i < BI; ++i) {
                    const MBH_LDS unsigned char* p = ab + 16 * ks * SAH + 64 * i;
                    const v4i16 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((MBH_LDS v4i16*)p);
                    const v4i16 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((MBH_LDS v4i16*)(p + 4 * SAH));
                    ah[i] = __builtin_bit_cast(bf16x8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
                }
#pragma unroll
                for (int jj = 0; jj < BJ; ++jj) {
                    const int row = wn0 + 32 * jj + lr;
                    bh[jj] = __builtin_bit_cast(bf16x8, (u32x4)b4[row * 4 + ((2 * ks + lh) ^ lds_swz(row))]);
                }
#pragma unroll
                for (int i = 0; i < BI; ++i)
#pragma unroll
                    for (int jj = 0; jj < BJ; ++jj)
                        acc[i][jj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[jj], ah[i], acc[i][jj], 0, 0, 0);
            }
        } else {
            // 8-deep k-blocks c (v_mfma_f32_32x32x2f32 over k = 8c + 4h + e,
            // lane half h), the fragments of block c + 1 read before block
            // c's MFMAs (always: rows past K hold zeros), MFMAs only for
            // blocks holding k < K.
            const MBH_LDS float* as = (const MBH_LDS float*)buf + 4 * lh * SAF + wm0 + lr;
            const MBH_LDS float* bs = (const MBH_LDS float*)(buf + L::A_BYTES) + (wn0 + lr) * SBF + 4 * lh;
            const int nc = nk >= 32 ? 4 : (nk + 7) >> 3;
            f32x4 bq[2][BJ];
            float af[2][4][BI];
            auto frag = [&](int c, f32x4 (&q)[BJ], float (&a)[4][BI]) {
#pragma unroll
                for (int jj = 0; jj < BJ; ++jj) q[jj] = *(const MBH_LDS f32x4*)(bs + 32 * jj * SBF + 8 * c);
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int i = 0; i < BI; ++i) a[e][i] = as[(8 * c + e) * SAF + 32 * i];
            };
            frag(0, bq[0], af[0]);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if (c < 3) frag(c + 1, bq[(c + 1) & 1], af[(c + 1) & 1]);
                __builtin_amdgcn_sched_barrier(0);
                if (c < nc) {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
#pragma unroll
                        for (int i = 0; i < BI; ++i)
#pragma unroll
                            for (int jj = 0; jj < BJ; ++jj)
                                acc[i][jj] = __builtin_amdgcn_mfma_f32_32x32x2f32(bq[c & 1][jj][e], af[c & 1][e][i],
                                                                                  acc[i][jj], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    };
    // C^T accumulators: register r of block (i, jj), lane l holds C(wm0 +
    // 32i + l%32, wn0 + 32jj + c(r) + 4(l/32)) from the tile origin Ct, c(r)
    // = 8(r/4) + r%4.  Each block goes out through the wave's 4 KB of the
    // stage buffer the item's last k-tile came from (every MFMA wave is past
    // it): written as [column][32 rows], read back as four rows of one column
    // per lane, and stored 16 bytes at a time (8 lanes per 128-byte column
    // segment); both LDS passes take one base address per lane and constant
    // offsets.  mr /
    // nr: rows / columns left for the wave; blocks past the edge store
    // nothing, partial rows and columns store per element.
    auto epilogue = [&](const f32x16 (&acc)[TW][TW], MBH_LDS unsigned char* buf, MBH_GLOBAL float* Ct, int ldc,
                        int mr, int nr, auto bi_c, auto bj_c, auto beta_c) {
        constexpr int BI = decltype(bi_c)::value, BJ = decltype(bj_c)::value;
        constexpr bool BETA = decltype(beta_c)::value;  // beta != 0: C is read
        MBH_LDS float* reg = (MBH_LDS float*)(buf + 4096 * wave);
        MBH_LDS float* wbase = reg + 128 * lh + lr;  // + 32 c(r): the lane's writes
        const int r4 = 4 * (lane & 7), cl = lane >> 3;
        const MBH_LDS float* rbase = reg + 32 * cl + r4;  // + 256 t: the lane's reads
#pragma unroll
        for (int i = 0; i < BI; ++i) {
#pragma unroll
            for (int jj = 0; jj < BJ; ++jj) {
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    wbase[32 * (8 * (r >> 2) + (r & 3))] = acc[i][jj][r];
                const int rows = mr - 32 * i - r4;  // rows of the lane's four below the edge
                auto ok = [&](int t) { return 32 * jj + 8 * t + cl < nr && rows > 0; };
                auto cp = [&](int t) {
                    return Ct + (wm0 + 32 * i + r4) + (int64_t)(wn0 + 32 * jj + 8 * t + cl) * ldc;
                };
                // beta != 0: the block's C loads all issue before its stores
                f32x4 c[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    c[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                    if (BETA && ok(t)) {
                        if (rows >= 4) {
                            c[t] = vb_ld16(cp(t));
                        } else {
#pragma unroll
                            for (int e = 0; e < 3; ++e)
                                if (e < rows) c[t][e] = cp(t)[e];
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    f32x4 v = *(const MBH_LDS f32x4*)(rbase + 256 * t);
                    v = BETA ? alpha * v + beta * c[t] : alpha * v;
                    if (!ok(t)) continue;
                    if (rows >= 4) {
                        vb_st16(cp(t), v);
                    } else {
#pragma unroll
                        for (int e = 0; e < 3; ++e)
                            if (e < rows) cp(t)[e] = v[e];
                    }
                }
            }
        }
    };
    // One item on the MFMA waves, with the wave's blocks fixed for the item:
    // its k-steps from LDS buffer b on (one barrier each), its C tile after
    // the last; returns the buffer of the next step.  The accumulators live
    // only within an item.
    // The MFMA waves walk the same list as the loaders, from the same
    // records (a scalar load of the C tile's fields per item), so they know
    // each step's k-depth and the C tile without reading anything the
    // loaders write but the k-tiles themselves.
    auto run_item = [&](int b, const VbWsTile& r, auto bi_c, auto bj_c, auto beta_c) {
        f32x16 acc[TW][TW];
#pragma unroll
        for (int i = 0; i < TW; ++i)
#pragma unroll
            for (int jj = 0; jj < TW; ++jj)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[i][jj][e] = 0.0f;
        for (int kt = 0;; ++kt) {
            const int nk = min(32, max(0, r.K - 32 * kt));
            const bool last = kt == r.nkt - 1;
            compute(acc, lbuf(b), nk, bi_c, bj_c);
            if (last) {
                // the C tile's origin and extent, opaque to the compiler here
                // so that the epilogue's per-lane addresses are formed here
                // and not held in registers across the k-loop
                uint64_t cb = (uint64_t)((MBH_GLOBAL float*)r.C + (r.m0 + (int64_t)r.n0 * r.ldc));
                int ldc = r.ldc, mr = r.m - r.m0 - wm0, nr = r.n - r.n0 - wn0;
                asm volatile("" : "+s"(cb), "+s"(ldc), "+s"(mr), "+s"(nr));
                __syncthreads();  // every MFMA wave is done with buffer b
                epilogue(acc, lbuf(b), (MBH_GLOBAL float*)cb, ldc, mr, nr, bi_c, bj_c, beta_c);
            }
            __syncthreads();
            b ^= 1;
            if (last) return b;
        }
    };

    // The two roles run separate loops (so neither holds the other's
    // registers) with the same barriers: one after the prologue and one per
    // step, steps taken in pairs; both leave at the first even step past the
    // end of the list (a list that ends on an odd step takes one empty step).
    // The loader's loop has no exit between its two halves, so the compiler
    // sees each set's loads in flight across exactly one half.
    if (loader) {
        Set S0, S1;
        fetch(S0);
        fetch(S1);
        stage(0, S0);
        // end / last flags of the k-tiles staged into buffers 0 and 1
        int end0 = S0.end, last0 = S0.end ? 0 : S0.last, last1;
        __syncthreads();
        // (steps that end an item take a second barrier: the MFMA waves'
        // C blocks pass through the stage buffer the step read)
        while (!end0) {
            fetch(S0);
            stage(1, S1);
            last1 = S1.end ? 0 : S1.last;
            if (last0) __syncthreads();
            __syncthreads();
            fetch(S1);
            stage(0, S0);
            end0 = S0.end;
            last0 = S0.end ? 0 : S0.last;
            if (last1) __syncthreads();
            __syncthreads();
        }
    } else {
        typedef std::integral_constant<int, 0> zero;
        typedef std::integral_constant<int, 1> one;
        typedef std::integral_constant<int, TW> tw;
        // The walk, in two copies: beta == 0 (C never read) and beta != 0.
        auto walk = [&](auto beta_c) {
            int jm = wsoff[blockIdx.x];
            const int jmend = wsoff[blockIdx.x + 1];
            for (int b = 0;;) {
                if (jm >= jmend) {
                    if (b == 0) break;
                    __syncthreads();  // the empty odd step
                    b = 0;
                    continue;
                }
                const VbWsTile r = vb_tile_rec(recs, jm++);
                // blocks of the wave below the item's edge (bf16: all of
                // them, its MFMAs cost little; those past the edge are never
                // stored)
                const int tmr = r.m - r.m0;
                const int tnr = r.n - r.n0;
                if (!BF16 && TW == 2 && tmr <= 64 && tnr > 64) {
                    wm0 = 0;
                    wn0 = 32 * wave;
                } else if (!BF16 && TW == 2 && tnr <= 64 && tmr > 64) {
                    wm0 = 32 * wave;
                    wn0 = 0;
                } else {
                    wm0 = (wave & 1) * (T / 2);
                    wn0 = ((wave >> 1) & 1) * (T / 2);
                }
                const int strip_r = !BF16 && TW == 2 && tmr <= 64 && tnr > 64 ? 1 : TW;  // block-columns at most
                const int strip_c = !BF16 && TW == 2 && tnr <= 64 && tmr > 64 ? 1 : TW;  // block-rows at most
                const int bi = min(strip_c, max(0, (tmr - wm0 + 31) >> 5));
                const int bj = min(strip_r, max(0, (tnr - wn0 + 31) >> 5));
                if (BF16 || (bi == TW && bj == TW))
                    b = run_item(b, r, tw(), tw(), beta_c);
                else if (bi == 0 || bj == 0)
                    b = run_item(b, r, zero(), zero(), beta_c);
                else if (TW == 2 && bi == 1 && bj == 2)
                    b = run_item(b, r, one(), tw(), beta_c);
                else if (TW == 2 && bi == 2 && bj == 1)
                    b = run_item(b, r, tw(), one(), beta_c);
                else
                    b = run_item(b, r, one(), one(), beta_c);
            }
        };
        __syncthreads();
        if (beta == 0.0f)
            walk(std::false_type());
        else
            walk(std::true_type());
    }
}

// A problem can take the lean tile when its element offsets (with the
// k-tile and tile overhang the tile computes before clamping) fit in int32
// and the byte offsets of the elements it loads fit in 32 bits.
bool vb_lean_ok(const VbProblem& q) {
    const int64_t l31 = (int64_t)1 << 31, l30 = (int64_t)1 << 30;
    return q.m < l30 && q.n < l30 && q.k < l30 && q.lda * (q.k + 128) + q.m + 128 < l31 &&
           q.ldb * (q.n + 128) + q.k + 128 < l31 && q.ldc * (q.n + 128) + q.m + 128 < l31 &&
           q.lda * q.k < l30 && q.ldb * q.n < l30;  // byte offsets of loaded elements fit 32 bits
}

// The warp-specialized tile also needs four elements in each operand of a
// problem with k > 0 (its 16-byte loads are clamped to the operand's last
// four elements).
bool vb_ws_ok(const VbProblem& q) {
    if (!vb_lean_ok(q)) return false;
    if (q.m == 0 || q.n == 0 || q.k == 0) return true;
    return (q.k - 1) * q.lda + q.m >= 4 && (q.n - 1) * q.ldb + q.k >= 4;
}

// MAXCLS: highest tile class compiled in.  The 64x64-only variant needs far
// fewer registers and LDS, so more workgroups stay resident per CU.
template <int MAXCLS>
__global__ __launch_bounds__(256, 2) void vb_gemm_f32_kernel(const VbProblem* __restrict__ probs,
                                                          const VbItem* __restrict__ items, int nitems, int* ctr,
                                                          float alpha, float beta, const float* const* A,
                                                          const float* const* B, float* const* C) {
    constexpr int LDS = MAXCLS == 0 ? FP_LDS64 : FP_LDS;
    __shared__ __attribute__((aligned(16))) unsigned char lds[LDS];
    __shared__ int slot;
    vb_walk(ctr, nitems, &slot, [&](int it) {
        const VbItem w = items[it];
        const VbProblem pr = probs[w.prob];
        GemmParams<float> p{};
        p.m = pr.m;
        p.n = pr.n;
        p.k = pr.k;
        p.A = A[w.prob];
        p.lda = pr.lda;
        p.B = B[w.prob];
        p.ldb = pr.ldb;
        p.C = C[w.prob];
        p.ldc = pr.ldc;
        p.alpha = alpha;
        p.beta = beta;
        p.a_vec = ((uintptr_t)p.A % 16 == 0) && (pr.lda % 4 == 0);
        p.b_vec = ((uintptr_t)p.B % 16 == 0) && (pr.ldb % 4 == 0);
        const int tm = w.tile / pr.tiles_n, tn = w.tile % pr.tiles_n;
        if (MAXCLS == 0 || pr.cls == 0)
            gemm_fp_tile_d2<float, float, float, float, false, false, 64, 64, VB_BK64>(p, tm, tn, lds);
        else
            gemm_fp_tile<float, float, float, float, false, false, 128, 128, VB_BK128>(p, tm, tn, lds);
    });
}

// bf16-input tile: C tile (tm, tn) of problem p, BT x BT, 2 x 2 waves.  The
// fp32 panels of A (m-contiguous) and B (k-contiguous, staged as op(B)^T
// rows) are loaded with bounds checks, rounded to bf16 and written in the
// packed LDS layout of the split-precision kernels (64-byte k-rows, chunk
// index XOR-swizzled by row), whose fragment reads and MFMAs are reused.
template <int BT>
__device__ inline void vb_bf16in_tile(const GemmParams<float>& p, int tm, int tn, u32x4* lds0, u32x4* lds1) {
    constexpr int NT = 256, BK = 32, CH = BT * 4;  // CH: 16-byte chunks per operand per buffer
    constexpr int TM = BT / 64, TN = BT / 64;     // 32x32 blocks per wave
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    PanelLoader<float, float, false, BT, BK, NT> la;  // A(i, k) at A[i + k*lda]
    PanelLoader<float, float, true, BT, BK, NT> lb;   // op(B)^T(j, k) at B[k + j*ldb]
    MBH_LDS u32x4* L[2] = {(MBH_LDS u32x4*)lds0, (MBH_LDS u32x4*)lds1};
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave & 1) * (BT / 2), wn0 = (wave >> 1) * (BT / 2);
    const int64_t m0 = (int64_t)tm * BT, n0 = (int64_t)tn * BT;

    // Registers -> bf16 in LDS.  A: a thread holds 4 consecutive rows at one
    // k (four 2-byte writes); B: 4 consecutive k of one row (one 8-byte write).
    auto stage = [&](MBH_LDS u32x4* buf) {
        MBH_LDS uint16_t* a16 = (MBH_LDS uint16_t*)buf;
        MBH_LDS uint16_t* b16 = (MBH_LDS uint16_t*)(buf + CH);
#pragma unroll
        for (int v = 0; v < la.NV; ++v) {
            const int idx = tid + NT * v;
            const int kk = idx / (BT / 4), row = (idx % (BT / 4)) * 4;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = row + e;
                a16[r * 32 + (((kk >> 3) ^ lds_swz(r)) << 3) + (kk & 7)] = (uint16_t)bf16_rn_bits(la.r[v][e]);
            }
        }
#pragma unroll
        for (int v = 0; v < lb.NV; ++v) {
            const int idx = tid + NT * v;
            const int row = idx / (BK / 4), kk = (idx % (BK / 4)) * 4;
            const u32x2 w = {bf16_rn_bits(lb.r[v][0]) | (bf16_rn_bits(lb.r[v][1]) << 16),
                             bf16_rn_bits(lb.r[v][2]) | (bf16_rn_bits(lb.r[v][3]) << 16)};
            *(MBH_LDS u32x2*)(b16 + row * 32 + (((kk >> 3) ^ lds_swz(row)) << 3) + (kk & 7)) = w;
        }
    };

    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int nkt = (int)((p.k + BK - 1) / BK);
    if (nkt > 0) {
        la.load(p.A, p.lda, m0, p.m, 0, p.k, p.a_vec, tid);
        lb.load(p.B, p.ldb, n0, p.n, 0, p.k, p.b_vec, tid);
        stage(L[0]);
    }
    __syncthreads();
    for (int kt = 0; kt < nkt; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nkt) {
            la.load(p.A, p.lda, m0, p.m, (int64_t)(kt + 1) * BK, p.k, p.a_vec, tid);
            lb.load(p.B, p.ldb, n0, p.n, (int64_t)(kt + 1) * BK, p.k, p.b_vec, tid);
        }
        packed_ktile_mfma<1, TM, TN>(L[cur], L[cur], L[cur] + CH, L[cur] + CH, wm0, wn0, lane, 0, 0, acc);
        if (kt + 1 < nkt) stage(L[cur ^ 1]);
        __syncthreads();
    }
    epilogue_32x32<TM, TN>(acc, p.C, p.ldc, p.m, p.n, m0 + wm0, n0 + wn0, lane, p.alpha, p.beta, TRI_NONE);
}

// Two waves per SIMD: the register cap keeps two workgroups resident per CU.
template <int MAXCLS>
__global__ __launch_bounds__(256, 2) void vb_gemm_bf16_kernel(const VbProblem* __restrict__ probs,
                                                           const VbItem* __restrict__ items, int nitems, int* ctr,
                                                           float alpha, float beta, const float* const* A,
                                                           const float* const* B, float* const* C) {
    constexpr int CH = (MAXCLS == 0 ? 64 : 128) * 4;
    __shared__ u32x4 lds0[2 * CH];
    __shared__ u32x4 lds1[2 * CH];
    __shared__ int slot;
    vb_walk(ctr, nitems, &slot, [&](int it) {
        const VbItem w = items[it];
        const VbProblem pr = probs[w.prob];
        GemmParams<float> p{};
        p.m = pr.m;
        p.n = pr.n;
        p.k = pr.k;
        p.A = A[w.prob];
        p.lda = pr.lda;
        p.B = B[w.prob];
        p.ldb = pr.ldb;
        p.C = C[w.prob];
        p.ldc = pr.ldc;
        p.alpha = alpha;
        p.beta = beta;
        p.a_vec = ((uintptr_t)p.A % 16 == 0) && (pr.lda % 4 == 0);
        p.b_vec = ((uintptr_t)p.B % 16 == 0) && (pr.ldb % 4 == 0);
        const int tm = w.tile / pr.tiles_n, tn = w.tile % pr.tiles_n;
        if (MAXCLS == 0 || pr.cls == 0) vb_bf16in_tile<64>(p, tm, tn, lds0, lds1);
        else vb_bf16in_tile<128>(p, tm, tn, lds0, lds1);
    });
}

// C_p := beta * C_p for every problem (alpha == 0: A and B are not read).
__global__ void vb_scale_kernel(const VbProblem* __restrict__ probs, int count, float beta, float* const* C) {
    for (int p = blockIdx.x; p < count; p += gridDim.x) {
        const VbProblem pr = probs[p];
        float* c = C[p];
        for (int64_t idx = threadIdx.x; idx < pr.m * pr.n; idx += blockDim.x) {
            const int64_t i = idx % pr.m, j = idx / pr.m;
            float* cp = c + i + j * pr.ldc;
            *cp = (beta == 0.0f) ? 0.0f : beta * *cp;
        }
    }
}

// ---- tuning state ------------------------------------------------------------
std::mutex g_tune_mu;
int g_class_table[2][VB_BUCKETS][VB_BUCKETS];
int g_wgs_per_cu[2];
bool g_tune_init = false;

void tune_init_locked() {
    if (g_tune_init) return;
    memcpy(g_class_table, VB_DEFAULT_CLASS, sizeof(g_class_table));
    memcpy(g_wgs_per_cu, VB_DEFAULT_WGS_PER_CU, sizeof(g_wgs_per_cu));
    g_tune_init = true;
}

int prec_slot(int precision) { return precision == MBH_PREC_BF16 ? 1 : 0; }

int bucket_of(int64_t x) {
    // 0: <=16, 1: <=32, 2: <=64, 3: <=128, 4: <=256, 5: larger
    int b = 0;
    int64_t lim = 16;
    while (b < VB_BUCKETS - 1 && x > lim) {
        ++b;
        lim *= 2;
    }
    return b;
}

}  // namespace

struct VbPlan {
    int64_t count = 0;
    int prec = 0;
    int nitems = 0, grid = 0;
    double flops = 0.0;
    VbProblem* probs = nullptr;
    VbItem* items = nullptr;
    int* ctr = nullptr;
    int* qoff = nullptr;  // lean path: per class, VB_QUEUES + 1 queue offsets
    int64_t class_count[2] = {0, 0};
    bool lean = true;  // every problem fits the lean tiles' int32 offsets
    // Lean path: items [0, nitems1) are 128x128 tiles (counters and queue
    // offsets of class 0, grid grid1), the rest 64x64 tiles (class 1, grid).
    int nitems1 = 0, grid1 = 0;
    // Warp-specialized path: items in per-workgroup lists (the 128x128 lists
    // first), wsoff[0 .. grid1] and wsoff[grid1 + 1 .. grid1 + 1 + grid] the
    // absolute list offsets of the two launches, recs one record per item.
    bool ws = false;
    int* wsoff = nullptr;
    VbWsRec* recs = nullptr;
    // First problem with k > 0 and an operand of fewer than eight elements
    // (-1: none); bf16-stored operands need eight.
    int64_t first_short16 = -1;
};

}  // namespace mbh

using namespace mbh;

namespace {

// The warp-specialized launches of one exec: the record fill, then the
// 128x128 and 64x64 classes.
template <bool BF16, bool SRC16>
int vb_ws_exec(const VbPlan* plan, float alpha, const float* const* a, const float* const* b, float beta,
               float* const* c, hipStream_t s) {
    vb_ws_fill_kernel<<<dim3((unsigned)((plan->nitems + 255) / 256)), 256, 0, s>>>(plan->probs, plan->items,
                                                                                   plan->nitems, a, b, c, plan->recs);
    if (plan->grid1 > 0)
        vb_gemm_ws_kernel<BF16, 128, SRC16><<<dim3((unsigned)plan->grid1), 512, 0, s>>>(plan->recs, plan->wsoff,
                                                                                         alpha, beta);
    if (plan->grid > 0)
        vb_gemm_ws_kernel<BF16, 64, SRC16><<<dim3((unsigned)plan->grid), 512, 0, s>>>(
            plan->recs, plan->wsoff + plan->grid1 + 1, alpha, beta);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "vbatched exec");
}

int vb_scale_exec(const VbPlan* plan, float beta, float* const* c, hipStream_t s) {
    if (beta == 1.0f) return 0;
    vb_scale_kernel<<<dim3((unsigned)std::min<int64_t>(plan->count, 65535)), 256, 0, s>>>(plan->probs,
                                                                                           (int)plan->count, beta, c);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "vbatched scale");
}

}  // namespace

extern "C" {

int mbh_vbatch_set_tuning(int precision, const int* class_table, int wgs_per_cu) {
    if (precision != MBH_PREC_FP32 && precision != MBH_PREC_BF16) {
        set_error("mbh_vbatch_set_tuning: precision %d (batched supports FP32 and BF16)", precision);
        return MBH_E_PRECISION;
    }
    if (wgs_per_cu < 0 || wgs_per_cu > 8) {
        set_error("mbh_vbatch_set_tuning: wgs_per_cu %d outside [0, 8]", wgs_per_cu);
        return MBH_E_ARG;
    }
    if (class_table)
        for (int i = 0; i < VB_BUCKETS * VB_BUCKETS; ++i)
            if (class_table[i] < 0 || class_table[i] > 1) {
                set_error("mbh_vbatch_set_tuning: class %d at %d (classes are 0 and 1)", class_table[i], i);
                return MBH_E_ARG;
            }
    std::lock_guard<std::mutex> g(g_tune_mu);
    tune_init_locked();
    const int ps = prec_slot(precision);
    memcpy(g_class_table[ps], class_table ? (const void*)class_table : (const void*)VB_DEFAULT_CLASS[ps],
           sizeof(g_class_table[ps]));
    g_wgs_per_cu[ps] = wgs_per_cu ? wgs_per_cu : VB_DEFAULT_WGS_PER_CU[ps];
    return 0;
}

int mbh_vbatch_plan_create(int64_t count, const int64_t* m, const int64_t* n, const int64_t* k, const int64_t* lda,
                           const int64_t* ldb, const int64_t* ldc, int precision, mbh_vbatch_plan* out) {
    if (!out) {
        set_error("mbh_vbatch_plan_create: null output handle");
        return MBH_E_ARG;
    }
    *out = nullptr;
    if (count < 0 || count > (int64_t)1 << 30) {
        set_error("mbh_vbatch_plan_create: count %lld out of range", (long long)count);
        return MBH_E_ARG;
    }
    if (precision != MBH_PREC_FP32 && precision != MBH_PREC_BF16) {
        set_error("mbh_vbatch_plan_create: precision %d (batched supports FP32 and BF16)", precision);
        return MBH_E_PRECISION;
    }
    if (count > 0 && (!m || !n || !k || !lda || !ldb || !ldc)) {
        set_error("mbh_vbatch_plan_create: null shape array");
        return MBH_E_ARG;
    }
    // Validation with gemm("N","N") semantics per problem (gemm.mojo:179-213).
    for (int64_t p = 0; p < count; ++p) {
        int info = check_gemm('N', 'N', m[p], n[p], k[p], lda[p], ldb[p], ldc[p]);
        if (info) {
            set_error("mbh_vbatch_plan_create: problem %lld fails gemm argument %d", (long long)p, info);
            return MBH_E_ARG;
        }
    }
    int class_table[VB_BUCKETS][VB_BUCKETS];
    int wgs;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        tune_init_locked();
        memcpy(class_table, g_class_table[prec_slot(precision)], sizeof(class_table));
        wgs = g_wgs_per_cu[prec_slot(precision)];
    }
    VbPlan* plan = new VbPlan();
    plan->count = count;
    plan->prec = precision;
    std::vector<VbProblem> probs((size_t)count);
    struct Item {
        int prob, tile;
        double cost;
    };
    int64_t ntiles = 0;
    bool all_ws = true;
    for (int64_t p = 0; p < count; ++p) {
        VbProblem& q = probs[(size_t)p];
        q.m = m[p];
        q.n = n[p];
        q.k = k[p];
        q.lda = lda[p];
        q.ldb = ldb[p];
        q.ldc = ldc[p];
        q.cls = class_table[bucket_of(q.m)][bucket_of(q.n)];
        const int T = vb_tile(q.cls);
        const int64_t tm = (q.m + T - 1) / T, tn = (q.n + T - 1) / T;
        q.tiles_n = (int)tn;
        plan->flops += 2.0 * (double)q.m * (double)q.n * (double)q.k;
        if (q.m == 0 || q.n == 0) continue;
        plan->class_count[q.cls]++;
        if (!vb_lean_ok(q)) plan->lean = false;
        if (!vb_ws_ok(q)) all_ws = false;
        if (plan->first_short16 < 0 && q.k > 0 && ((q.k - 1) * q.lda + q.m < 8 || (q.n - 1) * q.ldb + q.k < 8))
            plan->first_short16 = p;
        ntiles += tm * tn;
    }
    if (ntiles > (int64_t)INT32_MAX) {
        delete plan;
        set_error("mbh_vbatch_plan_create: too many tiles");
        return MBH_E_ARG;
    }
    const bool mixed = plan->class_count[1] > 0;
    const bool use_lean = plan->lean;
    // Cost of one tile of problem p (a k == 0 tile still scales C).
    auto tile_cost = [&](int64_t p) {
        const int T = vb_tile(probs[(size_t)p].cls);
        return (double)T * T * (double)std::max<int64_t>(probs[(size_t)p].k, 1);
    };
    auto tiles_of = [&](int64_t p) {
        const VbProblem& q = probs[(size_t)p];
        const int T = vb_tile(q.cls);
        return ((q.m + T - 1) / T) * ((q.n + T - 1) / T);
    };
    // Lean path: one queue per tile class (128x128 first); otherwise one
    // queue for the two-class kernels.
    std::vector<Item> queues[2];
    for (int64_t p = 0; p < count; ++p)
        if (probs[(size_t)p].m > 0 && probs[(size_t)p].n > 0)
            for (int64_t t = 0; t < tiles_of(p); ++t)
                queues[use_lean && probs[(size_t)p].cls == 0 ? 1 : 0].push_back({(int)p, (int)t, tile_cost(p)});
    // Longest-first (the tiles of a problem stay adjacent): a dynamic queue
    // then finishes with the short items.  The lean path splits each class
    // into VB_QUEUES queues of whole problems, longest-processing-time first
    // onto the least loaded queue (the cost of a tile counts about two
    // k-tiles of fixed work: claim, descriptor, epilogue); plan->qoff holds
    // each queue's first item, relative to its class.
    auto longest_first = [](std::vector<Item>& v) {
        std::stable_sort(v.begin(), v.end(), [](const Item& a, const Item& b) { return a.cost > b.cost; });
    };
    std::vector<VbItem> ditems;
    ditems.reserve((size_t)ntiles);
    std::vector<int> qoff(2 * (VB_QUEUES + 1), 0);
    int dev = 0, ncu = 256;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
    }
    const bool bf = precision == MBH_PREC_BF16;
    // Warp-specialized path unless MBH_VBATCH_KERNEL=lean asks for the
    // lean one (both give the same results; the switch is for comparisons).
    const char* kenv = getenv("MBH_VBATCH_KERNEL");
    plan->ws = use_lean && all_ws && !(kenv && strcmp(kenv, "lean") == 0);
    std::vector<int> wsoff;
    if (plan->ws) {
        // A persistent grid of the workgroups that stay resident (at most
        // wgs per CU); each workgroup gets a fixed list, balanced by a cost
        // per tile in units of one 32-deep k-step's fixed work (barrier,
        // staging): bf16, its k-steps plus one for the C tile (a
        // step's time varies little with the panels it stages: weighting
        // the staged bytes measured slower); fp32, its k-steps plus the
        // busiest MFMA wave's 32 x 32 blocks times the tile's 8-deep
        // k-blocks (four blocks of four a k-step weigh about two units) and
        // a share for the C tile.
        auto ws_tile_cost = [&](int p, int tile) {
            const VbProblem& q = probs[(size_t)p];
            const int T = vb_tile(q.cls);
            const int64_t mr = std::min<int64_t>(T, q.m - (int64_t)(tile / q.tiles_n) * T);
            const int64_t nr = std::min<int64_t>(T, q.n - (int64_t)(tile % q.tiles_n) * T);
            const double steps = (double)std::max<int64_t>(1, (q.k + 31) / 32);
            const double br = (double)((mr + 31) / 32), bc = (double)((nr + 31) / 32);
            if (bf) return steps + 1.0;
            double bmax;
            if (T == 128 && mr <= 64 && nr > 64)
                bmax = br;
            else if (T == 128 && nr <= 64 && mr > 64)
                bmax = bc;
            else
                bmax = std::min(br, T / 64.0) * std::min(bc, T / 64.0);
            return steps + bmax * (double)((q.k + 7) / 8) / 8.0 + bmax / 4.0;
        };
        auto ws_cost = [&](int p) {
            double c = 0.0;
            for (int64_t t = 0; t < tiles_of(p); ++t) c += ws_tile_cost(p, (int)t);
            return c;
        };
        for (int c = 0; c < 2; ++c) {
            const void* kfn = c == 0 ? (bf ? (const void*)vb_gemm_ws_kernel<true, 128> : (const void*)vb_gemm_ws_kernel<false, 128>)
                                     : (bf ? (const void*)vb_gemm_ws_kernel<true, 64> : (const void*)vb_gemm_ws_kernel<false, 64>);
            int resident = 0, w = wgs;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kfn, 512, 0) == hipSuccess && resident > 0)
                w = std::min(w, resident);
            const int G = (int)std::min<int64_t>((int64_t)queues[c].size(), (int64_t)ncu * std::max(w, 1));
            (c == 0 ? plan->grid1 : plan->grid) = G;
            // Problems onto the XCDs (workgroup b runs on XCD b % 8), longest
            // first onto the least loaded; then each XCD's items, longest
            // first, onto its least loaded workgroup.
            const int X = std::min(VB_QUEUES, std::max(G, 1));
            std::vector<std::vector<Item>> xitems(X);
            {
                struct Job {
                    size_t first;
                    int64_t tiles;
                    double cost;
                };
                std::vector<Job> jobs;
                for (size_t i = 0; i < queues[c].size();) {  // the tiles of a problem are consecutive
                    const int p = queues[c][i].prob;
                    const int64_t nt = tiles_of(p);
                    jobs.push_back({i, nt, ws_cost(p)});
                    i += (size_t)nt;
                }
                std::stable_sort(jobs.begin(), jobs.end(), [](const Job& a, const Job& b) { return a.cost > b.cost; });
                std::vector<double> xload((size_t)X, 0.0);
                for (const Job& jb : jobs) {
                    const int x = (int)(std::min_element(xload.begin(), xload.end()) - xload.begin());
                    xload[(size_t)x] += jb.cost;
                    for (int64_t t = 0; t < jb.tiles; ++t) {
                        Item it = queues[c][jb.first + (size_t)t];
                        it.cost = ws_tile_cost(it.prob, it.tile);
                        xitems[(size_t)x].push_back(it);
                    }
                }
            }
            std::vector<std::vector<VbItem>> lists((size_t)G);
            for (int x = 0; x < X && G > 0; ++x) {
                std::vector<int> members;
                for (int b = x; b < G; b += VB_QUEUES) members.push_back(b);
                std::vector<double> wl(members.size(), 0.0);
                std::stable_sort(xitems[(size_t)x].begin(), xitems[(size_t)x].end(),
                                 [](const Item& a, const Item& b) { return a.cost > b.cost; });
                for (const Item& it : xitems[(size_t)x]) {
                    const size_t u = (size_t)(std::min_element(wl.begin(), wl.end()) - wl.begin());
                    wl[u] += it.cost;
                    lists[(size_t)members[u]].push_back({it.prob, it.tile});
                }
            }
            for (int b = 0; b < G; ++b) {
                wsoff.push_back((int)ditems.size());
                for (const VbItem& it : lists[(size_t)b]) ditems.push_back(it);
            }
            wsoff.push_back((int)ditems.size());
        }
    }
    for (int c = 0; c < 2 && !plan->ws; ++c) {
        if (!use_lean) {
            longest_first(queues[c]);
            for (const Item& it : queues[c]) ditems.push_back({it.prob, it.tile});
            continue;
        }
        struct Job {
            int prob;
            int64_t tiles;
            double cost;
        };
        std::vector<Job> jobs;
        for (size_t i = 0; i < queues[c].size();) {  // the tiles of a problem are consecutive
            const int p = queues[c][i].prob;
            const int T = vb_tile(probs[(size_t)p].cls);
            const int64_t nt = tiles_of(p);
            jobs.push_back({p, nt, (double)nt * (queues[c][i].cost + 64.0 * T * T)});
            i += (size_t)nt;
        }
        std::stable_sort(jobs.begin(), jobs.end(), [](const Job& a, const Job& b) { return a.cost > b.cost; });
        std::vector<Item> sub[VB_QUEUES];
        double load[VB_QUEUES] = {};
        for (const Job& j : jobs) {
            const int q = (int)(std::min_element(load, load + VB_QUEUES) - load);
            load[q] += j.cost;
            for (int64_t t = 0; t < j.tiles; ++t) sub[q].push_back({j.prob, (int)t, tile_cost(j.prob)});
        }
        const int base = (int)ditems.size();
        for (int q = 0; q < VB_QUEUES; ++q) {
            qoff[c * (VB_QUEUES + 1) + q] = (int)ditems.size() - base;
            longest_first(sub[q]);
            for (const Item& it : sub[q]) ditems.push_back({it.prob, it.tile});
        }
        qoff[c * (VB_QUEUES + 1) + VB_QUEUES] = (int)ditems.size() - base;
    }
    plan->nitems = (int)ditems.size();
    plan->nitems1 = use_lean ? (int)queues[0].size() : 0;
    // A persistent grid never asks for more workgroups per CU than the kernel
    // it launches can keep resident (late workgroups would only find the
    // queue drained).
    auto grid_for = [&](const void* kfn, int nit) {
        int resident = 0, w = wgs;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kfn, 256, 0) == hipSuccess && resident > 0)
            w = std::min(w, resident);
        return std::min(nit, ncu * w);  // 0 for an empty queue (no launch)
    };
    if (plan->ws) {
        // grids set with the lists
    } else if (use_lean) {
        plan->grid1 = grid_for(bf ? (const void*)vb_gemm_lean_kernel<true, 128> : (const void*)vb_gemm_lean_kernel<false, 128>,
                               plan->nitems1);
        plan->grid = grid_for(bf ? (const void*)vb_gemm_lean_kernel<true, 64> : (const void*)vb_gemm_lean_kernel<false, 64>,
                              plan->nitems - plan->nitems1);
    } else if (bf) {
        plan->grid = grid_for(mixed ? (const void*)vb_gemm_bf16_kernel<1> : (const void*)vb_gemm_bf16_kernel<0>,
                              plan->nitems);
    } else {
        plan->grid = grid_for(mixed ? (const void*)vb_gemm_f32_kernel<1> : (const void*)vb_gemm_f32_kernel<0>,
                              plan->nitems);
    }
    auto fail = [&](hipError_t err, const char* what) {
        mbh_vbatch_plan_destroy((mbh_vbatch_plan)plan);
        return hip_fail(err, what);
    };
    if (inject_alloc_failure()) return fail(hipErrorOutOfMemory, "vbatch plan: counters");
    if ((e = hipMalloc(&plan->ctr, VB_COUNTERS * VB_CTR_STRIDE * sizeof(int))) != hipSuccess)
        return fail(e, "vbatch plan: counters");
    if ((e = hipMalloc(&plan->qoff, qoff.size() * sizeof(int))) != hipSuccess) return fail(e, "vbatch plan: queues");
    if ((e = hipMemcpy(plan->qoff, qoff.data(), qoff.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "vbatch plan: queues copy");
    if (count > 0) {
        if ((e = hipMalloc(&plan->probs, sizeof(VbProblem) * probs.size())) != hipSuccess)
            return fail(e, "vbatch plan: problems");
        if ((e = hipMemcpy(plan->probs, probs.data(), sizeof(VbProblem) * probs.size(), hipMemcpyHostToDevice)) !=
            hipSuccess)
            return fail(e, "vbatch plan: problems copy");
    }
    if (!ditems.empty()) {
        if ((e = hipMalloc(&plan->items, sizeof(VbItem) * ditems.size())) != hipSuccess)
            return fail(e, "vbatch plan: items");
        if ((e = hipMemcpy(plan->items, ditems.data(), sizeof(VbItem) * ditems.size(), hipMemcpyHostToDevice)) !=
            hipSuccess)
            return fail(e, "vbatch plan: items copy");
    }
    if (plan->ws) {
        if ((e = hipMalloc(&plan->wsoff, wsoff.size() * sizeof(int))) != hipSuccess)
            return fail(e, "vbatch plan: workgroup lists");
        if ((e = hipMemcpy(plan->wsoff, wsoff.data(), wsoff.size() * sizeof(int), hipMemcpyHostToDevice)) !=
            hipSuccess)
            return fail(e, "vbatch plan: workgroup lists copy");
        if (!ditems.empty() && (e = hipMalloc(&plan->recs, sizeof(VbWsRec) * ditems.size())) != hipSuccess)
            return fail(e, "vbatch plan: item records");
    }
    *out = (mbh_vbatch_plan)plan;
    return 0;
}

int mbh_vbatch_plan_info(mbh_vbatch_plan handle, double* info) {
    const VbPlan* plan = (const VbPlan*)handle;
    if (!plan || !info) {
        set_error("mbh_vbatch_plan_info: null argument");
        return MBH_E_ARG;
    }
    info[0] = (double)plan->count;
    info[1] = (double)plan->nitems;
    info[2] = (double)(plan->grid + plan->grid1);  // workgroups launched per exec
    info[3] = plan->flops;
    info[4] = (double)plan->class_count[0];
    info[5] = (double)plan->class_count[1];
    // 1: warp-specialized tiles on fixed lists, 0: work queues, 2: the
    // two-class kernels with 64-bit offsets
    info[6] = plan->ws ? 1.0 : plan->lean ? 0.0 : 2.0;
    info[7] = (double)plan->prec;
    return 0;
}

int mbh_bgemm_vbatched_exec(mbh_vbatch_plan handle, float alpha, const void* const* a, const void* const* b,
                            float beta, float* const* c, void* stream) {
    const VbPlan* plan = (const VbPlan*)handle;
    if (!plan) {
        set_error("mbh_bgemm_vbatched_exec: null plan");
        return MBH_E_ARG;
    }
    if (plan->prec != MBH_PREC_BF16) {
        set_error("mbh_bgemm_vbatched_exec: the plan's precision is %d (bf16 operands need a BF16 plan)", plan->prec);
        return MBH_E_PRECISION;
    }
    if (plan->count == 0 || plan->nitems == 0) return 0;
    if (!plan->ws) {
        set_error("mbh_bgemm_vbatched_exec: the plan runs the work-queue kernels (an operand under four elements, "
                  "offsets past 32 bits or MBH_VBATCH_KERNEL=lean), which read fp32 operands only");
        return MBH_E_ARG;
    }
    if (plan->first_short16 >= 0) {
        set_error("mbh_bgemm_vbatched_exec: problem %lld has an operand of fewer than 8 elements",
                  (long long)plan->first_short16);
        return MBH_E_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    if (alpha == 0.0f) return vb_scale_exec(plan, beta, c, s);
    return vb_ws_exec<true, true>(plan, alpha, (const float* const*)a, (const float* const*)b, beta, c, s);
}

int mbh_sgemm_vbatched_exec(mbh_vbatch_plan handle, float alpha, const float* const* a, const float* const* b,
                            float beta, float* const* c, void* stream) {
    const VbPlan* plan = (const VbPlan*)handle;
    if (!plan) {
        set_error("mbh_sgemm_vbatched_exec: null plan");
        return MBH_E_ARG;
    }
    if (plan->count == 0 || plan->nitems == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (alpha == 0.0f) return vb_scale_exec(plan, beta, c, s);
    const dim3 grid((unsigned)plan->grid), block(256);
    const bool mixed = plan->class_count[1] > 0;
    if (plan->ws) {
        return plan->prec == MBH_PREC_BF16 ? vb_ws_exec<true, false>(plan, alpha, a, b, beta, c, s)
                                           : vb_ws_exec<false, false>(plan, alpha, a, b, beta, c, s);
    }
    vb_reset_kernel<<<1, 64, 0, s>>>(plan->ctr);
    if (plan->lean) {
        // 128x128 items first (the longer ones), then the 64x64 queue.
        const bool bf = plan->prec == MBH_PREC_BF16;
        const int n1 = plan->nitems1, n0 = plan->nitems - plan->nitems1;
        int* ctr1 = plan->ctr + VB_QUEUES * VB_CTR_STRIDE;
        const int* qoff1 = plan->qoff + VB_QUEUES + 1;
        if (n1 > 0) {
            if (bf)
                vb_gemm_lean_kernel<true, 128><<<dim3((unsigned)plan->grid1), block, 0, s>>>(
                    plan->probs, plan->items, n1, plan->ctr, plan->qoff, alpha, beta, a, b, c);
            else
                vb_gemm_lean_kernel<false, 128><<<dim3((unsigned)plan->grid1), block, 0, s>>>(
                    plan->probs, plan->items, n1, plan->ctr, plan->qoff, alpha, beta, a, b, c);
        }
        if (n0 > 0) {
            if (bf)
                vb_gemm_lean_kernel<true, 64><<<grid, block, 0, s>>>(plan->probs, plan->items + n1, n0, ctr1, qoff1,
                                                                     alpha, beta, a, b, c);
            else
                vb_gemm_lean_kernel<false, 64><<<grid, block, 0, s>>>(plan->probs, plan->items + n1, n0, ctr1, qoff1,
                                                                      alpha, beta, a, b, c);
        }
    } else if (plan->prec == MBH_PREC_BF16) {
        if (mixed)
            vb_gemm_bf16_kernel<1><<<grid, block, 0, s>>>(plan->probs, plan->items, plan->nitems, plan->ctr, alpha,
                                                          beta, a, b, c);
        else
            vb_gemm_bf16_kernel<0><<<grid, block, 0, s>>>(plan->probs, plan->items, plan->nitems, plan->ctr, alpha,
                                                          beta, a, b, c);
    } else {
        if (mixed)
            vb_gemm_f32_kernel<1><<<grid, block, 0, s>>>(plan->probs, plan->items, plan->nitems, plan->ctr, alpha,
                                                         beta, a, b, c);
        else
            vb_gemm_f32_kernel<0><<<grid, block, 0, s>>>(plan->probs, plan->items, plan->nitems, plan->ctr, alpha,
                                                         beta, a, b, c);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "vbatched exec");
}

int mbh_vbatch_plan_destroy(mbh_vbatch_plan handle) {
    VbPlan* plan = (VbPlan*)handle;
    if (!plan) return 0;
    // The plan's buffers may still be read by queued work: wait for the device.
    hipError_t e = hipDeviceSynchronize();
    if (plan->probs) (void)hipFree(plan->probs);
    if (plan->items) (void)hipFree(plan->items);
    if (plan->ctr) (void)hipFree(plan->ctr);
    if (plan->qoff) (void)hipFree(plan->qoff);
    if (plan->wsoff) (void)hipFree(plan->wsoff);
    if (plan->recs) (void)hipFree(plan->recs);
    delete plan;
    return e == hipSuccess ? 0 : hip_fail(e, "mbh_vbatch_plan_destroy");
}

int mbh_sgemm_vbatched_dev(int64_t count, const int64_t* m, const int64_t* n, const int64_t* k, float alpha,
                           const float* const* a, const int64_t* lda, const float* const* b, const int64_t* ldb,
                           float beta, float* const* c, const int64_t* ldc, int precision, void* stream) {
    mbh_vbatch_plan plan = nullptr;
    int rc = mbh_vbatch_plan_create(count, m, n, k, lda, ldb, ldc, precision, &plan);
    if (rc) return rc;
    rc = mbh_sgemm_vbatched_exec(plan, alpha, a, b, beta, c, stream);
    int rc2 = mbh_vbatch_plan_destroy(plan);
    return rc ? rc : rc2;
}

}  // extern "C"
