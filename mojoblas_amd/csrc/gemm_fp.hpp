// Plain FP32 / FP64 macro-tile gemm on the gfx950 matrix cores.
//
//   C := alpha*op(A)*op(B) + beta*C, column-major, any lda/ldb/ldc,
//   op = N or T per operand (template AT / BT), edge tiles zero-filled.
//
// Layout in LDS: each operand panel is stored k-major, [BK][rows + pad], so an
// MFMA operand fetch is one ds_read_b32/b64 per lane with the 32 (or 16)
// lanes of a row group reading consecutive addresses.  Global loads are
// 16-byte vectors along whichever index is contiguous in HBM: along m/n for
// A 'N' / B 'T' (stored straight into the LDS row), along k for A 'T' /
// B 'N' (scattered into 4 LDS rows; the row pad is chosen so those 4-byte
// writes from one half-wave hit 32 distinct banks).
//
// MFMA: fp32 uses v_mfma_f32_32x32x2_f32 (32x32 output, 16 regs/lane),
// fp64 uses v_mfma_f64_16x16x4_f64 (16x16 output, 4 regs/lane).  The operands
// are passed swapped (B fragment as the instruction's A) so the instruction
// computes the C^T block: each accumulator register then holds 32 (16)
// consecutive ROWS of one C column across lanes, which makes the column-major
// epilogue store coalesced.
//
// Pipeline: two LDS buffers; the next k-tile's global loads are issued into
// registers before the current tile's MFMAs and written to the other buffer
// after them, one barrier per k-tile.
#pragma once
#include "common.hpp"

namespace mbh {

template <typename T>
struct MfmaFp;

template <>
struct MfmaFp<float> {
    static constexpr int MT = 32;   // output tile edge
    static constexpr int KPI = 2;   // k per instruction
    static constexpr int NACC = 16; // accumulator registers per lane
    typedef f32x16 acc_t;
    __device__ static inline acc_t mma(float x, float y, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(x, y, c, 0, 0, 0);
    }
};

template <>
struct MfmaFp<double> {
    static constexpr int MT = 16;
    static constexpr int KPI = 4;
    static constexpr int NACC = 4;
    typedef f64x4 acc_t;
    __device__ static inline acc_t mma(double x, double y, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
    }
};

// Row of the MFMA output held in register r of lane `lane` (with the swapped
// operands this is the C column offset; the C row offset is lane % MT).
//   v_mfma_f32_32x32x*  : 8*(r/4) + 4*(lane/32) + r%4
//   v_mfma_f64_16x16x4  : 4*r + lane/16   (measured on gfx950: the f64 form
//                         interleaves the four lane groups row by row)
template <int MT>
__device__ inline int acc_col(int r, int lane) {
    if (MT == 32) return (r >> 2) * 8 + 4 * (lane >> 5) + (r & 3);
    return 4 * r + (lane >> 4);
}

// Global -> register staging of one operand panel (ROWS x BK of op(X)) in
// its storage type TIN; LDS holds TIN as well and fragment reads convert to
// the MFMA type T (an fp32 operand of an fp64 product then costs half the
// LDS bytes of a converted one).
// KCONTIG: element (row, kk) lives at src[kk + row*ld]; else src[row + kk*ld].
template <typename T, typename TIN, bool KCONTIG, int ROWS, int BK, int NT>
struct PanelLoader {
    static constexpr int VEC = 16 / sizeof(TIN);  // elements per 16-byte load
    static constexpr int NV = ROWS * BK / VEC / NT;
    static_assert(NV * VEC * NT == ROWS * BK, "panel not divisible by the thread block");
    typedef TIN vin_t __attribute__((ext_vector_type(VEC)));
    TIN r[NV][VEC];

    __device__ inline void load(const TIN* __restrict__ src, int64_t ld, int64_t row0, int64_t nrows, int64_t k0,
                                int64_t ktot, bool vec_ok, int tid) {
        const bool full = vec_ok && (row0 + ROWS <= nrows) && (k0 + BK <= ktot);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int idx = tid + NT * v;
            int row, kk;
            if (KCONTIG) {
                row = idx / (BK / VEC);
                kk = (idx % (BK / VEC)) * VEC;
            } else {
                kk = idx / (ROWS / VEC);
                row = (idx % (ROWS / VEC)) * VEC;
            }
            if (full) {
                const TIN* p = KCONTIG ? src + (k0 + kk) + (row0 + row) * ld : src + (row0 + row) + (k0 + kk) * ld;
                const vin_t x = *(const MBH_GLOBAL vin_t*)p;
#pragma unroll
                for (int e = 0; e < VEC; ++e) r[v][e] = x[e];
            } else {
                // Branch-free masked loads: a masked element reads src[0]
                // (the panel has at least one element whenever it loads) and
                // is replaced by zero, so the loads issue back to back
                // (guarded loads each became a branch and were serialized).
                const MBH_GLOBAL TIN* g = (const MBH_GLOBAL TIN*)src;
                bool ok[VEC];
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                    const int64_t gr = row0 + row + (KCONTIG ? 0 : e);
                    const int64_t gk = k0 + kk + (KCONTIG ? e : 0);
                    ok[e] = gr < nrows && gk < ktot;
                    r[v][e] = g[ok[e] ? (KCONTIG ? gk + gr * ld : gr + gk * ld) : 0];
                }
#pragma unroll
                for (int e = 0; e < VEC; ++e) r[v][e] = ok[e] ? r[v][e] : TIN(0);
            }
        }
    }

    // Register -> LDS ([BK][S] layout).
    template <int S>
    __device__ inline void store(MBH_LDS TIN* lds, int tid) const {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int idx = tid + NT * v;
            if (KCONTIG) {
                const int row = idx / (BK / VEC);
                const int kk = (idx % (BK / VEC)) * VEC;
#pragma unroll
                for (int e = 0; e < VEC; ++e) lds[(kk + e) * S + row] = r[v][e];
            } else {
                const int kk = idx / (ROWS / VEC);
                const int row = (idx % (ROWS / VEC)) * VEC;
                vin_t x;
#pragma unroll
                for (int e = 0; e < VEC; ++e) x[e] = r[v][e];
                *reinterpret_cast<MBH_LDS vin_t*>(lds + kk * S + row) = x;
            }
        }
    }
};

// LDS row stride (elements of the storage type TIN) for a panel: rows + the
// vector width when rows are written as vectors; rows + 2 (4-byte) / rows + 1
// (8-byte elements) when written as scattered scalars (4 consecutive kk from
// 8 rows land in distinct banks).
template <typename T, typename TIN, bool KCONTIG, int ROWS>
constexpr int lds_stride() {
    return KCONTIG ? ROWS + (sizeof(TIN) == 4 ? 2 : 1) : ROWS + 16 / (int)sizeof(TIN);
}

// LDS bytes one tile's double-buffered panels need: [BK][SA] of TA and
// [BK][SB] of TB per buffer (every region a multiple of 16 bytes).
template <typename T, typename TA, typename TB, bool AT, bool BT, int BM, int BN, int BK>
struct FpTileLds {
    static constexpr int SA = lds_stride<T, TA, AT, BM>();
    static constexpr int SB = lds_stride<T, TB, !BT, BN>();
    static constexpr int A_BYTES = BK * SA * (int)sizeof(TA), B_BYTES = BK * SB * (int)sizeof(TB);
    static_assert(A_BYTES % 16 == 0 && B_BYTES % 16 == 0, "LDS regions must stay 16-byte aligned");
    static constexpr int BYTES = 2 * (A_BYTES + B_BYTES);
};

// One BM x BN tile (tm, tn) of C := alpha*op(A)*op(B) + beta*C on 256
// threads (2 x 2 waves); `lds` holds FpTileLds::BYTES, 16-byte aligned.
// Shared by the grid kernel below and the persistent batched one.
template <typename T, typename TA, typename TB, typename TC, bool AT, bool BT, int BM, int BN, int BK>
__device__ inline void gemm_fp_tile(const GemmArgs<TA, TB, TC, T>& p, int tm, int tn, void* lds) {
    typedef MfmaFp<T> M;
    typedef FpTileLds<T, TA, TB, AT, BT, BM, BN, BK> L;
    constexpr int NT = 256;
    constexpr int WM = BM / 2, WN = BN / 2;  // 2 x 2 waves
    constexpr int TM = WM / M::MT, TN = WN / M::MT;
    constexpr bool AK = AT;    // A 'T': op(A)(i,k) = A[k + i*lda] is k-contiguous
    constexpr bool BK_ = !BT;  // B 'N': op(B)(k,j) = B[k + j*ldb] is k-contiguous
    constexpr int SA = L::SA, SB = L::SB;
    MBH_LDS TA* const la3 = (MBH_LDS TA*)lds;
    MBH_LDS TB* const lb3 = (MBH_LDS TB*)((MBH_LDS char*)lds + 2 * L::A_BYTES);
    MBH_LDS TA* As[2] = {la3, la3 + BK * SA};
    MBH_LDS TB* Bs[2] = {lb3, lb3 + BK * SB};

    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;

    PanelLoader<T, TA, AK, BM, BK, NT> la;
    PanelLoader<T, TB, BK_, BN, BK, NT> lb;
    typename M::acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < M::NACC; ++r) acc[i][j][r] = T(0);

    int kt0, kt1;
    kt_range(p.krange, tm, tn, BM, BN, BK, (int)((p.k + BK - 1) / BK), kt0, kt1);
    if (kt0 < kt1) {
        la.load(p.A, p.lda, m0, p.m, (int64_t)kt0 * BK, p.k, p.a_vec, tid);
        lb.load(p.B, p.ldb, n0, p.n, (int64_t)kt0 * BK, p.k, p.b_vec, tid);
        la.template store<SA>(As[0], tid);
        lb.template store<SB>(Bs[0], tid);
    }
    __syncthreads();

    constexpr int G = 64 / M::MT;  // lane groups = k values per instruction
    const int lrow = lane % M::MT, lgrp = lane / M::MT;
    for (int kt = kt0; kt < kt1; ++kt) {
        const int cur = (kt - kt0) & 1;
        if (kt + 1 < kt1) {
            la.load(p.A, p.lda, m0, p.m, (int64_t)(kt + 1) * BK, p.k, p.a_vec, tid);
            lb.load(p.B, p.ldb, n0, p.n, (int64_t)(kt + 1) * BK, p.k, p.b_vec, tid);
        }
        const MBH_LDS TA* Ab = As[cur] + lgrp * SA + wm0 + lrow;
        const MBH_LDS TB* Bb = Bs[cur] + lgrp * SB + wn0 + lrow;
#pragma unroll
        for (int kk = 0; kk < BK; kk += G) {
            T a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = (T)Ab[kk * SA + i * M::MT];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = (T)Bb[kk * SB + j * M::MT];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = M::mma(b[j], a[i], acc[i][j]);
        }
        if (kt + 1 < kt1) {
            la.template store<SA>(As[cur ^ 1], tid);
            lb.template store<SB>(Bs[cur ^ 1], tid);
        }
        __syncthreads();
    }

    // Epilogue: register r of tile (i,j) holds C(row, col) with
    // row = m0 + wm0 + i*MT + lane%MT, col = n0 + wn0 + j*MT + acc_col(r, lane).
    const T alpha = p.alpha, beta = p.beta;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int64_t row = m0 + wm0 + i * M::MT + lrow;
        if (row >= p.m) continue;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
#pragma unroll
            for (int r = 0; r < M::NACC; ++r) {
                const int64_t col = n0 + wn0 + j * M::MT + acc_col<M::MT>(r, lane);
                if (col < p.n && tri_keep(p.tri, row, col)) {
                    MBH_GLOBAL TC* cp = (MBH_GLOBAL TC*)p.C + row + col * p.ldc;
                    T v = alpha * acc[i][j][r];
                    if (beta != T(0)) v += beta * (T)*cp;
                    *cp = (TC)v;
                }
            }
        }
    }
}


// The same tile with register prefetch distance two: the k-loop is unrolled
// by two with two named staging sets, so the global loads of k-tile kt+2 are
// issued while tile kt is computed and have two k-tiles of compute to land
// (tile kt+1 moves from registers to LDS after tile kt's MFMAs).  Used where
// k-loops are short and latency dominates (the persistent batched kernel).
template <typename T, typename TA, typename TB, typename TC, bool AT, bool BT, int BM, int BN, int BK>
__device__ inline void gemm_fp_tile_d2(const GemmArgs<TA, TB, TC, T>& p, int tm, int tn, void* lds) {
    typedef MfmaFp<T> M;
    typedef FpTileLds<T, TA, TB, AT, BT, BM, BN, BK> L;
    constexpr int NT = 256;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int TM = WM / M::MT, TN = WN / M::MT;
    constexpr bool AK = AT, BK_ = !BT;
    constexpr int SA = L::SA, SB = L::SB;
    MBH_LDS TA* const A0 = (MBH_LDS TA*)lds;
    MBH_LDS TA* const A1 = A0 + BK * SA;
    MBH_LDS TB* const B0 = (MBH_LDS TB*)((MBH_LDS char*)lds + 2 * L::A_BYTES);
    MBH_LDS TB* const B1 = B0 + BK * SB;

    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave & 1) * WM, wn0 = (wave >> 1) * WN;
    constexpr int G = 64 / M::MT;
    const int lrow = lane % M::MT, lgrp = lane / M::MT;

    PanelLoader<T, TA, AK, BM, BK, NT> la0, la1;
    PanelLoader<T, TB, BK_, BN, BK, NT> lb0, lb1;
    typename M::acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < M::NACC; ++r) acc[i][j][r] = T(0);

    auto mfma_tile = [&](const MBH_LDS TA* As, const MBH_LDS TB* Bs) {
        const MBH_LDS TA* Ab = As + lgrp * SA + wm0 + lrow;
        const MBH_LDS TB* Bb = Bs + lgrp * SB + wn0 + lrow;
#pragma unroll
        for (int kk = 0; kk < BK; kk += G) {
            T a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = (T)Ab[kk * SA + i * M::MT];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = (T)Bb[kk * SB + j * M::MT];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = M::mma(b[j], a[i], acc[i][j]);
        }
    };
#define MBH_LD(LA, LB_, kt)                                                            \
    do {                                                                              \
        LA.load(p.A, p.lda, m0, p.m, (int64_t)(kt) * BK, p.k, p.a_vec, tid);         \
        LB_.load(p.B, p.ldb, n0, p.n, (int64_t)(kt) * BK, p.k, p.b_vec, tid);        \
    } while (0)

    int kt0, kt1;
    kt_range(p.krange, tm, tn, BM, BN, BK, (int)((p.k + BK - 1) / BK), kt0, kt1);
    if (kt0 < kt1) {
        MBH_LD(la0, lb0, kt0);
        if (kt0 + 1 < kt1) MBH_LD(la1, lb1, kt0 + 1);
        la0.template store<SA>(A0, tid);
        lb0.template store<SB>(B0, tid);
    }
    __syncthreads();
    int kt = kt0;
    for (; kt + 1 < kt1; kt += 2) {
        // LDS 0: tile kt; registers 1: tile kt+1 (in flight)
        if (kt + 2 < kt1) MBH_LD(la0, lb0, kt + 2);
        mfma_tile(A0, B0);
        la1.template store<SA>(A1, tid);
        lb1.template store<SB>(B1, tid);
        __syncthreads();
        // LDS 1: tile kt+1; registers 0: tile kt+2 (in flight)
        if (kt + 3 < kt1) MBH_LD(la1, lb1, kt + 3);
        mfma_tile(A1, B1);
        if (kt + 2 < kt1) {
            la0.template store<SA>(A0, tid);
            lb0.template store<SB>(B0, tid);
        }
        __syncthreads();
    }
    if (kt < kt1) mfma_tile(A0, B0);
#undef MBH_LD

    const T alpha = p.alpha, beta = p.beta;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int64_t row = m0 + wm0 + i * M::MT + lrow;
        if (row >= p.m) continue;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
#pragma unroll
            for (int r = 0; r < M::NACC; ++r) {
                const int64_t col = n0 + wn0 + j * M::MT + acc_col<M::MT>(r, lane);
                if (col < p.n && tri_keep(p.tri, row, col)) {
                    MBH_GLOBAL TC* cp = (MBH_GLOBAL TC*)p.C + row + col * p.ldc;
                    T v = alpha * acc[i][j][r];
                    if (beta != T(0)) v += beta * (T)*cp;
                    *cp = (TC)v;
                }
            }
        }
    }
}

template <typename T, typename TA, typename TB, typename TC, bool AT, bool BT, int BM, int BN, int BK>
__global__ __launch_bounds__(256) void gemm_fp_kernel(GemmArgs<TA, TB, TC, T> p) {
    typedef FpTileLds<T, TA, TB, AT, BT, BM, BN, BK> L;
    __shared__ __attribute__((aligned(16))) unsigned char lds[L::BYTES];
    int tm, tn;
    if (p.tri_tiles) tri_tile_coords(p.tri_tiles, xcd_remap((int)blockIdx.x, (int)gridDim.x), p.tiles_m, tm, tn);
    else tile_coords_kr(p.krange, p.tiles_m, p.tiles_n, tm, tn);
    gemm_fp_tile<T, TA, TB, TC, AT, BT, BM, BN, BK>(p, tm, tn, lds);
}


// Interior tile with direct-to-LDS panel loads (no register staging), fp64
// or fp32, for a C of whole tiles (or the block of whole tiles of a plain
// product): 128 x 128 x BK, 2 x 2 waves of 64 x 64 (4 x 4 blocks of
// v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32, whose operands share one
// layout: lane l holds (row l%16, k l/16)), two LDS buffers, one vmcnt wait
// and one barrier per k-tile.  BK = 128 B / sizeof(T) (16 doubles, 32
// floats), so a k-contiguous panel row is one 128-byte line.  Only for
// interior tiles: operands 16-byte aligned with ld a multiple of the 16-byte
// vector, k a multiple of BK, the tile inside m x n (no masking at all).
// Each operand panel is copied by global_load_lds_dwordx4 (a wave-load moves
// 64 16-byte chunks to 1 KB of consecutive LDS) in one of two forms:
//   MN (the panel's rows are contiguous in HBM: op(A) for A 'N', op(B)^T for
//     B 'T'): k-major, one wave-load per KPW = 1 (fp64) or 2 (fp32) k rows of
//     128 elements, the wave-load groups 1152 B apart; lane l of a fragment
//     read takes (k = l/16, row = l%16): 16 consecutive elements per k.
//   K (k contiguous: op(A) for A 'T', op(B)^T for B 'N'): [row][BK], chunk c
//     of row r stored in slot c ^ (r % 8), one wave-load per 8 rows (each
//     lane fetches the global chunk that belongs in its slot), so the 16 rows
//     of a fragment read spread over distinct slots.
template <typename T, bool MN>
struct DmaPanel {
    static constexpr bool ROWS_CONTIGUOUS = MN;
    static constexpr int CE = 16 / (int)sizeof(T);        // elements per chunk
    static constexpr int BK = 8 * CE;                       // k per tile (128-byte rows)
    static constexpr int KPW = 1024 / (128 * (int)sizeof(T));  // MN: k rows per wave-load
    static constexpr int LPR = 64 / KPW;                    // MN: lanes per k row
    static constexpr int GS = 1152 / (int)sizeof(T);        // MN: wave-load group stride
    static constexpr int ELEMS = MN ? (BK / KPW) * GS : 128 * BK;
    // Global source of this lane for wave-load w of the k-tile at k0 (the
    // panel's first row r0); w in [0, 16).
    __device__ static inline const T* src(const T* X, int64_t ld, int64_t r0, int64_t k0, int w, int lane) {
        if (MN) return X + r0 + CE * (lane % LPR) + (k0 + w * KPW + lane / LPR) * ld;
        const int r = 8 * w + (lane >> 3), c = (lane & 7) ^ (r & 7);
        return X + k0 + CE * c + (r0 + r) * ld;
    }
    // LDS destination (wave-uniform) of wave-load w.
    __device__ static inline int dst(int w) { return MN ? w * GS : w * 8 * BK; }
    // Fragment element of lane l: (row = rb + l%16, k = kb + l/16).
    __device__ static inline int frag(int rb, int kb, int lane) {
        const int row = rb + (lane & 15), k = kb + (lane >> 4);
        if (MN) return (k / KPW) * GS + (k % KPW) * 128 + row;
        return row * BK + (((k / CE) ^ (row & 7)) * CE) + (k % CE);
    }
};

// Scheduling of one "read the next k-step while multiplying the current one"
// region: NREAD ds_reads spread evenly over NMFMA MFMAs.
template <int NREAD, int NMFMA>
__device__ inline void packed_interleave() {
    if constexpr (NREAD > NMFMA) {  // more reads than MFMAs: PER or PER + 1 before each
        constexpr int PER = NREAD / NMFMA, EXTRA = NREAD % NMFMA;
#pragma unroll
        for (int q = 0; q < EXTRA; ++q) {
            __builtin_amdgcn_sched_group_barrier(0x100, PER + 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        }
#pragma unroll
        for (int q = EXTRA; q < NMFMA; ++q) {
            __builtin_amdgcn_sched_group_barrier(0x100, PER, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        }
        return;
    }
    constexpr int PER = NMFMA / NREAD;
#pragma unroll
    for (int q = 0; q < NREAD; ++q) {
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);    // DS read
        __builtin_amdgcn_sched_group_barrier(0x008, PER, 0);  // MFMA
    }
    if (NMFMA > PER * NREAD) __builtin_amdgcn_sched_group_barrier(0x008, NMFMA - PER * NREAD, 0);
}

template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
    typedef f64x4 acc_t;
    __device__ static inline acc_t mma(double x, double y, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
    }
    // C column offset of register r (operands swapped: C^T is computed);
    // the f64 form interleaves the four lane groups row by row.
    __device__ static inline int col(int r, int lane) { return 4 * r + (lane >> 4); }
};
template <>
struct Mfma16<float> {
    typedef f32x4 acc_t;
    __device__ static inline acc_t mma(float x, float y, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
    }
    // The f32 form gives lane group g the four consecutive outputs 4g .. 4g+3.
    __device__ static inline int col(int r, int lane) { return 4 * (lane >> 4) + r; }
};

// One tile: panels of storage type T (both operands), MFMA in TS (T or the
// wider double, the fragment read widening each element), C in TC.
template <typename T, typename TS, typename TC, bool AT, bool BT>
struct DmaTileLds {
    static constexpr int BUF = DmaPanel<T, !AT>::ELEMS + DmaPanel<T, BT>::ELEMS;
    static constexpr int BYTES = 2 * BUF * (int)sizeof(T);
};

template <typename T, typename TS, typename TC, bool AT, bool BT>
__device__ inline void gemm_dma_tile(const GemmArgs<T, T, TC, TS>& p, void* lds) {
    constexpr int BM = 128, BN = 128;
    typedef DmaPanel<T, !AT> PA;  // op(A) rows contiguous iff A is 'N'
    typedef DmaPanel<T, BT> PB;   // op(B)^T rows contiguous iff B is 'T'
    typedef Mfma16<TS> M;
    constexpr int BK = PA::BK;
    constexpr int BUF = PA::ELEMS + PB::ELEMS;
    // Tile enumeration, k-range and store mask as in gemm_fp_kernel
    // (triangle grids for syrk / syr2k, triangular operands for trmm).
    int tm, tn, chunk = 0;
    if (p.ksplit > 1) {
        chunk = (int)blockIdx.x % p.ksplit;
        tile_coords_logical(p.tri_tiles, p.tile0 + (int)blockIdx.x / p.ksplit, p.tiles_m, p.tiles_n, tm, tn);
    } else if (p.tri_tiles) {
        tri_tile_coords(p.tri_tiles, xcd_remap((int)blockIdx.x, (int)gridDim.x), p.tiles_m, tm, tn);
    } else {
        tile_coords_kr(p.krange, p.tiles_m, p.tiles_n, tm, tn);
    }
    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 64;
    MBH_LDS T* const L = (MBH_LDS T*)lds;
    // 16 wave-loads per operand per k-tile, four per wave: w = wave + 4v.
    const T* ga[4];
    const T* gb[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        ga[v] = PA::src(p.A, p.lda, m0, 0, wave + 4 * v, lane);
        gb[v] = PB::src(p.B, p.ldb, n0, 0, wave + 4 * v, lane);
    }
    // Per k-tile the sources advance by BK: MN panels by BK rows of ld, K
    // panels by BK elements.
    const int64_t sa = PA::ROWS_CONTIGUOUS ? BK * p.lda : BK;
    const int64_t sb = PB::ROWS_CONTIGUOUS ? BK * p.ldb : BK;
#define MBH_DMAK(kt, BASE)                                                                                      \
    do {                                                                                                        \
        _Pragma("unroll") for (int v = 0; v < 4; ++v) {                                                         \
            __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(ga[v] + (int64_t)(kt) * sa),              \
                                             (MBH_LDS void*)(L + (BASE) + PA::dst(wave + 4 * v)), 16, 0, 0);    \
            __builtin_amdgcn_global_load_lds((const MBH_GLOBAL void*)(gb[v] + (int64_t)(kt) * sb),              \
                                             (MBH_LDS void*)(L + (BASE) + PA::ELEMS + PB::dst(wave + 4 * v)),   \
                                             16, 0, 0);                                                         \
        }                                                                                                       \
        __builtin_amdgcn_sched_barrier(0);                                                                      \
    } while (0)

    typename M::acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = TS(0);
    // fp32 MFMA: fragment reads software-pipelined inside a k-tile, two
    // register sets, the reads of k-step s+1 spread over the 16 MFMAs of
    // step s (the compiler's own order read each step right before its
    // MFMAs; 8192^3 135.6 -> 138.7 TFLOP/s, bit-identical).  fp64 MFMA
    // (64-cycle instructions) keeps the compiler's order, measured 1.5 %
    // faster there.
    constexpr int KS = BK / 4;
    static_assert(KS % 2 == 0, "k-steps come in pairs");
    auto read = [&](TS (&a)[4], TS (&b)[4], int base, int ks) {
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = (TS)L[base + PA::frag(wm0 + 16 * i, 4 * ks, lane)];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = (TS)L[base + PA::ELEMS + PB::frag(wn0 + 16 * j, 4 * ks, lane)];
    };
    auto mfma = [&](const TS (&a)[4], const TS (&b)[4]) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = M::mma(b[j], a[i], acc[i][j]);
    };
    auto ktile = [&](int base) {
        if constexpr (sizeof(TS) == 8) {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                TS a[4], b[4];
                read(a, b, base, ks);
                mfma(a, b);
            }
            return;
        }
        TS a0[4], b0[4], a1[4], b1[4];
        read(a0, b0, base, 0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int ks = 0; ks < KS; ks += 2) {
            read(a1, b1, base, ks + 1);
            mfma(a0, b0);
            packed_interleave<8, 16>();
            __builtin_amdgcn_sched_barrier(0);
            if (ks + 2 < KS) {
                read(a0, b0, base, ks + 2);
                mfma(a1, b1);
                packed_interleave<8, 16>();
            } else {
                mfma(a1, b1);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    int kt0, kt1;
    kt_range(p.krange, tm, tn, BM, BN, BK, (int)(p.k / BK), kt0, kt1);
    if (p.ksplit > 1) {  // this workgroup's share of the tile's k-tiles
        const int len = kt1 - kt0, c0 = kt0 + (int)((int64_t)len * chunk / p.ksplit);
        kt1 = kt0 + (int)((int64_t)len * (chunk + 1) / p.ksplit);
        kt0 = c0;
    }
    if (kt0 < kt1) MBH_DMAK(kt0, 0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int kt = kt0;
    for (; kt + 1 < kt1; kt += 2) {
        MBH_DMAK(kt + 1, BUF);
        ktile(0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (kt + 2 < kt1) MBH_DMAK(kt + 2, 0);
        ktile(BUF);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    if (kt < kt1) ktile(0);
#undef MBH_DMAK

    // Register r of block (i, j) holds C(m0 + wm0 + 16i + lane%16,
    // n0 + wn0 + 16j + M::col(r, lane)); the tile is interior.
    if (p.ksplit > 1) {  // raw partial sum of the chunk, the whole block
        MBH_GLOBAL TS* const part = (MBH_GLOBAL TS*)p.part + (size_t)blockIdx.x * BM * BN;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    part[(wm0 + 16 * i + (lane & 15)) + (wn0 + 16 * j + M::col(r, lane)) * BM] = acc[i][j][r];
        return;
    }
    const TS alpha = p.alpha, beta = p.beta;
    MBH_GLOBAL TC* const C = (MBH_GLOBAL TC*)p.C;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + wm0 + 16 * i + (lane & 15), col = n0 + wn0 + 16 * j + M::col(r, lane);
                if (!tri_keep(p.tri, row, col)) continue;
                MBH_GLOBAL TC* cp = C + row + col * p.ldc;
                TS v = alpha * acc[i][j][r];
                if (beta != TS(0)) v += beta * (TS)*cp;
                *cp = (TC)v;
            }
}

template <typename T, bool AT, bool BT>
__global__ __launch_bounds__(256) void gemm_dma_kernel(GemmArgs<T, T, T, T> p) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[DmaTileLds<T, T, T, AT, BT>::BYTES];
    gemm_dma_tile<T, T, T, AT, BT>(p, lds);
}

// Finishes the tiles of a split launch of gemm_dma_kernel: C = alpha * (sum
// of the ksplit partial blocks, added in chunk order) + beta * C, under the
// store mask.  Block (t, y): tile tile0 + t, columns 8y .. 8y + 7; a thread
// sums four consecutive rows of one column.
template <typename T>
__global__ __launch_bounds__(256) void dma_splitk_reduce_kernel(GemmArgs<T, T, T, T> p) {
    constexpr int BM = 128, BN = 128, TPC = BM / 4, CPB = 256 / TPC;
    static_assert(TPC * CPB == 256 && BN % CPB == 0, "block shape");
    const int t = (int)blockIdx.x;
    int tm, tn;
    tile_coords_logical(p.tri_tiles, p.tile0 + t, p.tiles_m, p.tiles_n, tm, tn);
    const int i = 4 * ((int)threadIdx.x % TPC), j = CPB * (int)blockIdx.y + (int)threadIdx.x / TPC;
    const MBH_GLOBAL T* part = (const MBH_GLOBAL T*)p.part + (size_t)t * p.ksplit * BM * BN + (size_t)j * BM + i;
    T sum[4] = {T(0), T(0), T(0), T(0)};
    for (int c = 0; c < p.ksplit; ++c)
#pragma unroll
        for (int e = 0; e < 4; ++e) sum[e] += part[(size_t)c * BM * BN + e];
    const int64_t col = (int64_t)tn * BN + j;
    if (col >= p.n) return;
    MBH_GLOBAL T* const C = (MBH_GLOBAL T*)p.C;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t row = (int64_t)tm * BM + i + e;
        if (row >= p.m || !tri_keep(p.tri, row, col)) continue;
        MBH_GLOBAL T* cp = C + row + col * p.ldc;
        T v = p.alpha * sum[e];
        if (p.beta != T(0)) v += p.beta * *cp;
        *cp = v;
    }
}

}  // namespace mbh
