// Launchers for the macro-tile gemm kernels and the beta-scale kernel.
#include <climits>

#include <atomic>
#include <algorithm>

#include "dispatch.hpp"
#include "gemm_fp.hpp"
#include "gemm_packed.hpp"
#include "pack.hpp"
#include "timing.hpp"

namespace mbh {

namespace {

template <typename T>
bool vec_ok(const T* p, int64_t ld) {
    constexpr int VEC = 16 / sizeof(T);
    return ((uintptr_t)p % 16 == 0) && (ld % VEC == 0);
}

template <typename P, int BM, int BN>
bool grid_for(P& p, int tri_grid, dim3& grid) {
    int64_t tm = (p.m + BM - 1) / BM, tn = (p.n + BN - 1) / BN;
    int64_t g = tri_grid ? tm * (tm + 1) / 2 : tm * tn;
    if (g <= 0 || g > INT_MAX || tm > INT_MAX || tn > INT_MAX) return false;
    p.tiles_m = (int)tm;
    p.tiles_n = (int)tn;
    p.tri_tiles = tri_grid;
    grid = dim3((unsigned)g);
    return true;
}

template <typename T, typename TA, typename TB, typename TC, int BM, int BN, int BK>
hipError_t launch_fp(bool at, bool bt, GemmArgs<TA, TB, TC, T> p, int tri_grid, hipStream_t s) {
    dim3 grid;
    if (!grid_for<GemmArgs<TA, TB, TC, T>, BM, BN>(p, tri_grid, grid)) return hipErrorInvalidValue;
    dim3 block(256);
    if (!at && !bt) gemm_fp_kernel<T, TA, TB, TC, false, false, BM, BN, BK><<<grid, block, 0, s>>>(p);
    else if (!at && bt) gemm_fp_kernel<T, TA, TB, TC, false, true, BM, BN, BK><<<grid, block, 0, s>>>(p);
    else if (at && !bt) gemm_fp_kernel<T, TA, TB, TC, true, false, BM, BN, BK><<<grid, block, 0, s>>>(p);
    else gemm_fp_kernel<T, TA, TB, TC, true, true, BM, BN, BK><<<grid, block, 0, s>>>(p);
    return hipGetLastError();
}

// Plain gemm (no triangle mask, full grid, every k-tile) on whole tiles: the
// block of whole 128 x 128 tiles runs the direct-to-LDS kernel (unmasked
// loads only; with the masked path present, the register-staged tile's
// k-loop was mostly branches around it) when both operands allow 16-byte
// loads and k is a multiple of its k-tile; the rows below it (all n
// columns) and the columns right of it go to the generic kernel as
// sub-problems, with EM x EN tiles.  Offsetting op(A) by a multiple of 128
// rows and op(B) by a multiple of 128 columns keeps the 16-byte alignment
// the vector flags record.
template <typename T, int EM, int EN, int EK>
hipError_t launch_fp_split(bool at, bool bt, const GemmParams<T>& p, hipStream_t s) {
    constexpr int BM = 128, BN = 128, IK = DmaPanel<T, true>::BK;
    const int64_t mi = p.m / BM * BM, ni = p.n / BN * BN;
    if (!p.a_vec || !p.b_vec || p.k % IK != 0 || mi == 0 || ni == 0)
        return launch_fp<T, T, T, T, EM, EN, EK>(at, bt, p, 0, s);
    GemmParams<T> q = p;
    q.m = mi;
    q.n = ni;
    dim3 grid;
    if (!grid_for<GemmParams<T>, BM, BN>(q, 0, grid)) return hipErrorInvalidValue;
    const dim3 block(256);
    if (!at && !bt) gemm_dma_kernel<T, false, false><<<grid, block, 0, s>>>(q);
    else if (!at && bt) gemm_dma_kernel<T, false, true><<<grid, block, 0, s>>>(q);
    else if (at && !bt) gemm_dma_kernel<T, true, false><<<grid, block, 0, s>>>(q);
    else gemm_dma_kernel<T, true, true><<<grid, block, 0, s>>>(q);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (mi < p.m) {  // rows [mi, m) x all columns
        GemmParams<T> r = p;
        r.m = p.m - mi;
        r.A = at ? p.A + mi * p.lda : p.A + mi;
        r.C = p.C + mi;
        if ((e = launch_fp<T, T, T, T, EM, EN, EK>(at, bt, r, 0, s)) != hipSuccess) return e;
    }
    if (ni < p.n) {  // rows [0, mi) x columns [ni, n)
        GemmParams<T> r = p;
        r.m = mi;
        r.n = p.n - ni;
        r.B = bt ? p.B + ni : p.B + ni * p.ldb;
        r.C = p.C + ni * p.ldc;
        if ((e = launch_fp<T, T, T, T, EM, EN, EK>(at, bt, r, 0, s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// Whole products on the direct-to-LDS kernel, with any triangle mask,
// triangle grid or k-range: for a C of whole 128 x 128 tiles, k a multiple
// of the kernel's k-tile and operands that allow 16-byte loads.
template <typename T>
bool dma_whole_ok(const GemmParams<T>& p) {
    return p.m % 128 == 0 && p.n % 128 == 0 && p.k % DmaPanel<T, true>::BK == 0 && p.a_vec && p.b_vec;
}
int device_cus();

// Workgroups of gemm_dma_kernel<T> resident at once on the current device
// (CUs x occupancy; the four transposes use the same LDS and registers).
template <typename T>
int64_t dma_round() {
    static std::atomic<int> occ{0};
    if (occ.load(std::memory_order_relaxed) == 0) {
        int r = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, gemm_dma_kernel<T, false, false>, 256, 0) != hipSuccess ||
            r <= 0)
            r = -1;
        occ.store(r, std::memory_order_relaxed);
    }
    return (int64_t)device_cus() * std::max(occ.load(std::memory_order_relaxed), 0);
}

// The direct-to-LDS kernel over every 128 x 128 tile of p.  When the tiles
// leave a last round of workgroups at most half full (ssyrk n = 8192,
// k = 2048: 2080 triangle tiles on 512 resident workgroups = four rounds
// and 32 tiles), the whole rounds run as one launch and the remaining tiles
// are split along k into up to W / tail chunks of >= 8 k-tiles, one
// workgroup each, whose partial sums go to slot-4 workspace and are added
// in chunk order by dma_splitk_reduce_kernel (deterministic).  Not for
// k-range launches, whose tiles differ in length.
template <typename T>
hipError_t launch_dma_whole(bool at, bool bt, GemmParams<T> p, int tri_grid, hipStream_t s) {
    dim3 grid;
    if (!grid_for<GemmParams<T>, 128, 128>(p, tri_grid, grid)) return hipErrorInvalidValue;
    const dim3 block(256);
    auto launch = [&](dim3 g) {
        if (!at && !bt) gemm_dma_kernel<T, false, false><<<g, block, 0, s>>>(p);
        else if (!at && bt) gemm_dma_kernel<T, false, true><<<g, block, 0, s>>>(p);
        else if (at && !bt) gemm_dma_kernel<T, true, false><<<g, block, 0, s>>>(p);
        else gemm_dma_kernel<T, true, true><<<g, block, 0, s>>>(p);
        return hipGetLastError();
    };
    const int64_t ntile = grid.x, W = dma_round<T>(), nk = p.k / DmaPanel<T, true>::BK;
    p.tile0 = 0;
    p.ksplit = 1;
    p.part = nullptr;
    if (p.krange == KR_ALL && W > 0 && ntile > W) {
        const int64_t t = ntile % W, ks = std::min<int64_t>(t > 0 ? W / t : 0, nk / 8);
        if (t > 0 && 2 * t <= W && ks >= 2) {
            void* pw = nullptr;
            hipError_t e = workspace(s, 4, (size_t)t * ks * 128 * 128 * sizeof(T), &pw);
            if (e == hipErrorStreamCaptureUnsupported) return launch(grid);  // capture: no split rather than growth
            if (e != hipSuccess) return e;
            if ((e = launch(dim3((unsigned)(ntile - t)))) != hipSuccess) return e;
            p.tile0 = (int)(ntile - t);
            p.ksplit = (int)ks;
            p.part = pw;
            if ((e = launch(dim3((unsigned)(t * ks)))) != hipSuccess) return e;
            dma_splitk_reduce_kernel<T><<<dim3((unsigned)t, 128 / 8), dim3(256), 0, s>>>(p);
            return hipGetLastError();
        }
    }
    return launch(grid);
}

// Split-precision path: pack op(A) and op(B)^T into bf16 planes, then run the
// packed macro-tile kernel.  256x256 tiles (8 waves) when that still gives one
// tile per CU, else 128x128 (4 waves).
// Pack planes of the split-precision modes: bytes for BT-row tiles.
template <int NPROD, int BT>
size_t packed_bytes(int64_t m, int64_t n, int64_t k) {
    constexpr int PK = 32;
    const int64_t tm = (m + BT - 1) / BT, tn = (n + BT - 1) / BT, ktiles = (k + PK - 1) / PK;
    return (size_t)(tm + tn) * BT * ktiles * PK * ((NPROD == 3) ? 2 : 1) * sizeof(uint16_t);
}

// e4m3 planes plus one int scale per 128x128 tile of each operand.
template <int BT>
size_t fp8_bytes(int64_t m, int64_t n, int64_t k) {
    constexpr int PK = 64;
    const int64_t tm = (m + BT - 1) / BT, tn = (n + BT - 1) / BT;
    const int64_t kblocks = (k + 127) / 128, ktiles = 2 * kblocks;
    return (size_t)(tm + tn) * BT * ktiles * PK + (size_t)(tm + tn) * (BT / 128) * kblocks * sizeof(int);
}

// Scratch for the pack planes: the caller's buffer when one is given, else
// slot 0 of the stream's workspace.
hipError_t pack_scratch(const Scratch* ext, hipStream_t s, size_t bytes, void** out) {
    if (!ext) return workspace(s, 0, bytes, out);
    if (bytes > ext->bytes) return hipErrorOutOfMemory;
    *out = ext->p;
    return hipSuccess;
}

// Compute units of the current device (cached per device).
int device_cus() {
    static std::atomic<int> cache[16];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return 0;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v == 0) {
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) return 0;
        cache[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}

// Split-precision path: pack op(A) and op(B)^T into bf16 planes, then run the
// packed macro-tile kernel.  256x256 tiles (8 waves) when that still gives one
// tile per CU, else 128x128 (4 waves).
//  * An operand pair that is one matrix (syrk: op(A) == op(B)^T, same
//    pointer and ld) is packed once and both sides read the same planes.
//  * Partial last round: when the tiles leave a last round of workgroups at
//    most half full (syrk n = 8192: 528 tiles on 256 CUs = two rounds and 16
//    tiles), the whole rounds run as usual and the remaining tiles are split
//    along k into up to W / tail chunks of >= 8 k-tiles, one workgroup each,
//    whose partial sums go to slot-4 workspace and are added in chunk order
//    by splitk_reduce_kernel (deterministic).  Not for k-range launches or a
//    caller's scratch.
// Resident workgroups per round of one packed-kernel instantiation on the
// current device (0 if unknown).
template <int NPROD, int BT, int WM_, int WN_>
int64_t packed_round() {
    static std::atomic<int> occ{0};
    if (occ.load(std::memory_order_relaxed) == 0) {
        int r = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, gemm_packed_kernel<NPROD, BT, BT, WM_, WN_>,
                                                         64 * WM_ * WN_, 0) != hipSuccess || r <= 0)
            r = -1;
        occ.store(r, std::memory_order_relaxed);
    }
    return (int64_t)device_cus() * std::max(occ.load(std::memory_order_relaxed), 0);
}

// Launch of the packed main kernel over all tiles of p, with the partial
// last round split along k (see run_packed) when `split` allows it.
template <int NPROD, int BT, int WM_, int WN_>
hipError_t packed_main(PackedGemmParams p, int64_t ntile, bool split, hipStream_t s) {
    constexpr int NT = 64 * WM_ * WN_;
    const int64_t W = packed_round<NPROD, BT, WM_, WN_>();
    int64_t nmain = ntile, tail = 0;
    int ksplit = 1;
    hipError_t e;
    if (split && p.krange == KR_ALL && W > 0 && ntile > W) {
        const int64_t t = ntile % W;
        const int64_t ks = std::min<int64_t>(t > 0 ? W / t : 0, p.ktiles / 8);
        if (t > 0 && 2 * t <= W && ks >= 2) {
            tail = t;
            ksplit = (int)ks;
            nmain = ntile - tail;
            void* pw = nullptr;
            if ((e = workspace(s, 4, (size_t)tail * ksplit * BT * BT * sizeof(float), &pw)) != hipSuccess) return e;
            p.part = (float*)pw;
        }
    }
    p.tile0 = 0;
    p.ksplit = 1;
    gemm_packed_kernel<NPROD, BT, BT, WM_, WN_><<<dim3((unsigned)nmain), dim3(NT), 0, s>>>(p);
    if (tail > 0) {
        if ((e = hipGetLastError()) != hipSuccess) return e;
        p.tile0 = (int)nmain;
        p.ksplit = ksplit;
        gemm_packed_kernel<NPROD, BT, BT, WM_, WN_><<<dim3((unsigned)(tail * ksplit)), dim3(NT), 0, s>>>(p);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        splitk_reduce_kernel<BT, BT><<<dim3((unsigned)tail, BT * BT / 1024), dim3(256), 0, s>>>(p);
    }
    return hipGetLastError();
}

// Split-precision path: pack op(A) and op(B)^T into bf16 planes, then run the
// packed macro-tile kernel.  256x256 tiles (8 waves) when that still gives one
// tile per CU, else 128x128 (4 waves).
//  * An operand pair that is one matrix (syrk: op(A) == op(B)^T, same
//    pointer and ld) is packed once and both sides read the same planes.
//  * syr2k (pair2 != nullptr): a second product op(B) op(A)^T -- the same
//    two planes, swapped -- is added into C (beta 1) by a second launch, so
//    both operands are packed once for the two products.
//  * Partial last round: when the tiles leave a last round of workgroups at
//    most half full (syrk n = 8192: 528 tiles on 256 CUs = two rounds and 16
//    tiles), the whole rounds run as usual and the remaining tiles are split
//    along k into up to W / tail chunks of >= 8 k-tiles, one workgroup each,
//    whose partial sums go to slot-4 workspace and are added in chunk order
//    by splitk_reduce_kernel (deterministic).  Not for k-range launches or a
//    caller's scratch.
//  * aview / bview (PackView): the pack reads op(A) / op(B)^T as a triangle
//    or from one stored half of a symmetric matrix (trmm, symm), so that no
//    copy of the operand is made first.
template <int NPROD, int BT_, int WM_, int WN_>
hipError_t run_packed(bool at, bool bt, const GemmParams<float>& g, int tri_grid, hipStream_t s,
                      const Scratch* ext, bool pair2 = false, int aview = PV_PLAIN, int bview = PV_PLAIN) {
    constexpr int BT = BT_, PK = 32;
    PackedGemmParams p{};
    const int64_t tm = (g.m + BT - 1) / BT, tn = (g.n + BT - 1) / BT;
    const int64_t ktiles = (g.k + PK - 1) / PK;
    const int64_t ntile = tri_grid ? tm * (tm + 1) / 2 : tm * tn;
    if (ntile <= 0 || ntile > INT_MAX || ktiles > INT_MAX) return hipErrorInvalidValue;
    const int planes = (NPROD == 3) ? 2 : 1;
    const bool shared = g.A == g.B && g.lda == g.ldb && g.m == g.n && at != bt && aview == PV_PLAIN &&
                        bview == PV_PLAIN;
    const size_t a_el = (size_t)tm * BT * ktiles * PK, b_el = shared ? 0 : (size_t)tn * BT * ktiles * PK;
    void* ws = nullptr;
    hipError_t e = pack_scratch(ext, s, (a_el + b_el) * planes * sizeof(uint16_t), &ws);
    if (e != hipSuccess) return e;
    uint16_t* base = (uint16_t*)ws;
    PackParams pa{g.A, g.lda, g.m, g.k, (int)ktiles, base, planes == 2 ? base + a_el : nullptr, g.a_vec, aview};
    uint16_t* bbase = base + a_el * planes;
    PackParams pb{g.B, g.ldb, g.n, g.k, (int)ktiles, bbase, planes == 2 ? bbase + b_el : nullptr, g.b_vec, bview};
    if (shared) {
        pb.hi = pa.hi;
        pb.lo = pa.lo;
    }
    TimingMark tmk;
    tmk.active = timing_on();
    timing_mark(tmk, 0, s);
    // One launch packs both operands: 64-row x 128-k blocks, op(A) first.
    // op(A) rows = m: k-contiguous iff A is transposed; op(B)^T rows = n:
    // k-contiguous iff B is NOT transposed.  A shared pair: op(A)'s blocks only.
    const int64_t kblk = (ktiles + PACK_NKB - 1) / PACK_NKB;
    const int64_t gx_a = tm * BT / 64, gx_b = tn * BT / 64;
    const int64_t nblk_a = gx_a * kblk, nblk = nblk_a + (shared ? 0 : gx_b * kblk);
    if (nblk > INT_MAX) return hipErrorInvalidValue;
    dim3 pgrid((unsigned)nblk), pblock(256);
    const int ga = (int)gx_a, na = (int)nblk_a, gb = (int)gx_b;
    if (at && !bt) pack_split_kernel<true, true, BT, PK><<<pgrid, pblock, 0, s>>>(pa, pb, ga, na, gb);
    else if (at && bt) pack_split_kernel<true, false, BT, PK><<<pgrid, pblock, 0, s>>>(pa, pb, ga, na, gb);
    else if (!at && !bt) pack_split_kernel<false, true, BT, PK><<<pgrid, pblock, 0, s>>>(pa, pb, ga, na, gb);
    else pack_split_kernel<false, false, BT, PK><<<pgrid, pblock, 0, s>>>(pa, pb, ga, na, gb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    p.m = g.m;
    p.n = g.n;
    p.ktiles = (int)ktiles;
    p.Ah = pa.hi;
    p.Al = pa.lo;
    p.Bh = pb.hi;
    p.Bl = pb.lo;
    p.C = g.C;
    p.ldc = g.ldc;
    p.alpha = g.alpha;
    p.beta = g.beta;
    p.tiles_m = (int)tm;
    p.tiles_n = (int)tn;
    p.tri = g.tri;
    p.tri_tiles = tri_grid;
    p.krange = g.krange;
    timing_mark(tmk, 1, s);
    if ((e = packed_main<NPROD, BT, WM_, WN_>(p, ntile, !ext, s)) != hipSuccess) return e;
    if (pair2) {  // + op(B) op(A)^T from the same planes, added to C
        std::swap(p.Ah, p.Bh);
        std::swap(p.Al, p.Bl);
        p.beta = 1.0f;
        if ((e = packed_main<NPROD, BT, WM_, WN_>(p, ntile, !ext, s)) != hipSuccess) return e;
    }
    timing_mark(tmk, 2, s);
    return hipGetLastError();
}

// FP8 path: per-128x128-tile scaled e4m3 pack, then the packed kernel in fp8
// mode (64-deep k-tiles; K padded to a multiple of 128).
template <int BT_, int WM_, int WN_>
hipError_t run_fp8(bool at, bool bt, const GemmParams<float>& g, int tri_grid, hipStream_t s, const Scratch* ext) {
    constexpr int BT = BT_, PK = 64;
    const int64_t tm = (g.m + BT - 1) / BT, tn = (g.n + BT - 1) / BT;
    const int64_t kblocks = (g.k + 127) / 128, ktiles = 2 * kblocks;
    const int64_t ntile = tri_grid ? tm * (tm + 1) / 2 : tm * tn;
    if (ntile <= 0 || ntile > INT_MAX || ktiles > INT_MAX) return hipErrorInvalidValue;
    const size_t qa = (size_t)tm * BT * ktiles * PK, qb = (size_t)tn * BT * ktiles * PK;
    const size_t na = (size_t)tm * (BT / 128) * kblocks;
    void* ws = nullptr;
    hipError_t e = pack_scratch(ext, s, fp8_bytes<BT>(g.m, g.n, g.k), &ws);
    if (e != hipSuccess) return e;
    uint8_t* base = (uint8_t*)ws;
    int* sa = (int*)(base + qa + qb);
    int* sb = sa + na;
    PackFp8Params pa{g.A, g.lda, g.m, g.k, (int)ktiles, (int)kblocks, base, sa, g.a_vec};
    PackFp8Params pb{g.B, g.ldb, g.n, g.k, (int)ktiles, (int)kblocks, base + qa, sb, g.b_vec};
    TimingMark tmk;
    tmk.active = timing_on();
    timing_mark(tmk, 0, s);
    // One launch packs both operands in 128x128 blocks, op(A) first: op(A)
    // is k-contiguous iff A is transposed, op(B)^T iff B is not.
    const int64_t gx_a = tm * BT / 128, gx_b = tn * BT / 128;
    const int64_t nblk_a = gx_a * kblocks, nblk = nblk_a + gx_b * kblocks;
    if (nblk > INT_MAX) return hipErrorInvalidValue;
    const dim3 pgrid((unsigned)nblk);
    const int xa = (int)gx_a, na_ = (int)nblk_a, xb = (int)gx_b;
    if (at && !bt) pack_fp8_kernel<true, true, BT><<<pgrid, 256, 0, s>>>(pa, pb, xa, na_, xb);
    else if (at && bt) pack_fp8_kernel<true, false, BT><<<pgrid, 256, 0, s>>>(pa, pb, xa, na_, xb);
    else if (!at && !bt) pack_fp8_kernel<false, true, BT><<<pgrid, 256, 0, s>>>(pa, pb, xa, na_, xb);
    else pack_fp8_kernel<false, false, BT><<<pgrid, 256, 0, s>>>(pa, pb, xa, na_, xb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    PackedGemmParams p{};
    p.m = g.m;
    p.n = g.n;
    p.ktiles = (int)ktiles;
    p.Ah = (const uint16_t*)pa.q;
    p.Bh = (const uint16_t*)pb.q;
    p.C = g.C;
    p.ldc = g.ldc;
    p.alpha = g.alpha;
    p.beta = g.beta;
    p.tiles_m = (int)tm;
    p.tiles_n = (int)tn;
    p.tri = g.tri;
    p.tri_tiles = tri_grid;
    p.krange = g.krange;
    p.sA = sa;
    p.sB = sb;
    p.kblocks = (int)kblocks;
    timing_mark(tmk, 1, s);
    gemm_packed_kernel<8, BT, BT, WM_, WN_><<<dim3((unsigned)ntile), dim3(64 * WM_ * WN_), 0, s>>>(p);
    timing_mark(tmk, 2, s);
    return hipGetLastError();
}

// Workgroups a launch with BM x BN tiles would have (block triangle when
// tri_grid is set).
int64_t tile_count(int64_t m, int64_t n, int BM, int BN, int tri_grid) {
    const int64_t tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    return tri_grid ? tm * (tm + 1) / 2 : tm * tn;
}

// Tile size by grid fill: 256x256 (8 waves) while that gives a tile per CU,
// then 128x128, then 64x64 for the small products of the level-3 recursions.
// A k-range launch (triangular operand) asks for four tiles per CU before it
// takes the larger tile: its tiles differ in length, and with one tile per CU
// the longest one would set the time.
int packed_tile(int nprod, int64_t m, int64_t n, int tri_grid, int krange) {
    if (tile_count(m, n, 256, 256, tri_grid) >= (krange ? 1024 : 256)) return 256;
    if (nprod == 8 || tile_count(m, n, 128, 128, tri_grid) >= 128) return 128;
    return 64;
}

template <int NPROD>
hipError_t launch_packed(bool at, bool bt, const GemmParams<float>& g, int tri_grid, hipStream_t s,
                         const Scratch* ext, bool pair2 = false, int aview = PV_PLAIN, int bview = PV_PLAIN) {
    const int bt_ = packed_tile(NPROD, g.m, g.n, tri_grid, g.krange);
    if constexpr (NPROD == 8) {
        if (aview != PV_PLAIN || bview != PV_PLAIN) return hipErrorNotSupported;
        if (bt_ == 256) return run_fp8<256, 2, 4>(at, bt, g, tri_grid, s, ext);
        return run_fp8<128, 2, 2>(at, bt, g, tri_grid, s, ext);
    } else {
        if (bt_ == 256) return run_packed<NPROD, 256, 2, 4>(at, bt, g, tri_grid, s, ext, pair2, aview, bview);
        if (bt_ == 128) return run_packed<NPROD, 128, 2, 2>(at, bt, g, tri_grid, s, ext, pair2, aview, bview);
        return run_packed<NPROD, 64, 2, 2>(at, bt, g, tri_grid, s, ext, pair2, aview, bview);
    }
}


// Native 16-bit path.  op(A) (rows m) is k-contiguous in memory iff A is
// transposed, op(B)^T (rows n) iff B is not; such an operand with K a
// multiple of 64, a 16-byte aligned base and ld % 8 == 0 is read in place by
// the main kernel.  So is an operand whose rows are the contiguous index (A
// not transposed, B transposed) under the same conditions with its row count
// a multiple of 8: it is loaded as k-rows (MN form) and its fragments are
// taken with transposing LDS reads.  Any other operand is first re-laid into
// the packed single plane (bits unchanged; one launch packs whichever
// operands need it).
template <bool F16, int BT, int WM_, int WN_>
hipError_t run_packed16(bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t* A,
                        int64_t lda, const uint16_t* B, int64_t ldb, float beta, uint16_t* C, int64_t ldc,
                        hipStream_t s) {
    constexpr int PK = GEMM16_BK;
    const int64_t tm = (m + BT - 1) / BT, tn = (n + BT - 1) / BT;
    const int64_t ktiles = (k + PK - 1) / PK;
    const int64_t ntile = tm * tn;
    if (ntile <= 0 || ntile > INT_MAX || ktiles > INT_MAX) return hipErrorInvalidValue;
    const bool mn_a = !at && k % PK == 0 && m % 8 == 0 && vec_ok(A, lda);
    const bool mn_b = bt && k % PK == 0 && n % 8 == 0 && vec_ok(B, ldb);
    const bool direct_a = mn_a || (at && k % PK == 0 && vec_ok(A, lda));
    const bool direct_b = mn_b || (!bt && k % PK == 0 && vec_ok(B, ldb));
    const size_t a_el = direct_a ? 0 : (size_t)tm * BT * ktiles * PK;
    const size_t b_el = direct_b ? 0 : (size_t)tn * BT * ktiles * PK;
    uint16_t* base = nullptr;
    hipError_t e;
    if (a_el + b_el) {
        void* ws = nullptr;
        if ((e = workspace(s, 0, (a_el + b_el) * sizeof(uint16_t), &ws)) != hipSuccess) return e;
        base = (uint16_t*)ws;
    }
    Pack16Params pa{A, lda, m, k, tm * BT, (int)ktiles, base, (int)vec_ok(A, lda)};
    Pack16Params pb{B, ldb, n, k, tn * BT, (int)ktiles, base + a_el, (int)vec_ok(B, ldb)};
    TimingMark tmk;
    tmk.active = timing_on();
    timing_mark(tmk, 0, s);
    // 128-row x 128-k pack blocks; an operand read in place gets none.
    const int64_t kblk = (ktiles * PK + 127) / 128;
    const int64_t gx_a = (tm * BT + 127) / 128, gx_b = (tn * BT + 127) / 128;
    const int64_t nblk_a = direct_a ? 0 : gx_a * kblk, nblk = nblk_a + (direct_b ? 0 : gx_b * kblk);
    if (nblk > INT_MAX) return hipErrorInvalidValue;
    if (nblk > 0) {
        const dim3 pgrid((unsigned)nblk), pblock(256);
        const int ga = (int)gx_a, na = (int)nblk_a, gb = (int)gx_b;
        if (at && !bt) pack16_kernel<true, true, BT, PK><<<pgrid, pblock, 0, s>>>(pa, pb, ga, na, gb);
        else if (at && bt) pack16_kernel<true, false, BT, PK><<<pgrid, pblock, 0, s>>>(pa, pb, ga, na, gb);
        else if (!at && !bt) pack16_kernel<false, true, BT, PK><<<pgrid, pblock, 0, s>>>(pa, pb, ga, na, gb);
        else pack16_kernel<false, false, BT, PK><<<pgrid, pblock, 0, s>>>(pa, pb, ga, na, gb);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    PackedGemmParams p{};
    p.m = m;
    p.n = n;
    p.ktiles = (int)ktiles;
    p.Ah = direct_a ? A : pa.out;
    p.Bh = direct_b ? B : pb.out;
    p.a_ld = direct_a ? lda : 0;
    p.b_ld = direct_b ? ldb : 0;
    p.C = C;
    p.ldc = ldc;
    p.alpha = alpha;
    p.beta = beta;
    p.tiles_m = (int)tm;
    p.tiles_n = (int)tn;
    timing_mark(tmk, 1, s);
    const dim3 grid((unsigned)ntile), block(64 * WM_ * WN_);
    if (mn_a && mn_b) gemm_packed16_kernel<F16, BT, BT, WM_, WN_, true, true><<<grid, block, 0, s>>>(p);
    else if (mn_a) gemm_packed16_kernel<F16, BT, BT, WM_, WN_, true, false><<<grid, block, 0, s>>>(p);
    else if (mn_b) gemm_packed16_kernel<F16, BT, BT, WM_, WN_, false, true><<<grid, block, 0, s>>>(p);
    else gemm_packed16_kernel<F16, BT, BT, WM_, WN_><<<grid, block, 0, s>>>(p);
    timing_mark(tmk, 2, s);
    return hipGetLastError();
}

// C := beta*C on 16-bit storage (beta == 0 writes zeros without reading).
template <bool F16>
__global__ void scale16_kernel(int64_t m, int64_t n, float beta, uint16_t* C, int64_t ldc) {
    typedef CElem<F16 ? 1 : 2> E;
    const int64_t total = m * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        uint16_t* c = C + idx % m + (idx / m) * ldc;
        *c = (beta == 0.0f) ? (uint16_t)0 : E::store(beta * E::load(*c));
    }
}

template <typename T>
__global__ void scale_kernel(int64_t m, int64_t n, T beta, T* C, int64_t ldc, int tri) {
    const int64_t total = m * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % m, j = idx / m;
        if (!tri_keep(tri, i, j)) continue;
        T* c = C + i + j * ldc;
        *c = (beta == T(0)) ? T(0) : beta * *c;
    }
}

template <typename T>
GemmParams<T> make_params(int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B,
                          int64_t ldb, T beta, T* C, int64_t ldc, int tri, int krange) {
    GemmParams<T> p{};
    p.m = m;
    p.n = n;
    p.k = k;
    p.A = A;
    p.lda = lda;
    p.B = B;
    p.ldb = ldb;
    p.C = C;
    p.ldc = ldc;
    p.alpha = alpha;
    p.beta = beta;
    p.tri = tri;
    p.krange = krange;
    p.a_vec = vec_ok(A, lda);
    p.b_vec = vec_ok(B, ldb);
    return p;
}

}  // namespace

size_t sgemm_scratch_bytes(int64_t m, int64_t n, int64_t k, float alpha, int prec) {
    if (m <= 0 || n <= 0 || k <= 0 || alpha == 0.0f) return 0;
    if (prec != PREC_BF16X3 && prec != PREC_BF16 && prec != PREC_FP8) return 0;
    const int nprod = prec == PREC_FP8 ? 8 : (prec == PREC_BF16X3 ? 3 : 1);
    const int t = packed_tile(nprod, m, n, 0, 0);
    if (nprod == 8) return t == 256 ? fp8_bytes<256>(m, n, k) : fp8_bytes<128>(m, n, k);
    if (nprod == 3)
        return t == 256 ? packed_bytes<3, 256>(m, n, k) : t == 128 ? packed_bytes<3, 128>(m, n, k)
                                                             : packed_bytes<3, 64>(m, n, k);
    return t == 256 ? packed_bytes<1, 256>(m, n, k) : t == 128 ? packed_bytes<1, 128>(m, n, k)
                                                         : packed_bytes<1, 64>(m, n, k);
}

template <typename T>
hipError_t scale(int64_t m, int64_t n, T beta, T* C, int64_t ldc, hipStream_t s, int tri) {
    if (m <= 0 || n <= 0 || beta == T(1)) return hipSuccess;
    int64_t blocks = (m * n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    scale_kernel<T><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(m, n, beta, C, ldc, tri);
    return hipGetLastError();
}
template hipError_t scale<float>(int64_t, int64_t, float, float*, int64_t, hipStream_t, int);
template hipError_t scale<double>(int64_t, int64_t, double, double*, int64_t, hipStream_t, int);

hipError_t sgemm(bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                 const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int prec, hipStream_t s, int tri,
                 int tri_grid, int krange, const Scratch* scratch) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (alpha == 0.0f || k == 0) return scale<float>(m, n, beta, C, ldc, s, tri);
    GemmParams<float> p = make_params<float>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, krange);
    switch (prec) {
        case PREC_FP32:
            if (tile_count(m, n, 128, 128, tri_grid) >= (krange ? 1024 : 256)) {
                // Plain products: interior kernel plus edge strips (8192^3:
                // 126.8 -> 131.7 TFLOP/s, bit-identical).
                if (!tri && !tri_grid && !krange)
                    return launch_fp_split<float, 64, 64, 16>(at, bt, p, s);
                if (dma_whole_ok(p)) return launch_dma_whole<float>(at, bt, p, tri_grid, s);
                return launch_fp<float, float, float, float, 128, 128, 16>(at, bt, p, tri_grid, s);
            }
            return launch_fp<float, float, float, float, 64, 64, 16>(at, bt, p, tri_grid, s);
        case PREC_BF16X3:
            return launch_packed<3>(at, bt, p, tri_grid, s, scratch);
        case PREC_BF16:
            return launch_packed<1>(at, bt, p, tri_grid, s, scratch);
        case PREC_FP8:
            return launch_packed<8>(at, bt, p, tri_grid, s, scratch);
        default:
            return hipErrorInvalidValue;
    }
}

hipError_t sgemm_syr2k(bool at, bool bt, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                       const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int prec, hipStream_t s, int tri,
                       int tri_grid) {
    if (n <= 0) return hipSuccess;
    if (alpha == 0.0f || k == 0) return scale<float>(n, n, beta, C, ldc, s, tri);
    if (prec == PREC_BF16X3 || prec == PREC_BF16) {
        GemmParams<float> p = make_params<float>(n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, 0);
        return prec == PREC_BF16X3 ? launch_packed<3>(at, bt, p, tri_grid, s, nullptr, true)
                                   : launch_packed<1>(at, bt, p, tri_grid, s, nullptr, true);
    }
    hipError_t e = sgemm(at, bt, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec, s, tri, tri_grid);
    if (e != hipSuccess) return e;
    return sgemm(at, bt, n, n, k, alpha, B, ldb, A, lda, 1.0f, C, ldc, prec, s, tri, tri_grid);
}

bool sgemm_has_views(int prec) { return prec == PREC_BF16X3 || prec == PREC_BF16; }

hipError_t sgemm_view(bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                      int aview, const float* B, int64_t ldb, int bview, float beta, float* C, int64_t ldc, int prec,
                      hipStream_t s, int krange) {
    if (!sgemm_has_views(prec)) return hipErrorNotSupported;
    if (m <= 0 || n <= 0) return hipSuccess;
    if (alpha == 0.0f || k == 0) return scale<float>(m, n, beta, C, ldc, s);
    GemmParams<float> p = make_params<float>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0, krange);
    return prec == PREC_BF16X3 ? launch_packed<3>(at, bt, p, 0, s, nullptr, false, aview, bview)
                               : launch_packed<1>(at, bt, p, 0, s, nullptr, false, aview, bview);
}

hipError_t dgemm(bool at, bool bt, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                 const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s, int tri,
                 int tri_grid, int krange) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (alpha == 0.0 || k == 0) return scale<double>(m, n, beta, C, ldc, s, tri);
    GemmParams<double> p = make_params<double>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, krange);
    // Plain products large enough to fill the chip with 128 x 128 tiles take
    // them (interior kernel plus edge strips: 66.3 TFLOP/s at 8192^3; the
    // generic kernel measured 63.5 with 128 x 128 tiles, 60.6 with 64 x 64
    // and 62.1 with 128 x 64).
    if (!tri && !tri_grid && !krange && tile_count(m, n, 128, 128, 0) >= 256)
        return launch_fp_split<double, 64, 64, 16>(at, bt, p, s);
    if (tile_count(m, n, 128, 128, tri_grid) >= (krange ? 1024 : 256) && dma_whole_ok(p))
        return launch_dma_whole<double>(at, bt, p, tri_grid, s);
    return launch_fp<double, double, double, double, 64, 64, 16>(at, bt, p, tri_grid, s);
}

template <bool F16>
static hipError_t gemm16_t(bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t* A,
                           int64_t lda, const uint16_t* B, int64_t ldb, float beta, uint16_t* C, int64_t ldc,
                           hipStream_t s) {
    if (alpha == 0.0f || k == 0) {
        if (beta == 1.0f) return hipSuccess;
        int64_t blocks = (m * n + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        scale16_kernel<F16><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(m, n, beta, C, ldc);
        return hipGetLastError();
    }
    // Tile size by grid fill, as for the split-precision path.
    if (tile_count(m, n, 256, 256, 0) >= 256)
        return run_packed16<F16, 256, 2, 4>(at, bt, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
    if (tile_count(m, n, 128, 128, 0) >= 128)
        return run_packed16<F16, 128, 2, 2>(at, bt, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
    return run_packed16<F16, 64, 2, 2>(at, bt, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

hipError_t gemm16(bool f16, bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t* A,
                  int64_t lda, const uint16_t* B, int64_t ldb, float beta, uint16_t* C, int64_t ldc, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (f16) return gemm16_t<true>(at, bt, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
    return gemm16_t<false>(at, bt, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

}  // namespace mbh
