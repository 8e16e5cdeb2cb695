// Internal (C++) interface between the C-ABI layer and the kernel launchers.
// Arguments are already validated; pointers are device pointers; all calls
// are asynchronous on `s`.  Return hipSuccess or the first HIP error.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbh {

// C := alpha*op(A)*op(B) + beta*C with the tile kernel of `prec`.
//  tri      : TriMask, restricts the stored entries to one triangle of C;
//  tri_grid : 0 full tile grid, 1/2 only the lower/upper block-triangle tiles
//             (square C);
//  krange   : KRange, each tile visits only the k-tiles where a triangular
//             op(A) / op(B) (starting at the operand's origin) is nonzero.
//  scratch  : device buffer for the split-precision pack planes in place of
//             the stream's workspace (nullptr = the workspace); a product
//             that needs more than it holds returns hipErrorOutOfMemory
//             before anything is queued.
struct Scratch {
    void* p;
    size_t bytes;
};
hipError_t sgemm(bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                 const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int prec, hipStream_t s,
                 int tri = 0, int tri_grid = 0, int krange = 0, const Scratch* scratch = nullptr);
// syr2k's two products C := alpha*op(A)*op(B) + alpha*op(B)*op(A) + beta*C
// (n x n, op() = (at, bt) of the first; tri / tri_grid as for sgemm): in the
// split-precision modes both operands are packed once for the two products.
hipError_t sgemm_syr2k(bool at, bool bt, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                       const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int prec, hipStream_t s, int tri,
                       int tri_grid);
// sgemm in a split-precision mode whose pack reads op(A) / op(B)^T through a
// PackView (aview / bview: a triangle or one stored half of a symmetric
// matrix), so the caller makes no full copy of that operand.  C is written
// only after both operands are packed, so it may alias A or B.  Modes
// without a pack pass of their own (fp32, FP8) return hipErrorNotSupported
// with nothing queued: sgemm_has_views(prec) says which.
bool sgemm_has_views(int prec);
hipError_t sgemm_view(bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                      int aview, const float* B, int64_t ldb, int bview, float beta, float* C, int64_t ldc, int prec,
                      hipStream_t s, int krange = 0);
// Bytes of pack-plane scratch that sgemm needs for a plain product (no
// triangle grid or k-range) in precision `prec`: 0 for fp32, alpha == 0 or
// k == 0.
size_t sgemm_scratch_bytes(int64_t m, int64_t n, int64_t k, float alpha, int prec);
hipError_t dgemm(bool at, bool bt, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                 const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s, int tri = 0,
                 int tri_grid = 0, int krange = 0);

// Native 16-bit gemm: A, B and C hold f16 (f16 = true) or bf16 bit patterns;
// products on the 16-bit MFMA with fp32 accumulation, alpha*acc + beta*C
// formed in fp32 and rounded once to the 16-bit format.
hipError_t gemm16(bool f16, bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t* A,
                  int64_t lda, const uint16_t* B, int64_t ldb, float beta, uint16_t* C, int64_t ldc, hipStream_t s);

// Device scratch owned by the library: grow-only buffers per (device, stream,
// slot); slot 0 belongs to the gemm pack pass, slot 1 to the level-3 drivers,
// slots 2 and 3 to gemv partials and trsv, slot 4 to the split-k partial sums
// of a gemm's partial last round.  The calling thread must hold the
// stream's lease (StreamLease) for the whole call; a buffer stays valid until
// the next request on the same (stream, slot) that asks for more.  Returns
// hipErrorStreamCaptureUnsupported when a slot would have to grow while the
// stream is being captured, hipErrorInvalidHandle without a lease.
hipError_t workspace(hipStream_t s, int slot, size_t bytes, void** out);

// Holds the lease on a stream's workspace for one library call: other
// threads' calls on the same stream wait, so calls queue their kernels whole.
// Nested leases on the same stream by one thread are no-ops.
class StreamLease {
   public:
    explicit StreamLease(hipStream_t s);
    ~StreamLease();
    StreamLease(const StreamLease&) = delete;
    StreamLease& operator=(const StreamLease&) = delete;
    hipError_t error() const { return err_; }

   private:
    void* held_ = nullptr;
    void* prev_ = nullptr;
    hipError_t err_ = hipSuccess;
};

// Test hook for the error paths: the (n+1)-th internal device allocation
// after this call (workspace growth, host staging, batched plans) fails with
// hipErrorOutOfMemory, as a real exhausted device would, once; n < 0 turns
// the hook off.  inject_alloc_failure() is consulted before each of them.
void inject_alloc_failure_after(long n);
bool inject_alloc_failure();

// Frees every workspace buffer of a stream (after synchronising it),
// including those retired while graphs captured on it may still use them.
hipError_t workspace_release(hipStream_t s);
// Bytes of workspace a stream holds.
hipError_t workspace_bytes(hipStream_t s, size_t* bytes);
// trsv's status word (nonzero when a block's bounded wait gave up): recorded
// per stream by the queued trsv, read back after synchronising the stream.
void workspace_note_trsv(hipStream_t s, int* fail);
hipError_t workspace_trsv_status(hipStream_t s, int* failed);

// C := beta*C on an m x n block (beta == 0 writes zeros without reading);
// `tri` restricts to one triangle.
template <typename T>
hipError_t scale(int64_t m, int64_t n, T beta, T* C, int64_t ldc, hipStream_t s, int tri = 0);

}  // namespace mbh
