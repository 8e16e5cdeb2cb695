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
hipError_t sgemm(bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                 const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int prec, hipStream_t s,
                 int tri = 0, int tri_grid = 0, int krange = 0);
hipError_t dgemm(bool at, bool bt, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                 const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s, int tri = 0,
                 int tri_grid = 0, int krange = 0);

// Native 16-bit gemm: A, B and C hold f16 (f16 = true) or bf16 bit patterns;
// products on the 16-bit MFMA with fp32 accumulation, alpha*acc + beta*C
// formed in fp32 and rounded once to the 16-bit format.
hipError_t gemm16(bool f16, bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t* A,
                  int64_t lda, const uint16_t* B, int64_t ldb, float beta, uint16_t* C, int64_t ldc, hipStream_t s);

// fp64-MFMA gemm over mixed storage types (fp32/fp64 operands, fp32/fp64 C),
// all arithmetic in fp64.  shape 0: 64x64 tiles; 1: 128x64 (a <=128-row C is
// one tile row, so C may alias B); 2: 64x128 (C may alias A when n <= 128).
template <typename TA, typename TB, typename TC>
hipError_t gemm_f64acc(bool at, bool bt, int64_t m, int64_t n, int64_t k, double alpha, const TA* A, int64_t lda,
                       const TB* B, int64_t ldb, double beta, TC* C, int64_t ldc, hipStream_t s, int shape,
                       int krange = 0);

// Device scratch owned by the library: grow-only buffers per (device, stream,
// slot); slot 0 belongs to the gemm pack pass, slot 1 to the level-3 drivers.
// A buffer stays valid until the next request on the same (stream, slot) that
// asks for more.
hipError_t workspace(hipStream_t s, int slot, size_t bytes, void** out);

// C := beta*C on an m x n block (beta == 0 writes zeros without reading);
// `tri` restricts to one triangle.
template <typename T>
hipError_t scale(int64_t m, int64_t n, T beta, T* C, int64_t ldc, hipStream_t s, int tri = 0);

}  // namespace mbh
