// Internal (C++) interface between the C-ABI layer and the kernel launchers.
// Arguments are already validated; pointers are device pointers; all calls
// are asynchronous on `s`.  Return hipSuccess or the first HIP error.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mbh {

// C := alpha*op(A)*op(B) + beta*C with the tile kernel of `prec`.
// `tri` restricts the stored entries to one triangle of C (TriMask) and, with
// tri_grid, the grid to the tiles touching it (square C only).
hipError_t sgemm(bool at, bool bt, int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                 const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int prec, hipStream_t s,
                 int tri = 0, bool tri_grid = false);
hipError_t dgemm(bool at, bool bt, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                 const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s, int tri = 0,
                 bool tri_grid = false);

// Device scratch owned by the library, one grow-only buffer per
// (device, stream); valid until the next call on the same stream that asks
// for more.
hipError_t workspace(hipStream_t s, size_t bytes, void** out);

// C := beta*C on an m x n block (beta == 0 writes zeros without reading);
// `tri` restricts to one triangle.
template <typename T>
hipError_t scale(int64_t m, int64_t n, T beta, T* C, int64_t ldc, hipStream_t s, int tri = 0);

}  // namespace mbh
