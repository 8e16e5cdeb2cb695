// fp32-data gemm for callers that cannot allocate (code under graph capture,
// tight serving loops, harnesses that forbid device allocation): the same
// kernels as mbh_sgemm_dev, with the split-precision pack planes in a
// file-scope device array instead of the stream's workspace.  No device or
// host allocation, no synchronisation.
//
// Each GPU holds its own copy of `scratch` (MBH_NOALLOC_SCRATCH_BYTES) from
// the first call on it until the process ends; its address on a device comes
// from hipGetSymbolAddress with that device current, once per device.  Every
// call on one GPU shares that array and leaves its contents behind (the pack
// kernels rewrite every byte a call reads, padding included, so nothing has
// to start at zero); calls that overlap on one GPU race.
#include <mutex>

#include "../../include/mojoblas_hip.h"
#include "abi_internal.hpp"
#include "dispatch.hpp"

namespace mbh {
namespace {

__device__ __align__(256) unsigned char scratch[MBH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDev = 16;
std::mutex g_mu;
void* g_ptr[kMaxDev];

// Device address of `scratch` on device `dev` (which is current).
hipError_t scratch_ptr(int dev, void** out) {
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return hipSuccess;
}

}  // namespace
}  // namespace mbh

using namespace mbh;

extern "C" {

int mbh_sgemm_dev_noalloc(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float* a,
                          int64_t lda, const float* b, int64_t ldb, float beta, float* c, int64_t ldc,
                          int precision, void* stream) {
    int info = check_gemm(transa, transb, m, n, k, lda, ldb, ldc);
    if (info) return info;
    if (!gemm_precision_ok(precision)) {
        set_error("mbh_sgemm_dev_noalloc: invalid precision %d", precision);
        return MBH_E_PRECISION;
    }
    if (m == 0 || n == 0) return 0;
    const size_t need = sgemm_scratch_bytes(m, n, k, alpha, precision);
    if (need > (size_t)MBH_NOALLOC_SCRATCH_BYTES) {
        set_error("mbh_sgemm_dev_noalloc: this product needs %zu bytes of pack scratch; the no-allocation entry "
                  "holds %zu (use mbh_sgemm_dev)", need, (size_t)MBH_NOALLOC_SCRATCH_BYTES);
        return MBH_E_ARG;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "mbh_sgemm_dev_noalloc");
    if (dev < 0 || dev >= kMaxDev) {
        set_error("mbh_sgemm_dev_noalloc: device %d outside the %d this entry supports", dev, kMaxDev);
        return MBH_E_ARG;
    }
    if (stream) {
        int sdev = 0;
        if ((e = hipStreamGetDevice((hipStream_t)stream, &sdev)) != hipSuccess)
            return hip_fail(e, "mbh_sgemm_dev_noalloc");
        if (sdev != dev) {
            set_error("mbh_sgemm_dev_noalloc: the stream belongs to device %d, device %d is current", sdev, dev);
            return MBH_E_ARG;
        }
    }
    Scratch sc{nullptr, (size_t)MBH_NOALLOC_SCRATCH_BYTES};
    if (need > 0 && (e = scratch_ptr(dev, &sc.p)) != hipSuccess) return hip_fail(e, "mbh_sgemm_dev_noalloc: scratch");
    e = sgemm(!flag_n(transa), !flag_n(transb), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, precision,
              (hipStream_t)stream, 0, 0, 0, &sc);
    return e == hipSuccess ? 0 : hip_fail(e, "mbh_sgemm_dev_noalloc");
}

}  // extern "C"
