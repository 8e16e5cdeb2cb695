// Device-resident buffers for callers that keep operands in HBM across calls
// and have no HIP bindings of their own: allocation, column-major matrix
// copies with leading dimensions, and streams.  Arguments are validated
// before any HIP call; failures return MBH_E_* with a message in
// mbh_last_error().
#include <hip/hip_runtime.h>

#include "../../include/mojoblas_hip.h"
#include "abi_internal.hpp"

namespace {

int check_matrix(const char* who, int64_t rows, int64_t cols, int64_t elem_size, const void* src, int64_t lds,
                 const void* dst, int64_t ldd) {
    if (rows < 0 || cols < 0) {
        mbh::set_error("%s: rows %lld / cols %lld negative", who, (long long)rows, (long long)cols);
        return MBH_E_ARG;
    }
    if (elem_size < 1 || elem_size > 16) {
        mbh::set_error("%s: elem_size %lld outside [1, 16]", who, (long long)elem_size);
        return MBH_E_ARG;
    }
    const int64_t minld = rows > 1 ? rows : 1;
    if (lds < minld || ldd < minld) {
        mbh::set_error("%s: leading dimensions %lld / %lld below max(1, rows) = %lld", who, (long long)lds,
                       (long long)ldd, (long long)minld);
        return MBH_E_ARG;
    }
    if (rows > 0 && cols > 0 && (!src || !dst)) {
        mbh::set_error("%s: null buffer", who);
        return MBH_E_ARG;
    }
    return 0;
}

int copy_matrix(const char* who, int64_t rows, int64_t cols, int64_t es, const void* src, int64_t lds, void* dst,
                int64_t ldd, hipMemcpyKind kind, void* stream) {
    if (int rc = check_matrix(who, rows, cols, es, src, lds, dst, ldd)) return rc;
    if (rows == 0 || cols == 0) return 0;
    const hipError_t e = hipMemcpy2DAsync(dst, (size_t)(ldd * es), src, (size_t)(lds * es), (size_t)(rows * es),
                                          (size_t)cols, kind, (hipStream_t)stream);
    return e == hipSuccess ? 0 : mbh::hip_fail(e, who);
}

}  // namespace

extern "C" {

int mbh_dev_alloc(size_t bytes, void** ptr) {
    if (!ptr) {
        mbh::set_error("mbh_dev_alloc: null output pointer");
        return MBH_E_ARG;
    }
    *ptr = nullptr;
    if (bytes == 0) return 0;
    const hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        mbh::set_error("mbh_dev_alloc: hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        return MBH_E_ALLOC;
    }
    return 0;
}

int mbh_dev_free(void* ptr) {
    if (!ptr) return 0;
    const hipError_t e = hipFree(ptr);
    return e == hipSuccess ? 0 : mbh::hip_fail(e, "mbh_dev_free");
}

int mbh_host_alloc(size_t bytes, void** ptr) {
    if (!ptr) {
        mbh::set_error("mbh_host_alloc: null output pointer");
        return MBH_E_ARG;
    }
    *ptr = nullptr;
    if (bytes == 0) return 0;
    const hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        mbh::set_error("mbh_host_alloc: hipHostMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        return MBH_E_ALLOC;
    }
    return 0;
}

int mbh_host_free(void* ptr) {
    if (!ptr) return 0;
    const hipError_t e = hipHostFree(ptr);
    return e == hipSuccess ? 0 : mbh::hip_fail(e, "mbh_host_free");
}

int mbh_set_matrix(int64_t rows, int64_t cols, int64_t elem_size, const void* host, int64_t ldh, void* dev,
                   int64_t ldd, void* stream) {
    return copy_matrix("mbh_set_matrix", rows, cols, elem_size, host, ldh, dev, ldd, hipMemcpyHostToDevice, stream);
}

int mbh_get_matrix(int64_t rows, int64_t cols, int64_t elem_size, const void* dev, int64_t ldd, void* host,
                   int64_t ldh, void* stream) {
    return copy_matrix("mbh_get_matrix", rows, cols, elem_size, dev, ldd, host, ldh, hipMemcpyDeviceToHost, stream);
}

int mbh_stream_create(void** stream) {
    if (!stream) {
        mbh::set_error("mbh_stream_create: null output pointer");
        return MBH_E_ARG;
    }
    *stream = nullptr;
    hipStream_t s = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return mbh::hip_fail(e, "mbh_stream_create");
    *stream = (void*)s;
    return 0;
}

int mbh_stream_destroy(void* stream) {
    if (!stream) return 0;
    // The stream's workspace goes with it (release synchronises the stream).
    hipError_t e = mbh::workspace_release((hipStream_t)stream);
    if (e != hipSuccess) return mbh::hip_fail(e, "mbh_stream_destroy: workspace");
    e = hipStreamDestroy((hipStream_t)stream);
    return e == hipSuccess ? 0 : mbh::hip_fail(e, "mbh_stream_destroy");
}

int mbh_stream_sync(void* stream) {
    const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    return e == hipSuccess ? 0 : mbh::hip_fail(e, "mbh_stream_sync");
}

}  // extern "C"

extern "C" {

int mbh_workspace_release(void* stream) {
    const hipError_t e = mbh::workspace_release((hipStream_t)stream);
    return e == hipSuccess ? 0 : mbh::hip_fail(e, "mbh_workspace_release");
}

int mbh_workspace_bytes(void* stream, size_t* bytes) {
    if (!bytes) {
        mbh::set_error("mbh_workspace_bytes: null output pointer");
        return MBH_E_ARG;
    }
    const hipError_t e = mbh::workspace_bytes((hipStream_t)stream, bytes);
    return e == hipSuccess ? 0 : mbh::hip_fail(e, "mbh_workspace_bytes");
}

}  // extern "C"
