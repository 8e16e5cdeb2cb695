// C-ABI entry points (include/mojoblas_hip.h): argument validation with the
// reference's Netlib info codes, quick returns, host-pointer staging, and
// dispatch to the kernel launchers.  No entry runs any GPU work before its
// arguments are validated.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include <sys/syscall.h>
#include <unistd.h>

#include "../../include/mojoblas_hip.h"
#include "abi_internal.hpp"
#include "dispatch.hpp"

namespace mbh {

static thread_local std::string g_err;
static int g_default_prec = -1;
static std::mutex g_prec_mu;

void set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int hip_fail(hipError_t e, const char* where) {
    if (e == hipErrorStreamCaptureUnsupported) {
        set_error("%s: its workspace would have to grow while the stream is being captured; run the same call "
                  "(or a larger one) on this stream once before capture", where);
        return MBH_E_ARG;
    }
    set_error("%s: %s", where, hipGetErrorString(e));
    return MBH_E_HIP;
}

bool flag_n(char c) { return c == 'N' || c == 'n'; }
bool flag_t(char c) { return c == 'T' || c == 't' || c == 'C' || c == 'c'; }
bool flag_l(char c) { return c == 'L' || c == 'l'; }
bool flag_r(char c) { return c == 'R' || c == 'r'; }
bool flag_u(char c) { return c == 'U' || c == 'u'; }
static int64_t max1(int64_t v) { return v > 1 ? v : 1; }

// gemm.mojo:179-213
int check_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc) {
    if (!flag_n(ta) && !flag_t(ta)) return 1;
    if (!flag_n(tb) && !flag_t(tb)) return 2;
    if (m < 0) return 3;
    if (n < 0) return 4;
    if (k < 0) return 5;
    if (lda < max1(flag_n(ta) ? m : k)) return 8;
    if (ldb < max1(flag_n(tb) ? k : n)) return 10;
    if (ldc < max1(m)) return 13;
    return 0;
}

// trsm.mojo:62-89 and trmm.mojo:62-89 (identical numbering)
int check_trxm(char side, char uplo, char ta, char diag, int64_t m, int64_t n, int64_t lda, int64_t ldb) {
    if (!flag_l(side) && !flag_r(side)) return 1;
    if (!flag_u(uplo) && !flag_l(uplo)) return 2;
    if (!flag_n(ta) && !flag_t(ta)) return 3;
    if (!flag_u(diag) && !flag_n(diag)) return 4;
    if (m < 0) return 5;
    if (n < 0) return 6;
    if (lda < max1(flag_l(side) ? m : n)) return 9;
    if (ldb < max1(m)) return 11;
    return 0;
}

// syrk.mojo:60-83
int check_syrk(char uplo, char trans, int64_t n, int64_t k, int64_t lda, int64_t ldc) {
    if (!flag_u(uplo) && !flag_l(uplo)) return 1;
    if (!flag_n(trans) && !flag_t(trans)) return 2;
    if (n < 0) return 3;
    if (k < 0) return 4;
    if (lda < max1(flag_n(trans) ? n : k)) return 7;
    if (ldc < max1(n)) return 10;
    return 0;
}

// syr2k.mojo: same order as syrk with ldb at position 9 and ldc at 12
int check_syr2k(char uplo, char trans, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc) {
    if (!flag_u(uplo) && !flag_l(uplo)) return 1;
    if (!flag_n(trans) && !flag_t(trans)) return 2;
    if (n < 0) return 3;
    if (k < 0) return 4;
    if (lda < max1(flag_n(trans) ? n : k)) return 7;
    if (ldb < max1(flag_n(trans) ? n : k)) return 9;
    if (ldc < max1(n)) return 12;
    return 0;
}

// symm.mojo:68-86
int check_symm(char side, char uplo, int64_t m, int64_t n, int64_t lda, int64_t ldb, int64_t ldc) {
    if (!flag_l(side) && !flag_r(side)) return 1;
    if (!flag_u(uplo) && !flag_l(uplo)) return 2;
    if (m < 0) return 3;
    if (n < 0) return 4;
    if (lda < max1(flag_l(side) ? m : n)) return 7;
    if (ldb < max1(m)) return 9;
    if (ldc < max1(m)) return 12;
    return 0;
}

int default_precision() {
    std::lock_guard<std::mutex> g(g_prec_mu);
    if (g_default_prec < 0) {
        g_default_prec = MBH_PREC_FP32;
        const char* e = getenv("MOJOBLAS_HIP_PRECISION");
        if (e) {
            if (!strcmp(e, "bf16x3") || !strcmp(e, "BF16X3")) g_default_prec = MBH_PREC_BF16X3;
            else if (!strcmp(e, "fp8") || !strcmp(e, "FP8")) g_default_prec = MBH_PREC_FP8;
        }
    }
    return g_default_prec;
}

static thread_local int t_prec = -1;

// Precision of the host entries on the calling thread: the thread's own
// setting (mbh_set_thread_precision) if it has one, else the library default.
int host_precision() { return t_prec >= 0 ? t_prec : default_precision(); }

bool gemm_precision_ok(int prec) { return prec == MBH_PREC_FP32 || prec == MBH_PREC_BF16X3 || prec == MBH_PREC_FP8; }

// ---- device staging for the host-pointer entries ----------------------------
// Each host thread stages through its own HIP stream (non-blocking, default
// priority) and its own grow-only device buffers, so host calls from
// different threads run concurrently and share nothing; the stream's
// workspace is leased for the call like any other.  A thread's stream and
// buffers are released when the thread exits (the main thread's at process
// exit, by the operating system).
namespace {
constexpr int kMaxDev = 16;
constexpr int kStageSlots = 8;
struct Slot {
    void* p = nullptr;
    size_t cap = 0;
};
struct DevStage {
    hipStream_t s = nullptr;
    Slot slots[kStageSlots];
};
struct HostCtx {
    DevStage dev[kMaxDev];
    ~HostCtx() {
        // The main thread's context dies in exit(), when the HIP runtime may
        // already be going away: leave it to the operating system.
        if (syscall(SYS_gettid) == getpid()) return;
        for (int d = 0; d < kMaxDev; ++d) {
            DevStage& st = dev[d];
            if (!st.s) continue;
            if (hipSetDevice(d) != hipSuccess) continue;
            (void)workspace_release(st.s);
            for (Slot& sl : st.slots)
                if (sl.p) (void)hipFree(sl.p);
            (void)hipStreamDestroy(st.s);
        }
    }
};
thread_local HostCtx t_host;

DevStage* current_stage(hipError_t* err) {
    int d = 0;
    if ((*err = hipGetDevice(&d)) != hipSuccess) return nullptr;
    if (d < 0 || d >= kMaxDev) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    DevStage& st = t_host.dev[d];
    if (!st.s && (*err = hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking)) != hipSuccess) {
        st.s = nullptr;
        return nullptr;
    }
    return &st;
}
}  // namespace

HostStage::HostStage() {
    DevStage* st = current_stage(&err_);
    if (!st) return;
    s_ = st->s;
    lease_.reset(new StreamLease(s_));
    err_ = lease_->error();
}

hipError_t slot_get(int i, size_t bytes, void** out) {
    hipError_t e = hipSuccess;
    DevStage* st = current_stage(&e);
    if (!st) return e;
    if (i < 0 || i >= kStageSlots) return hipErrorInvalidValue;
    Slot& sl = st->slots[i];
    if (bytes == 0) bytes = 16;
    if (sl.cap < bytes) {
        if (sl.p) {
            // earlier work on this thread's stream may still read the buffer
            if ((e = hipStreamSynchronize(st->s)) != hipSuccess) return e;
            if ((e = hipFree(sl.p)) != hipSuccess) return e;
            sl.p = nullptr;
            sl.cap = 0;
        }
        if (inject_alloc_failure()) return hipErrorOutOfMemory;
        if ((e = hipMalloc(&sl.p, bytes)) != hipSuccess) return e;
        sl.cap = bytes;
    }
    *out = sl.p;
    return hipSuccess;
}

// Copy an r x c column-major host matrix (leading dim ld) to a packed device
// buffer (leading dim max(1, r)) on stream s; and back, after which s is
// synchronised so the host buffer is final.
hipError_t put_matrix(void* dev, const void* host, int64_t r, int64_t c, int64_t ld, size_t es, hipStream_t s) {
    if (r <= 0 || c <= 0) return hipSuccess;
    return hipMemcpy2DAsync(dev, (size_t)r * es, host, (size_t)ld * es, (size_t)r * es, (size_t)c,
                            hipMemcpyHostToDevice, s);
}
hipError_t get_matrix(void* host, const void* dev, int64_t r, int64_t c, int64_t ld, size_t es, hipStream_t s) {
    if (r > 0 && c > 0) {
        hipError_t e = hipMemcpy2DAsync(host, (size_t)ld * es, dev, (size_t)r * es, (size_t)r * es, (size_t)c,
                                        hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
    }
    return hipStreamSynchronize(s);
}

template <typename T>
static int gemm_host(char ta, char tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, const T* b,
                     int64_t ldb, T beta, T* c, int64_t ldc, int prec) {
    int info = check_gemm(ta, tb, m, n, k, lda, ldb, ldc);
    if (info) return info;
    if (m == 0 || n == 0) return 0;
    const bool at = !flag_n(ta), bt = !flag_n(tb);
    const bool need_ab = !(alpha == T(0) || k == 0);
    const int64_t ar = at ? k : m, ac = at ? m : k, br = bt ? n : k, bc = bt ? k : n;
    HostStage st;
    if (st.error() != hipSuccess) return hip_fail(st.error(), "host staging");
    const hipStream_t hs = st.stream();
    void *dA = nullptr, *dB = nullptr, *dC = nullptr;
    hipError_t e;
    const size_t es = sizeof(T);
    if ((e = slot_get(2, (size_t)m * n * es, &dC)) != hipSuccess) return hip_fail(e, "hipMalloc(C)");
    if (need_ab) {
        if ((e = slot_get(0, (size_t)ar * ac * es, &dA)) != hipSuccess) return hip_fail(e, "hipMalloc(A)");
        if ((e = slot_get(1, (size_t)br * bc * es, &dB)) != hipSuccess) return hip_fail(e, "hipMalloc(B)");
        if ((e = put_matrix(dA, a, ar, ac, lda, es, hs)) != hipSuccess) return hip_fail(e, "copy A");
        if ((e = put_matrix(dB, b, br, bc, ldb, es, hs)) != hipSuccess) return hip_fail(e, "copy B");
    }
    if (beta != T(0))
        if ((e = put_matrix(dC, c, m, n, ldc, es, hs)) != hipSuccess) return hip_fail(e, "copy C");
    if constexpr (sizeof(T) == 4)
        e = sgemm(at, bt, m, n, k, alpha, (const float*)dA, max1(ar), (const float*)dB, max1(br), beta, (float*)dC, m,
                  prec, hs);
    else
        e = dgemm(at, bt, m, n, k, alpha, (const double*)dA, max1(ar), (const double*)dB, max1(br), beta,
                  (double*)dC, m, hs);
    if (e != hipSuccess) return hip_fail(e, "gemm launch");
    if ((e = get_matrix(c, dC, m, n, ldc, es, hs)) != hipSuccess) return hip_fail(e, "copy C back");
    return 0;
}

// 16-bit (f16 / bf16) gemm on host buffers: same staging as gemm_host.
static int gemm16_host(bool f16, char ta, char tb, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t* a,
                       int64_t lda, const uint16_t* b, int64_t ldb, float beta, uint16_t* c, int64_t ldc) {
    int info = check_gemm(ta, tb, m, n, k, lda, ldb, ldc);
    if (info) return info;
    if (m == 0 || n == 0) return 0;
    const bool at = !flag_n(ta), bt = !flag_n(tb);
    const bool need_ab = !(alpha == 0.0f || k == 0);
    const int64_t ar = at ? k : m, ac = at ? m : k, br = bt ? n : k, bc = bt ? k : n;
    HostStage st;
    if (st.error() != hipSuccess) return hip_fail(st.error(), "host staging");
    const hipStream_t hs = st.stream();
    void *dA = nullptr, *dB = nullptr, *dC = nullptr;
    hipError_t e;
    const size_t es = sizeof(uint16_t);
    if ((e = slot_get(2, (size_t)m * n * es, &dC)) != hipSuccess) return hip_fail(e, "hipMalloc(C)");
    if (need_ab) {
        if ((e = slot_get(0, (size_t)ar * ac * es, &dA)) != hipSuccess) return hip_fail(e, "hipMalloc(A)");
        if ((e = slot_get(1, (size_t)br * bc * es, &dB)) != hipSuccess) return hip_fail(e, "hipMalloc(B)");
        if ((e = put_matrix(dA, a, ar, ac, lda, es, hs)) != hipSuccess) return hip_fail(e, "copy A");
        if ((e = put_matrix(dB, b, br, bc, ldb, es, hs)) != hipSuccess) return hip_fail(e, "copy B");
    }
    if (beta != 0.0f)
        if ((e = put_matrix(dC, c, m, n, ldc, es, hs)) != hipSuccess) return hip_fail(e, "copy C");
    e = gemm16(f16, at, bt, m, n, k, alpha, (const uint16_t*)dA, max1(ar), (const uint16_t*)dB, max1(br), beta,
               (uint16_t*)dC, m, hs);
    if (e != hipSuccess) return hip_fail(e, "gemm launch");
    if ((e = get_matrix(c, dC, m, n, ldc, es, hs)) != hipSuccess) return hip_fail(e, "copy C back");
    return 0;
}

static int gemm16_dev(bool f16, const char* name, char ta, char tb, int64_t m, int64_t n, int64_t k, float alpha,
                      const uint16_t* a, int64_t lda, const uint16_t* b, int64_t ldb, float beta, uint16_t* c,
                      int64_t ldc, void* stream) {
    int info = check_gemm(ta, tb, m, n, k, lda, ldb, ldc);
    if (info) return info;
    if (m == 0 || n == 0) return 0;
    StreamLease lease((hipStream_t)stream);
    if (lease.error() != hipSuccess) return hip_fail(lease.error(), name);
    hipError_t e = gemm16(f16, !flag_n(ta), !flag_n(tb), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                          (hipStream_t)stream);
    return e == hipSuccess ? 0 : hip_fail(e, name);
}

}  // namespace mbh

using namespace mbh;

extern "C" {

const char* mbh_version(void) { return "mojoblas_hip 0.1 (gfx950)"; }
const char* mbh_last_error(void) { return g_err.c_str(); }

int mbh_set_default_precision(int precision) {
    int old = default_precision();
    if (!gemm_precision_ok(precision)) {
        set_error("mbh_set_default_precision: invalid precision %d", precision);
        return MBH_E_PRECISION;
    }
    std::lock_guard<std::mutex> g(g_prec_mu);
    g_default_prec = precision;
    return old;
}
int mbh_get_default_precision(void) { return default_precision(); }

int mbh_set_thread_precision(int precision) {
    if (precision != -1 && !gemm_precision_ok(precision)) {
        set_error("mbh_set_thread_precision: invalid precision %d", precision);
        return MBH_E_PRECISION;
    }
    const int old = t_prec;
    t_prec = precision;
    return old;
}

int mbh_debug_fail_allocation(long n) {
    inject_alloc_failure_after(n);
    return 0;
}

int mbh_device_sync(void) {
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? 0 : hip_fail(e, "hipDeviceSynchronize");
}

int mbh_sgemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float* a,
                   int64_t lda, const float* b, int64_t ldb, float beta, float* c, int64_t ldc) {
    return gemm_host<float>(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, host_precision());
}

int mbh_dgemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                   int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc) {
    return gemm_host<double>(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, MBH_PREC_FP32);
}

int mbh_sgemm_dev(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const float* a,
                  int64_t lda, const float* b, int64_t ldb, float beta, float* c, int64_t ldc, int precision,
                  void* stream) {
    int info = check_gemm(transa, transb, m, n, k, lda, ldb, ldc);
    if (info) return info;
    if (!gemm_precision_ok(precision)) {
        set_error("mbh_sgemm_dev: invalid precision %d", precision);
        return MBH_E_PRECISION;
    }
    if (m == 0 || n == 0) return 0;
    StreamLease lease((hipStream_t)stream);
    if (lease.error() != hipSuccess) return hip_fail(lease.error(), "mbh_sgemm_dev");
    hipError_t e = sgemm(!flag_n(transa), !flag_n(transb), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, precision,
                         (hipStream_t)stream);
    return e == hipSuccess ? 0 : hip_fail(e, "mbh_sgemm_dev");
}

int mbh_dgemm_dev(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                  int64_t lda, const double* b, int64_t ldb, double beta, double* c, int64_t ldc, void* stream) {
    int info = check_gemm(transa, transb, m, n, k, lda, ldb, ldc);
    if (info) return info;
    if (m == 0 || n == 0) return 0;
    StreamLease lease((hipStream_t)stream);
    if (lease.error() != hipSuccess) return hip_fail(lease.error(), "mbh_dgemm_dev");
    hipError_t e = dgemm(!flag_n(transa), !flag_n(transb), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                         (hipStream_t)stream);
    return e == hipSuccess ? 0 : hip_fail(e, "mbh_dgemm_dev");
}

int mbh_hgemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t* a,
                   int64_t lda, const uint16_t* b, int64_t ldb, float beta, uint16_t* c, int64_t ldc) {
    return gemm16_host(true, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

int mbh_bgemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t* a,
                   int64_t lda, const uint16_t* b, int64_t ldb, float beta, uint16_t* c, int64_t ldc) {
    return gemm16_host(false, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

int mbh_hgemm_dev(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t* a,
                  int64_t lda, const uint16_t* b, int64_t ldb, float beta, uint16_t* c, int64_t ldc, void* stream) {
    return gemm16_dev(true, "mbh_hgemm_dev", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}

int mbh_bgemm_dev(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t* a,
                  int64_t lda, const uint16_t* b, int64_t ldb, float beta, uint16_t* c, int64_t ldc, void* stream) {
    return gemm16_dev(false, "mbh_bgemm_dev", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}

}  // extern "C"
