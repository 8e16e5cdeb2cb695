// Per-(device, stream) grow-only device workspace for kernels that stage
// operands (the split-precision pack pass, blocked trsm scratch).  Growing a
// buffer first synchronises its stream, so work still reading the old buffer
// has finished before it is freed.
#include <map>
#include <mutex>
#include <utility>

#include "dispatch.hpp"

namespace mbh {

namespace {
struct Buf {
    void* p = nullptr;
    size_t cap = 0;
};
std::mutex g_ws_mu;
std::map<std::pair<int, hipStream_t>, Buf> g_ws;
}  // namespace

hipError_t workspace(hipStream_t s, size_t bytes, void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> g(g_ws_mu);
    Buf& b = g_ws[std::make_pair(dev, s)];
    if (b.cap < bytes) {
        if (b.p) {
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
            if ((e = hipFree(b.p)) != hipSuccess) return e;
            b.p = nullptr;
            b.cap = 0;
        }
        size_t want = bytes + bytes / 8;  // headroom against repeated small growth
        if ((e = hipMalloc(&b.p, want)) != hipSuccess) return e;
        b.cap = want;
    }
    *out = b.p;
    return hipSuccess;
}

}  // namespace mbh
