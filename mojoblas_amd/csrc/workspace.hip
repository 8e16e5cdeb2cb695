// Grow-only device workspace per (device, stream, slot) for kernels that stage
// operands (the split-precision pack pass, level-3 driver scratch).  Growing a
// buffer first synchronises its stream, so work still reading the old buffer
// has finished before it is freed.
#include <map>
#include <mutex>
#include <tuple>

#include "dispatch.hpp"

namespace mbh {

namespace {
struct Buf {
    void* p = nullptr;
    size_t cap = 0;
};
std::mutex g_ws_mu;
std::map<std::tuple<int, hipStream_t, int>, Buf> g_ws;
}  // namespace

hipError_t workspace(hipStream_t s, int slot, size_t bytes, void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> g(g_ws_mu);
    Buf& b = g_ws[std::make_tuple(dev, s, slot)];
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
