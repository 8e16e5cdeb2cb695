// Device workspace per (device, stream): grow-only scratch slots for kernels
// that stage operands (the split-precision pack pass, level-3 driver scratch,
// level-2 partials and flags).
//
// A library call leases its stream's workspace for the whole call
// (StreamLease at the C entry): two host threads that issue calls on one
// stream queue their kernels one call after the other, never interleaved, so
// no call's kernels read another call's scratch.  Growing a slot first
// synchronises the stream, so work still reading the old buffer has finished
// before it is freed.  A slot handed out while the stream was being captured
// into a graph is never freed on growth (the graph keeps its address): it is
// retired and freed by workspace_release.  Growth itself is refused during
// capture (no allocation or synchronisation may be captured).
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "dispatch.hpp"

namespace mbh {

namespace {
constexpr int kSlots = 5;

struct Buf {
    void* p = nullptr;
    size_t cap = 0;
    bool captured = false;  // handed out during graph capture
};

struct StreamWs {
    std::mutex mu;  // held by the thread whose call leases this stream
    Buf slot[kSlots];
    std::vector<std::pair<void*, size_t>> retired;
    int* trsv_fail = nullptr;  // status word of the last trsv queued here
};

std::mutex g_map_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<StreamWs>> g_ws;
thread_local StreamWs* t_held = nullptr;

StreamWs* state(int dev, hipStream_t s) {
    std::lock_guard<std::mutex> g(g_map_mu);
    std::unique_ptr<StreamWs>& u = g_ws[std::make_pair(dev, s)];
    if (!u) u.reset(new StreamWs());
    return u.get();
}

hipError_t stream_device(hipStream_t s, int* dev) {
    if (s) return hipStreamGetDevice(s, dev);
    return hipGetDevice(dev);
}

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}
std::atomic<long> g_fail_countdown{-1};
}  // namespace

void inject_alloc_failure_after(long n) { g_fail_countdown.store(n < 0 ? -1 : n); }

bool inject_alloc_failure() {
    if (g_fail_countdown.load(std::memory_order_relaxed) < 0) return false;
    return g_fail_countdown.fetch_sub(1) == 0;  // exactly one failure, then off
}

StreamLease::StreamLease(hipStream_t s) {
    int dev = 0;
    if ((err_ = stream_device(s, &dev)) != hipSuccess) return;
    StreamWs* w = state(dev, s);
    if (t_held == w) return;  // nested call on the same stream: already leased
    w->mu.lock();
    prev_ = t_held;
    t_held = w;
    held_ = w;
}

StreamLease::~StreamLease() {
    if (!held_) return;
    StreamWs* w = static_cast<StreamWs*>(held_);
    t_held = static_cast<StreamWs*>(prev_);
    w->mu.unlock();
}

hipError_t workspace(hipStream_t s, int slot, size_t bytes, void** out) {
    int dev = 0;
    hipError_t e = stream_device(s, &dev);
    if (e != hipSuccess) return e;
    StreamWs* w = state(dev, s);
    if (t_held != w || slot < 0 || slot >= kSlots) return hipErrorInvalidHandle;  // caller holds no lease
    Buf& b = w->slot[slot];
    const bool cap = capturing(s);
    if (b.cap < bytes) {
        if (cap) return hipErrorStreamCaptureUnsupported;
        if (b.p) {
            if (b.captured) {
                w->retired.emplace_back(b.p, b.cap);
            } else {
                if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
                if ((e = hipFree(b.p)) != hipSuccess) return e;
            }
            b.p = nullptr;
            b.cap = 0;
            b.captured = false;
        }
        size_t want = bytes + bytes / 8;  // headroom against repeated small growth
        if (inject_alloc_failure()) return hipErrorOutOfMemory;
        if ((e = hipMalloc(&b.p, want)) != hipSuccess) return e;
        b.cap = want;
    }
    if (cap) b.captured = true;
    *out = b.p;
    return hipSuccess;
}

void workspace_note_trsv(hipStream_t s, int* fail) {
    int dev = 0;
    if (stream_device(s, &dev) != hipSuccess) return;
    StreamWs* w = state(dev, s);
    if (t_held == w) w->trsv_fail = fail;
}

hipError_t workspace_trsv_status(hipStream_t s, int* failed) {
    *failed = 0;
    int dev = 0;
    hipError_t e = stream_device(s, &dev);
    if (e != hipSuccess) return e;
    StreamLease lease(s);
    if (lease.error() != hipSuccess) return lease.error();
    StreamWs* w = state(dev, s);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if (!w->trsv_fail) return hipSuccess;
    return hipMemcpy(failed, w->trsv_fail, sizeof(int), hipMemcpyDeviceToHost);
}

hipError_t workspace_release(hipStream_t s) {
    int dev = 0;
    hipError_t e = stream_device(s, &dev);
    if (e != hipSuccess) return e;
    StreamLease lease(s);
    if (lease.error() != hipSuccess) return lease.error();
    StreamWs* w = state(dev, s);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    hipError_t first = hipSuccess;
    for (Buf& b : w->slot) {
        if (b.p) {
            hipError_t f = hipFree(b.p);
            if (first == hipSuccess) first = f;
        }
        b = Buf();
    }
    for (auto& r : w->retired) {
        hipError_t f = hipFree(r.first);
        if (first == hipSuccess) first = f;
    }
    w->retired.clear();
    w->trsv_fail = nullptr;
    return first;
}

hipError_t workspace_bytes(hipStream_t s, size_t* bytes) {
    *bytes = 0;
    int dev = 0;
    hipError_t e = stream_device(s, &dev);
    if (e != hipSuccess) return e;
    StreamLease lease(s);
    if (lease.error() != hipSuccess) return lease.error();
    StreamWs* w = state(dev, s);
    for (const Buf& b : w->slot) *bytes += b.cap;
    for (const auto& r : w->retired) *bytes += r.second;
    return hipSuccess;
}

}  // namespace mbh
