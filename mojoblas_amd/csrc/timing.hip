// Opt-in kernel timing for measurement harnesses: when enabled, the
// split-precision gemm records HIP events on its launch stream around its
// pack kernels and its main kernel, so a caller can read their durations
// without a profiler.  Off by default (no events are recorded).
#include <mutex>
#include <vector>

#include "../../include/mojoblas_hip.h"
#include "abi_internal.hpp"
#include "timing.hpp"

namespace mbh {
namespace {
std::mutex g_mu;
bool g_on = false;
struct Triple {
    hipEvent_t e[3];  // before pack, before main, after main
};
std::vector<Triple> g_done;
std::vector<hipEvent_t> g_free;

hipEvent_t take() {
    if (!g_free.empty()) {
        hipEvent_t e = g_free.back();
        g_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

bool timing_on() {
    std::lock_guard<std::mutex> g(g_mu);
    return g_on;
}

void timing_mark(TimingMark& t, int which, hipStream_t s) {
    if (!t.active) return;
    std::lock_guard<std::mutex> g(g_mu);
    t.e[which] = take();
    if (t.e[which]) (void)hipEventRecord((hipEvent_t)t.e[which], s);
    if (which == 2) {
        Triple tr{{(hipEvent_t)t.e[0], (hipEvent_t)t.e[1], (hipEvent_t)t.e[2]}};
        if (tr.e[0] && tr.e[1] && tr.e[2]) g_done.push_back(tr);
    }
}

}  // namespace mbh

using namespace mbh;

extern "C" {

int mbh_timing_enable(int on) {
    std::lock_guard<std::mutex> g(g_mu);
    int old = g_on ? 1 : 0;
    g_on = on != 0;
    return old;
}

int mbh_timing_read(double* out, int reset) {
    if (!out) {
        set_error("mbh_timing_read: null output");
        return MBH_E_ARG;
    }
    std::lock_guard<std::mutex> g(g_mu);
    double pack = 0.0, main = 0.0;
    for (const Triple& t : g_done) {
        hipError_t e = hipEventSynchronize(t.e[2]);
        if (e != hipSuccess) return hip_fail(e, "mbh_timing_read");
        float a = 0.0f, b = 0.0f;
        if ((e = hipEventElapsedTime(&a, t.e[0], t.e[1])) != hipSuccess) return hip_fail(e, "mbh_timing_read");
        if ((e = hipEventElapsedTime(&b, t.e[1], t.e[2])) != hipSuccess) return hip_fail(e, "mbh_timing_read");
        pack += a;
        main += b;
    }
    out[0] = main;
    out[1] = (double)g_done.size();
    out[2] = pack;
    out[3] = (double)g_done.size();
    if (reset) {
        for (const Triple& t : g_done)
            for (int i = 0; i < 3; ++i) g_free.push_back(t.e[i]);
        g_done.clear();
    }
    return 0;
}

}  // extern "C"
