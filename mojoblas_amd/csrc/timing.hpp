// Opt-in event timing of the split-precision gemm (see timing.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace mbh {

struct TimingMark {
    bool active = false;
    void* e[3] = {nullptr, nullptr, nullptr};
};

bool timing_on();
// which: 0 = before the pack kernels, 1 = before the main kernel, 2 = after it.
void timing_mark(TimingMark& t, int which, hipStream_t s);

}  // namespace mbh
