// hiphost.h -- host-side HIP helpers shared by the library's .hip files:
// status-returning error checks, device allocation, event timing.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "host.h"

namespace fornos {
namespace {

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            return set_error(e_ == hipErrorOutOfMemory ? FORNOS_OOM : FORNOS_HIP_ERROR,   \
                             std::string(#expr) + ": " + hipGetErrorString(e_));          \
        }                                                                                 \
    } while (0)

template <typename T>
fornos_status dev_alloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return set_error(e == hipErrorOutOfMemory ? FORNOS_OOM : FORNOS_HIP_ERROR,
                         std::string("hipMalloc(") + std::to_string(sizeof(T) * count) +
                             " bytes): " + hipGetErrorString(e));
    }
    return FORNOS_OK;
}

void dev_free(void* p)
{
    if (p) (void)hipFree(p);
}

fornos_status require_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return set_error(FORNOS_UNSUPPORTED, "no HIP device is visible");
    }
    return FORNOS_OK;
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st) {}
    fornos_status start()
    {
        HIP_TRY(hipEventCreate(&a));
        HIP_TRY(hipEventCreate(&b));
        HIP_TRY(hipEventRecord(a, s));
        return FORNOS_OK;
    }
    fornos_status stop(float* ms)
    {
        HIP_TRY(hipEventRecord(b, s));
        HIP_TRY(hipEventSynchronize(b));
        float v = 0;
        HIP_TRY(hipEventElapsedTime(&v, a, b));
        if (ms) *ms = v;
        return FORNOS_OK;
    }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Completion marks of work submitted without waiting: an event recorded
// after each part, with the amount of work finished once it has completed.
// poll() queries the events in order and never blocks.
struct Marks {
    std::vector<hipEvent_t> ev;
    std::vector<uint64_t> work;  // work finished when ev[i] has completed
    size_t reached = 0;          // leading events known to have completed
    ~Marks() { clear(); }
    void clear()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
        work.clear();
        reached = 0;
    }
    bool empty() const { return ev.empty(); }
    bool complete() const { return reached == ev.size(); }
    fornos_status add(hipStream_t st, uint64_t w)
    {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ev.push_back(e);
        work.push_back(w);
        HIP_TRY(hipEventRecord(e, st));
        return FORNOS_OK;
    }
    // Work finished so far (0 before the first mark).
    fornos_status poll(uint64_t* finished)
    {
        while (reached < ev.size()) {
            const hipError_t r = hipEventQuery(ev[reached]);
            if (r == hipErrorNotReady) break;
            HIP_TRY(r);
            ++reached;
        }
        *finished = reached ? work[reached - 1] : 0;
        return FORNOS_OK;
    }
};

}  // namespace
}  // namespace fornos
