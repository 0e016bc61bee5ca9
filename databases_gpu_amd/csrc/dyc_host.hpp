// dyc_host.hpp -- host-side helpers shared by the single-table and group halves of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <utility>

#include "dyc.h"
#include "dyc_device.hpp"
#include "dyc_threads.hpp"

namespace dyc {
namespace host {

// Record a message for dyc_last_error() and return `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();

// Keeps the calling thread's last error message across a step whose failure the caller
// recovers from (a fallback), so that a call that succeeds leaves no message of its own.
class ErrorKeep {
  public:
    ErrorKeep() { snprintf(msg_, sizeof msg_, "%s", last_error()); }
    void restore() const { (void)fail(0, "%s", msg_); }

  private:
    char msg_[512];
};

constexpr uint32_t kMaxBuckets = 1u << 29;  // packed bucket locations use 29 bits

// Wait for a kernel on stream s to post seq to post[0] (host-pinned). Spins for up to spin_us,
// which covers the small launch groups whose latency this saves; past that, or if the kernel
// never posts (a fault), the stream is synchronised, which also returns the stream's error.
inline hipError_t await_post(const volatile unsigned long long *post, unsigned long long seq, hipStream_t s,
                             double spin_us = 2000.0)
{
    const auto t0 = std::chrono::steady_clock::now();
    while (post[0] != seq) {
        _mm_pause();
        if (std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() > spin_us) {
            const hipError_t e = hipStreamSynchronize(s);
            if (e != hipSuccess) return e;
            if (post[0] != seq) return hipErrorUnknown;
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return hipSuccess;
}

// Buckets per subtable of a d-subtable table for a slot-count hint (api.cpp:47-56, which
// divides by TABLE_NUM * BUCKET_SIZE).
uint32_t initial_buckets(uint64_t hint, uint32_t d);

// The subtable an upsize doubles (api.cpp:116-122): the first i with nb[i] == 2 nb[i+1] gives
// i + 1, else subtable 0.
inline uint32_t choose_up(const uint32_t *nb, uint32_t d)
{
    for (uint32_t i = 0; i + 1 < d; ++i)
        if (nb[i] == 2u * nb[i + 1]) return i + 1;
    return 0;
}

// The subtable a downsize halves (api.cpp:147-153): the first i with nb[i] > nb[i+1], else the
// last one.
inline uint32_t choose_down(const uint32_t *nb, uint32_t d)
{
    for (uint32_t i = 0; i + 1 < d; ++i)
        if (nb[i] > nb[i + 1]) return i;
    return d - 1;
}

// The subtable count a configuration asks for (0 means the default).
inline uint32_t config_subtables(const dyc_config &cfg)
{
    return cfg.subtables ? cfg.subtables : (uint32_t)kDefaultSubtables;
}

// Frees device allocations on a helper thread (ReclaimerT, dyc_threads.hpp). A hipFree costs
// about 300 us on MI355X whatever the size (it also waits for the device), and a resize retires
// one subtable; handed to the helper, the free no longer sits in the calling thread's path (a
// helper-thread free slowed a loop of launch + synchronise on the calling thread by about 3 us
// per free).
struct HipFree {
    void operator()(int device, void *p) const
    {
        (void)hipSetDevice(device);
        (void)hipFree(p);
    }
};
using Reclaimer = ReclaimerT<HipFree>;

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace host
}  // namespace dyc

#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return ::dyc::host::fail(DYC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
