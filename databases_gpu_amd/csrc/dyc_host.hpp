// dyc_host.hpp -- host-side helpers shared by the single-table and group halves of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dyc.h"
#include "dyc_device.hpp"

namespace dyc {
namespace host {

// Record a message for dyc_last_error() and return `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();

constexpr uint32_t kMaxBuckets = 1u << 30;  // packed bucket locations use 30 bits

// Buckets per subtable for a slot-count hint (api.cpp:47-56).
uint32_t initial_buckets(uint64_t hint);

// The subtable an upsize doubles (api.cpp:116-122): the first i with nb[i] == 2 nb[i+1] gives
// i + 1, else subtable 0.
inline uint32_t choose_up(const uint32_t nb[4])
{
    for (uint32_t i = 0; i + 1 < kSubtables; ++i)
        if (nb[i] == 2u * nb[i + 1]) return i + 1;
    return 0;
}

// The subtable a downsize halves (api.cpp:147-153): the first i with nb[i] > nb[i+1], else 3.
inline uint32_t choose_down(const uint32_t nb[4])
{
    for (uint32_t i = 0; i + 1 < kSubtables; ++i)
        if (nb[i] > nb[i + 1]) return i;
    return kSubtables - 1;
}

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
