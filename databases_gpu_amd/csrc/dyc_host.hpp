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
