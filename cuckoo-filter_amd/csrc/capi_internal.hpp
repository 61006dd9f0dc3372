// Pieces of the C ABI shared by its translation units (capi.hip, server.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/cuckoo_filter.h"
#include "capi_table.hpp"

namespace cfabi::internal {

inline thread_local std::string g_lastError;

inline int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_lastError = buf;
    return code;
}

inline int hipFail(hipError_t e, const char* what) {
    return fail(CF_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), static_cast<int>(e));
}

/// Makes the filter's device current for the duration of a call.
class DeviceGuard {
   public:
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev_) != hipSuccess) prev_ = -1;
        if (prev_ != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev_ >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev_) (void)hipSetDevice(prev_);
    }
    DeviceGuard(const DeviceGuard&) = delete;
    DeviceGuard& operator=(const DeviceGuard&) = delete;

   private:
    int prev_ = -1;
};

}  // namespace cfabi::internal

struct cf_filter {
    cf_config cfg{};
    const cfabi::Ops* ops = nullptr;
    int device = 0;
    size_t numBuckets = 0;
    void* buckets = nullptr;
    unsigned long long* counters = nullptr;  // [0] occupied, [1] evictions
    unsigned int* hist = nullptr;            // max_evictions + 2 bins, or nullptr

    size_t tableBytes() const { return numBuckets * ops->bucketBytes; }

    cfabi::AnyView view() const {
        cfabi::AnyView v;
        v.words = buckets;
        v.numBuckets = numBuckets;
        v.occupied = counters;
        v.evictions = counters + 1;
        v.chainHist = hist;
        v.maxEvictions = cfg.max_evictions;
        v.evictionPolicy = cfg.eviction;
        return v;
    }
};


namespace cfabi::internal {

/// The batch descriptor of one operation on `f`.
inline cuckoo::BatchOpDesc describe(const cf_filter* f, uint32_t op, const void* keys, size_t n, bool* out) {
    cuckoo::BatchOpDesc d{};
    d.words = f->buckets;
    d.numBuckets = f->numBuckets;
    d.occupied = f->counters;
    d.evictions = f->counters + 1;
    d.chainHist = f->hist;
    d.keys = keys;
    d.out = op == CF_OP_INSERT ? nullptr : out;
    d.n = n;
    d.kind = op;
    d.maxEvictions = f->cfg.max_evictions;
    d.evictionPolicy = f->cfg.eviction;
    d.pad = 0;
    return d;
}

}  // namespace cfabi::internal
