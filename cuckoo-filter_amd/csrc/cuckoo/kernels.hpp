// gfx950 kernels of the cuckoo filter and the error-returning launch layer
// shared by the C++ template API (CuckooFilter.hpp) and the C ABI.
//
// Kernel names and argument order follow the reference's free kernel
// templates (CuckooFilter.cuh:82-128, bodies :1119-1266); the filter argument
// is a DeviceView passed by value instead of a host `this` pointer.
//
// Launch shape: grid-stride loops over keys with the grid capped at the
// number of blocks the device keeps resident, so a 2^28-key call is a few
// thousand workgroups rather than a million, and the occupancy counter gets
// one atomic per workgroup instead of one per 256 keys.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "device_ops.hpp"
#include "radix.hpp"

namespace cuckoo {

inline constexpr int kMaxBlock = 1024;

namespace detail {

/// Adds (or subtracts) a per-thread count to a device counter with one
/// global atomic per workgroup.  Every thread of the block must call it.
__device__ __forceinline__ void blockCountAdd(uint32_t local, unsigned long long* target, bool subtract) {
    __shared__ uint32_t sum;
    if (threadIdx.x == 0) sum = 0;
    __syncthreads();
    // wave64 reduction first, then one LDS atomic per wave
    uint32_t v = local;
    const bool fullWaves = (blockDim.x % 64) == 0;
    if (fullWaves) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (__lane_id() == 0 && v) atomicAdd(&sum, v);
    } else if (v) {
        atomicAdd(&sum, v);
    }
    __syncthreads();
    if (threadIdx.x == 0 && sum) {
        const unsigned long long d = subtract ? 0ull - static_cast<unsigned long long>(sum)
                                              : static_cast<unsigned long long>(sum);
        atomicAdd(target, d);
    }
}

}  // namespace detail

/// insertKernel (CuckooFilter.cuh:1119-1141).
template <typename Config>
__global__ __launch_bounds__(kMaxBlock) void insertKernel(const typename Config::KeyType* __restrict__ keys,
                                                          size_t n, DeviceView<Config> v) {
    using Ops = FilterOps<Config>;
    uint32_t ok = 0;
    uint32_t ev = 0;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t chain;
        const bool s = Ops::insert(v, keys[i], chain);
        ok += s;
        ev += chain;
        if (v.chainHist != nullptr && (chain != 0 || !s)) {
            atomicAdd(&v.chainHist[s ? chain : v.maxEvictions + 1], 1u);
        }
    }
    detail::blockCountAdd(ok, v.occupied, false);
    if (v.evictions != nullptr) detail::blockCountAdd(ev, v.evictions, false);
}

namespace detail {

/// The lookup with kLookupLanes (L) consecutive lanes per key, each reading
/// 32 bytes of the key's buckets; a ballot over the group gives the result,
/// which the group's first lane stores.  Virtual thread t handles part t % L
/// of key t / L; with blockDim a multiple of L a group never straddles a wave
/// and every lane of a group takes the same branches.
template <typename Config>
__device__ __forceinline__ void containsGrouped(const typename Config::KeyType* __restrict__ keys,
                                                bool* __restrict__ output, size_t n, const DeviceView<Config>& v) {
    using Ops = FilterOps<Config>;
    using Policy = typename Config::AltBucketPolicy;
    using Tag = typename Config::TagType;
    constexpr uint32_t L = Ops::kLookupLanes;
    const uint64_t group = ((uint64_t(1) << L) - 1) << (__lane_id() & ~(L - 1));
    const size_t total = n * L;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += stride) {
        const size_t i = t / L;
        const uint32_t part = static_cast<uint32_t>(t % L);
        const auto c = Policy::getCandidateBuckets(keys[i], v.numBuckets);
        bool found = (__ballot(Ops::bucketPartHas(v, c.i1, part, c.fp)) & group) != 0;
        if (!found) {
            Tag fp2 = c.fp;
            if constexpr (Policy::usesChoiceBit) fp2 = Ops::altWithNewFp(c.i1, c.fp, v.numBuckets).fp;
            if (c.i2 != c.i1 || fp2 != c.fp) {
                found = (__ballot(Ops::bucketPartHas(v, c.i2, part, fp2)) & group) != 0;
            }
        }
        if (part == 0) output[i] = found;
    }
}

}  // namespace detail

/// containsKernel (CuckooFilter.cuh:1143-1155).  Buckets of 64 and 128 bytes
/// are read by several lanes per key where FilterOps::groupedLookup says so;
/// any grid covers all n keys.
template <typename Config>
__global__ __launch_bounds__(kMaxBlock) void containsKernel(const typename Config::KeyType* __restrict__ keys,
                                                            bool* __restrict__ output, size_t n,
                                                            DeviceView<Config> v) {
    using Ops = FilterOps<Config>;
    if constexpr (Ops::kLookupLanes > 1) {
        if (Ops::groupedLookup(v.numBuckets, blockDim.x)) {  // uniform across the grid
            detail::containsGrouped<Config>(keys, output, n, v);
            return;
        }
    }
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        output[i] = Ops::contains(v, keys[i]);
    }
}

/// deleteKernel (CuckooFilter.cuh:1157-1183); output may be nullptr.
template <typename Config>
__global__ __launch_bounds__(kMaxBlock) void deleteKernel(const typename Config::KeyType* __restrict__ keys,
                                                          bool* __restrict__ output, size_t n,
                                                          DeviceView<Config> v) {
    using Ops = FilterOps<Config>;
    uint32_t ok = 0;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const bool s = Ops::remove(v, keys[i]);
        ok += s;
        if (output != nullptr) output[i] = s;
    }
    detail::blockCountAdd(ok, v.occupied, true);
}

/// PackedTagType of a configuration (CuckooFilter.cuh:159).
template <typename Config>
using PackedTagOf = std::conditional_t<Config::bitsPerTag <= 8, uint32_t, uint64_t>;

/// computePackedTagsKernel (CuckooFilter.cuh:1185-1207): (i1 << bitsPerTag) | fp
/// in PackedTagType (so i1 is truncated exactly as the reference truncates it).
template <typename Config>
__global__ __launch_bounds__(kMaxBlock) void computePackedTagsKernel(
    const typename Config::KeyType* __restrict__ keys, PackedTagOf<Config>* __restrict__ packedTags, size_t n,
    size_t numBuckets) {
    using P = PackedTagOf<Config>;
    using Policy = typename Config::AltBucketPolicy;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const auto c = Policy::getCandidateBuckets(keys[i], numBuckets);
        packedTags[i] = (static_cast<P>(c.i1) << Config::bitsPerTag) | static_cast<P>(c.fp);
    }
}

/// insertKernelSorted (CuckooFilter.cuh:1209-1266); `keys` is unused, as there.
template <typename Config>
__global__ __launch_bounds__(kMaxBlock) void insertKernelSorted(const typename Config::KeyType* /*keys*/,
                                                                const PackedTagOf<Config>* __restrict__ packedTags,
                                                                size_t n, DeviceView<Config> v) {
    using Ops = FilterOps<Config>;
    using Tag = typename Config::TagType;
    using P = PackedTagOf<Config>;
    constexpr P fpMask = static_cast<P>((1ULL << Config::bitsPerTag) - 1);
    uint32_t ok = 0;
    uint32_t ev = 0;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const P packed = packedTags[i];
        uint32_t chain;
        const bool s = Ops::insertFrom(v, static_cast<size_t>(packed >> Config::bitsPerTag),
                                       static_cast<Tag>(packed & fpMask), chain);
        ok += s;
        ev += chain;
        if (v.chainHist != nullptr && (chain != 0 || !s)) {
            atomicAdd(&v.chainHist[s ? chain : v.maxEvictions + 1], 1u);
        }
    }
    detail::blockCountAdd(ok, v.occupied, false);
    if (v.evictions != nullptr) detail::blockCountAdd(ev, v.evictions, false);
}

/// Per-key (i1, i2, fp) exactly as getCandidateBuckets computes them.
template <typename Config>
__global__ __launch_bounds__(kMaxBlock) void candidatesKernel(const typename Config::KeyType* __restrict__ keys,
                                                              size_t n, size_t numBuckets,
                                                              uint64_t* __restrict__ i1, uint64_t* __restrict__ i2,
                                                              uint32_t* __restrict__ fp) {
    using Policy = typename Config::AltBucketPolicy;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const auto c = Policy::getCandidateBuckets(keys[i], numBuckets);
        i1[i] = c.i1;
        i2[i] = c.i2;
        fp[i] = c.fp;
    }
}

/// Device memory the sorted insert needs, grown on demand and kept between
/// calls -- or, when `borrowed`, a fixed buffer the workspace neither grows
/// nor frees (a request beyond it fails with hipErrorOutOfMemory).
struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool borrowed = false;

    hipError_t reserve(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (borrowed) return hipErrorOutOfMemory;
        if (ptr != nullptr) {
            const hipError_t e = hipFree(ptr);
            ptr = nullptr;
            bytes = 0;
            if (e != hipSuccess) return e;
        }
        const hipError_t e = hipMalloc(&ptr, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }

    void release() {
        if (ptr != nullptr && !borrowed) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

namespace detail {

/// Resident workgroups of `kernel` at `block` threads and `lds` bytes of
/// dynamic LDS on device `dev`: the occupancy query runs once per (kernel,
/// block, LDS, device) and thread, not per launch.
inline uint32_t residentBlocks(const void* kernel, uint32_t block, int dev, size_t lds = 0) {
    struct Entry {
        const void* kernel;
        uint32_t block;
        int dev;
        size_t lds;
        uint32_t resident;
    };
    constexpr int kEntries = 64;
    static thread_local Entry cache[kEntries] = {};
    static thread_local int next = 0;
    for (const Entry& e : cache) {
        if (e.kernel == kernel && e.block == block && e.dev == dev && e.lds == lds) return e.resident;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 1;
    int perCu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, static_cast<int>(block), lds) != hipSuccess ||
        perCu <= 0) {
        perCu = 1;
    }
    const uint32_t resident = static_cast<uint32_t>(perCu) * static_cast<uint32_t>(cus);
    cache[next] = Entry{kernel, block, dev, lds, resident};
    next = (next + 1) % kEntries;
    return resident;
}

/// Compute units of device `dev` (queried once per device and thread).
inline uint32_t computeUnits(int dev) {
    constexpr int kDevices = 64;
    static thread_local int cached[kDevices] = {};
    if (dev >= 0 && dev < kDevices && cached[dev] > 0) return static_cast<uint32_t>(cached[dev]);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 1;
    if (dev >= 0 && dev < kDevices) cached[dev] = cus;
    return static_cast<uint32_t>(cus);
}

/// Grid for a grid-stride kernel: enough blocks for n keys, capped at the
/// number of workgroups the device keeps resident.
template <typename Kernel>
inline uint32_t gridFor(Kernel kernel, size_t n, uint32_t block) {
    const size_t need = (n + block - 1) / block;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const size_t cap = residentBlocks(reinterpret_cast<const void*>(kernel), block, dev);
    return static_cast<uint32_t>(std::max<size_t>(1, std::min(need, cap)));
}

inline int bitsFor(size_t nb) {  // ceil(log2(nb))
    int b = 0;
    while ((size_t(1) << b) < nb) ++b;
    return b;
}

}  // namespace detail

}  // namespace cuckoo
