// Measurement kernels for the roofline that bounds the cuckoo filter on one
// MI355X (built into libcuckoo_roofline.so; no part of the filter library):
//
//  * rl_gather: per key, read `gathers` uniformly random, aligned blocks of
//    `gather_bytes` (32, 64 or 128) from a table, write one byte.  The same key
//    stream in and byte out as containsMany, with the hashing reduced to one
//    multiply so the memory system is what is measured.  `dependent` makes
//    each further block's address depend on the data of the previous one
//    (the shape of a lookup's second probe).
//  * rl_gather_coop: one random block per key read by 1-8 lanes together,
//    each lane loading its part (fewer load instructions per lane and page).
//  * rl_cas: per key, one relaxed device-scope 64-bit CAS at a random word
//    (the write half of an insert).
//  * rl_cas_xcd: the same CAS, but each workgroup reads the XCD it runs on
//    (hardware register XCC_ID) and only touches that XCD's eighth of the
//    table, so no line is shared between XCDs; with workgroup scope the CAS
//    can then complete in the XCD's own L2 (agent scope for comparison).
//  * rl_gather_region: per key, one random 32-byte read confined to the table
//    region its position in the key array falls in (keys grouped by region,
//    as a lookup whose keys were partitioned by bucket range would see them).
//    Routed: each XCD takes the contiguous eighth of the regions it owns
//    (XCC_ID) from its own work counter, so a region's lines stay in that
//    XCD's L2; otherwise all workgroups take positions in order.
//  * rl_store_bytes: per key, one byte stored at a random position (the
//    output of a lookup or delete whose keys were reordered).
//  * rl_copy: a plain streaming copy, used to calibrate the FETCH_SIZE /
//    WRITE_SIZE counters against a known byte count.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

__device__ __forceinline__ uint64_t mix(uint64_t z) {
    z ^= z >> 31;
    z *= 0x9E3779B97F4A7C15ull;
    return z ^ (z >> 29);
}

using V4 = unsigned int __attribute__((ext_vector_type(4)));

template <int GB, bool NT>
__device__ __forceinline__ unsigned int readBlock(const V4* p) {
    V4 v[GB / 16];
#pragma unroll
    for (int i = 0; i < GB / 16; ++i) v[i] = NT ? __builtin_nontemporal_load(p + i) : p[i];
    V4 acc = v[0];
#pragma unroll
    for (int i = 1; i < GB / 16; ++i) acc ^= v[i];
    return acc.x ^ acc.y ^ acc.z ^ acc.w;
}

template <int GB, bool NT>
__global__ __launch_bounds__(1024) void gatherKernel(const V4* __restrict__ table, uint64_t blockMask,
                                                     const uint64_t* __restrict__ keys, size_t n,
                                                     uint8_t* __restrict__ out, int gathers, int dependent) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        uint64_t h = mix(k);
        unsigned int acc = readBlock<GB, NT>(table + (h & blockMask) * (GB / 16));
        for (int g = 1; g < gathers; ++g) {
            h = mix(h + g + (dependent ? acc : 0u));
            acc ^= readBlock<GB, NT>(table + (h & blockMask) * (GB / 16));
        }
        out[i] = acc == 0x5A5A5A5Au;
    }
}

/// rl_gather with `L` consecutive lanes per key, each reading its GB/L-byte
/// part of the key's block (one translation per lane instead of GB/16), the
/// parts combined with cross-lane shuffles.
template <int GB, int L>
__global__ __launch_bounds__(1024) void gatherCoopKernel(const V4* __restrict__ table, uint64_t blockMask,
                                                         const uint64_t* __restrict__ keys, size_t n,
                                                         uint8_t* __restrict__ out) {
    static_assert(GB / L >= 16 && 64 % L == 0, "a part is at least one 16-byte load");
    const size_t total = n * L;
    const size_t stride = size_t(gridDim.x) * blockDim.x;  // a multiple of 64: a key's lanes move together
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += stride) {
        const size_t i = t / L;
        const uint32_t part = static_cast<uint32_t>(t % L);
        const uint64_t h = mix(keys[i]);
        unsigned int acc = readBlock<GB / L, false>(table + (h & blockMask) * (GB / 16) + part * (GB / L / 16));
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) acc ^= __shfl_xor(acc, off, 64);
        if (part == 0) out[i] = acc == 0x5A5A5A5Au;
    }
}

__global__ __launch_bounds__(1024) void casKernel(unsigned long long* __restrict__ words, uint64_t wordMask,
                                                  const uint64_t* __restrict__ keys, size_t n) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        unsigned long long* p = words + (mix(k) & wordMask);
        unsigned long long expected = 0;
        __hip_atomic_compare_exchange_strong(p, &expected, k | 1ull, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    }
}

/// XCC_ID (hwreg 20, bits [3:0]) of the XCD this wave runs on.
__device__ __forceinline__ unsigned xccId() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20); }

template <int SCOPE>
__global__ __launch_bounds__(1024) void casXcdKernel(unsigned long long* __restrict__ words, uint64_t regionMask,
                                                     int regionShift, unsigned regions,
                                                     const uint64_t* __restrict__ keys, size_t n) {
    unsigned long long* base = words + (static_cast<uint64_t>(xccId() % regions) << regionShift);
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        unsigned long long* p = base + (mix(k) & regionMask);
        unsigned long long expected = 0;
        __hip_atomic_compare_exchange_strong(p, &expected, k | 1ull, __ATOMIC_RELAXED, __ATOMIC_RELAXED, SCOPE);
    }
}

constexpr uint32_t kRegionChunk = 1024;
__device__ uint32_t g_regionNext[8];  // per-XCD work counters of regionGatherKernel

__global__ void zeroRegionCountersKernel() {
    if (threadIdx.x < 8) g_regionNext[threadIdx.x] = 0;
}

// Workgroups take kRegionChunk positions at a time from their queue's counter;
// routed, a workgroup starts at its own XCD's queue and then helps the others,
// so every position is read whatever the placement of workgroups on XCDs.
__global__ __launch_bounds__(256) void regionGatherKernel(const V4* __restrict__ table, uint64_t regionBlocks,
                                                          uint64_t regions, const uint64_t* __restrict__ keys,
                                                          size_t n, uint8_t* __restrict__ out, int routed) {
    __shared__ uint32_t sBase;
    const unsigned queues = routed ? 8u : 1u;
    const unsigned home = routed ? (xccId() & 7u) : 0u;
    for (unsigned q = 0; q < queues; ++q) {
        const unsigned x = (home + q) % queues;
        const size_t lo = n * x / queues;
        const size_t hi = n * (x + 1) / queues;
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) sBase = atomicAdd(&g_regionNext[x], kRegionChunk);
            __syncthreads();
            const size_t b = sBase;
            if (b >= hi - lo) break;
            const size_t end = hi - lo - b < kRegionChunk ? hi - lo - b : kRegionChunk;
            for (size_t j = threadIdx.x; j < end; j += blockDim.x) {
                const size_t i = lo + b + j;
                const uint64_t r = static_cast<uint64_t>(i) * regions / n;
                const uint64_t blk = r * regionBlocks + (mix(keys[i]) & (regionBlocks - 1));
                out[i] = readBlock<32, false>(table + blk * 2) == 0x5A5A5A5Au;
            }
        }
    }
}

__global__ __launch_bounds__(1024) void storeBytesKernel(uint8_t* __restrict__ out, uint64_t mask,
                                                         const uint64_t* __restrict__ keys, size_t n) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        out[mix(k) & mask] = static_cast<uint8_t>(k);
    }
}

__global__ __launch_bounds__(1024) void copyKernel(const V4* __restrict__ src, V4* __restrict__ dst, size_t n16) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

uint32_t gridFor(const void* k, size_t n, int block) {
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, block, 0);
    if (per < 1) per = 1;
    const size_t need = (n + block - 1) / block;
    const size_t cap = size_t(per) * size_t(cus > 0 ? cus : 1);
    return uint32_t(need < cap ? (need ? need : 1) : cap);
}

bool pow2(uint64_t v) { return v && !(v & (v - 1)); }

}  // namespace

extern "C" {

// Returns 0 on success, -1 on a bad argument, else the HIP error code.
// gather_bytes is 32, 64 or 128; nontemporal selects loads with the NT bit.
int rl_gather(const void* table, size_t table_bytes, int gather_bytes, const uint64_t* keys, size_t n,
              uint8_t* out, int gathers, int dependent, int nontemporal, int block, void* stream) {
    if (!table || !keys || !out || (gather_bytes != 32 && gather_bytes != 64 && gather_bytes != 128) ||
        gathers < 1 || block < 64 || block > 1024 || table_bytes % gather_bytes ||
        !pow2(table_bytes / gather_bytes))
        return -1;
    const uint64_t mask = table_bytes / gather_bytes - 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const V4* t = static_cast<const V4*>(table);
#define RL_GATHER(GB, NT)                                                                                   \
    gatherKernel<GB, NT><<<gridFor((const void*)gatherKernel<GB, NT>, n, block), block, 0, s>>>(t, mask, keys, \
                                                                                              n, out, gathers, dependent)
    if (gather_bytes == 32) {
        if (nontemporal) RL_GATHER(32, true); else RL_GATHER(32, false);
    } else if (gather_bytes == 64) {
        if (nontemporal) RL_GATHER(64, true); else RL_GATHER(64, false);
    } else {
        if (nontemporal) RL_GATHER(128, true); else RL_GATHER(128, false);
    }
#undef RL_GATHER
    return static_cast<int>(hipGetLastError());
}

// lanes: 1, 2, 4 or 8 lanes per key, with gather_bytes / lanes at least 16.
int rl_gather_coop(const void* table, size_t table_bytes, int gather_bytes, int lanes, const uint64_t* keys,
                   size_t n, uint8_t* out, int block, void* stream) {
    if (!table || !keys || !out || (gather_bytes != 32 && gather_bytes != 64 && gather_bytes != 128) ||
        (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8) || gather_bytes / lanes < 16 || block < 64 ||
        block > 1024 || block % 64 || table_bytes % gather_bytes || !pow2(table_bytes / gather_bytes))
        return -1;
    const uint64_t mask = table_bytes / gather_bytes - 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const V4* t = static_cast<const V4*>(table);
#define RL_COOP(GB, L)                                                                                          \
    gatherCoopKernel<GB, L><<<gridFor((const void*)gatherCoopKernel<GB, L>, n * L, block), block, 0, s>>>(t, mask, \
                                                                                                    keys, n, out)
    const int key = gather_bytes * 16 + lanes;
    switch (key) {
        case 32 * 16 + 1: RL_COOP(32, 1); break;
        case 32 * 16 + 2: RL_COOP(32, 2); break;
        case 64 * 16 + 1: RL_COOP(64, 1); break;
        case 64 * 16 + 2: RL_COOP(64, 2); break;
        case 64 * 16 + 4: RL_COOP(64, 4); break;
        case 128 * 16 + 1: RL_COOP(128, 1); break;
        case 128 * 16 + 2: RL_COOP(128, 2); break;
        case 128 * 16 + 4: RL_COOP(128, 4); break;
        case 128 * 16 + 8: RL_COOP(128, 8); break;
        default: return -1;
    }
#undef RL_COOP
    return static_cast<int>(hipGetLastError());
}

int rl_cas(void* table, size_t table_bytes, const uint64_t* keys, size_t n, int block, void* stream) {
    if (!table || !keys || block < 64 || block > 1024 || table_bytes % 8 || !pow2(table_bytes / 8)) return -1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    casKernel<<<gridFor((const void*)casKernel, n, block), block, 0, s>>>(
        static_cast<unsigned long long*>(table), table_bytes / 8 - 1, keys, n);
    return static_cast<int>(hipGetLastError());
}

// regions: the table is split into this many equal parts (a power of two),
// one per XCD; workgroup_scope selects the CAS scope (else agent).
int rl_cas_xcd(void* table, size_t table_bytes, unsigned regions, int workgroup_scope, const uint64_t* keys,
               size_t n, int block, void* stream) {
    if (!table || !keys || block < 64 || block > 1024 || table_bytes % 8 || !pow2(table_bytes / 8) ||
        !pow2(regions) || table_bytes / 8 < regions)
        return -1;
    const uint64_t perRegion = table_bytes / 8 / regions;
    int shift = 0;
    while ((uint64_t(1) << shift) < perRegion) ++shift;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto* w = static_cast<unsigned long long*>(table);
    if (workgroup_scope) {
        auto k = casXcdKernel<__HIP_MEMORY_SCOPE_WORKGROUP>;
        k<<<gridFor((const void*)k, n, block), block, 0, s>>>(w, perRegion - 1, shift, regions, keys, n);
    } else {
        auto k = casXcdKernel<__HIP_MEMORY_SCOPE_AGENT>;
        k<<<gridFor((const void*)k, n, block), block, 0, s>>>(w, perRegion - 1, shift, regions, keys, n);
    }
    return static_cast<int>(hipGetLastError());
}

// region_bytes: a power of two, at least 32 bytes, with table_bytes a multiple
// of 8 regions; n below 2^31.
int rl_gather_region(const void* table, size_t table_bytes, size_t region_bytes, const uint64_t* keys, size_t n,
                     uint8_t* out, int routed, int grid, void* stream) {
    if (!table || !keys || !out || n == 0 || n >= (size_t(1) << 31) || grid < 1 || !pow2(region_bytes) ||
        region_bytes < 32 || table_bytes % region_bytes || (table_bytes / region_bytes) % 8)
        return -1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    zeroRegionCountersKernel<<<1, 64, 0, s>>>();
    regionGatherKernel<<<grid, 256, 0, s>>>(static_cast<const V4*>(table), region_bytes / 32,
                                            table_bytes / region_bytes, keys, n, out, routed);
    return static_cast<int>(hipGetLastError());
}

int rl_store_bytes(void* out, size_t out_bytes, const uint64_t* keys, size_t n, int block, void* stream) {
    if (!out || !keys || block < 64 || block > 1024 || !pow2(out_bytes)) return -1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    storeBytesKernel<<<gridFor((const void*)storeBytesKernel, n, block), block, 0, s>>>(
        static_cast<uint8_t*>(out), out_bytes - 1, keys, n);
    return static_cast<int>(hipGetLastError());
}

int rl_copy(const void* src, void* dst, size_t bytes, int block, void* stream) {
    if (!src || !dst || bytes % 16 || block < 64 || block > 1024) return -1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    copyKernel<<<gridFor((const void*)copyKernel, bytes / 16, block), block, 0, s>>>(
        static_cast<const V4*>(src), static_cast<V4*>(dst), bytes / 16);
    return static_cast<int>(hipGetLastError());
}

}  // extern "C"
