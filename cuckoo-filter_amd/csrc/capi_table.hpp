// Type-erased per-configuration operations behind the C ABI.  Each
// instantiation unit (capi_inst.hip, built once per policy and key width)
// fills a table of these; capi.hip picks the entry matching a cf_config.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "cuckoo/batch.hpp"
#include "cuckoo/engine.hpp"

namespace cfabi {

/// DeviceView with the word pointer erased (same fields, same meaning).
struct AnyView {
    void* words;
    size_t numBuckets;
    unsigned long long* occupied;
    unsigned long long* evictions;
    unsigned int* chainHist;
    uint32_t maxEvictions;
    uint32_t evictionPolicy;
};

struct Ops {
    uint32_t policy, keyBytes, bitsPerTag, bucketSize, wordBytes;
    size_t bucketBytes;
    size_t (*numBuckets)(size_t capacity);
    void (*hostCandidates)(size_t nb, const void* keys, size_t n, uint64_t* i1, uint64_t* i2, uint32_t* fp);
    hipError_t (*insert)(const AnyView&, const void* keys, size_t n, uint32_t block, hipStream_t,
                         cuckoo::Workspace&);
    // device workspace insert() takes for n keys into a table of nb buckets (0 on the direct path)
    size_t (*insertWorkspaceBytes)(size_t n, size_t nb);
    hipError_t (*insertSorted)(const AnyView&, const void* keys, size_t n, uint32_t block, hipStream_t,
                               cuckoo::Workspace&);
    hipError_t (*contains)(const AnyView&, const void* keys, size_t n, bool* out, uint32_t block, hipStream_t);
    hipError_t (*remove)(const AnyView&, const void* keys, size_t n, bool* out, uint32_t block, hipStream_t,
                         cuckoo::Workspace&);
    hipError_t (*candidates)(const void* keys, size_t n, size_t nb, uint64_t* i1, uint64_t* i2, uint32_t* fp,
                             uint32_t block, hipStream_t);
    hipError_t (*packedTags)(const void* keys, size_t n, size_t nb, void* out, uint32_t block, hipStream_t);
    // the presorted route: ascending sort of packed tags, then insertKernelSorted over them
    hipError_t (*sortPacked)(void* packed, size_t n, size_t nb, hipStream_t, cuckoo::Workspace&);
    hipError_t (*insertPacked)(const AnyView&, const void* packed, size_t n, uint32_t block, hipStream_t);
    // batched many-filter execution; *state is this configuration's executor (created on first use)
    hipError_t (*batchExecute)(void** state, const cuckoo::BatchOpDesc* ops, size_t n, hipStream_t,
                               uint64_t* occupiedOut);
    // asynchronous form: launches on the stream; batchResults gives the
    // post-op occupancies once the stream has passed the launch
    hipError_t (*batchLaunch)(void** state, const cuckoo::BatchOpDesc* ops, size_t n, hipStream_t, bool useGraph,
                              uint32_t chunkKeys, bool phased, hipEvent_t beforeLast);
    const unsigned long long* (*batchResults)(const void* state);
    uint32_t (*batchLevels)(const void* state);
    void (*batchRelease)(void* state);
};

}  // namespace cfabi

// One table per (policy, key width); defined in capi_inst.hip.
#define CFABI_TABLE_DECL(P, K) const cfabi::Ops* cfabi_table_##P##_##K(size_t* count)
CFABI_TABLE_DECL(0, 4);
CFABI_TABLE_DECL(0, 8);
CFABI_TABLE_DECL(1, 4);
CFABI_TABLE_DECL(1, 8);
CFABI_TABLE_DECL(2, 4);
CFABI_TABLE_DECL(2, 8);
