// Instantiates the filter kernels for one alternate-bucket policy and one key
// width (CF_INST_POLICY, CF_INST_KEY) over every valid
// (bitsPerTag, bucketSize, WordType) with bitsPerTag in {8,16,32}, bucketSize
// in {4,8,16,32} and WordType in {uint32_t, uint64_t}.  maxEvictions,
// eviction policy and block size are run-time values in the C ABI.
#include "capi_table.hpp"
#include "cuckoo/CuckooFilter.hpp"

#ifndef CF_INST_POLICY
#error "CF_INST_POLICY must be defined (0 = XOR, 1 = AddSub, 2 = Offset)"
#endif
#ifndef CF_INST_KEY
#error "CF_INST_KEY must be defined (4 or 8)"
#endif

namespace {

#if CF_INST_POLICY == 0
template <typename K, typename Tg, size_t B, size_t S>
using Pol = XorAltBucketPolicy<K, Tg, B, S>;
#elif CF_INST_POLICY == 1
template <typename K, typename Tg, size_t B, size_t S>
using Pol = AddSubAltBucketPolicy<K, Tg, B, S>;
#else
template <typename K, typename Tg, size_t B, size_t S>
using Pol = OffsetAltBucketPolicy<K, Tg, B, S>;
#endif

#if CF_INST_KEY == 8
using KeyT = uint64_t;
#else
using KeyT = uint32_t;
#endif

template <typename Cfg>
struct Entry {
    using Key = typename Cfg::KeyType;
    using Tag = typename Cfg::TagType;
    using View = cuckoo::DeviceView<Cfg>;
    using E = cuckoo::Engine<Cfg>;

    static View view(const cfabi::AnyView& a) {
        View v;
        v.words = static_cast<typename View::Word*>(a.words);
        v.numBuckets = a.numBuckets;
        v.occupied = a.occupied;
        v.evictions = a.evictions;
        v.chainHist = a.chainHist;
        v.maxEvictions = a.maxEvictions;
        v.evictionPolicy = a.evictionPolicy;
        return v;
    }

    static size_t numBuckets(size_t capacity) { return Cfg::AltBucketPolicy::calculateNumBuckets(capacity); }

    static void hostCandidates(size_t nb, const void* keys, size_t n, uint64_t* i1, uint64_t* i2, uint32_t* fp) {
        const Key* k = static_cast<const Key*>(keys);
        for (size_t i = 0; i < n; ++i) {
            const auto c = Cfg::AltBucketPolicy::getCandidateBuckets(k[i], nb);
            i1[i] = c.i1;
            i2[i] = c.i2;
            fp[i] = c.fp;
        }
    }

    static hipError_t insert(const cfabi::AnyView& a, const void* keys, size_t n, uint32_t block, hipStream_t s,
                             cuckoo::Workspace& ws) {
        return E::insert(view(a), static_cast<const Key*>(keys), n, block, s, ws);
    }
    static size_t insertWorkspaceBytes(size_t n, size_t nb) { return E::insertWorkspaceBytes(n, nb); }
    static hipError_t insertSorted(const cfabi::AnyView& a, const void* keys, size_t n, uint32_t block,
                                   hipStream_t s, cuckoo::Workspace& ws) {
        return E::insertSorted(view(a), static_cast<const Key*>(keys), n, block, s, ws);
    }
    static hipError_t contains(const cfabi::AnyView& a, const void* keys, size_t n, bool* out, uint32_t block,
                               hipStream_t s) {
        return E::contains(view(a), static_cast<const Key*>(keys), n, out, block, s);
    }
    static hipError_t remove(const cfabi::AnyView& a, const void* keys, size_t n, bool* out, uint32_t block,
                             hipStream_t s, cuckoo::Workspace& ws) {
        return E::remove(view(a), static_cast<const Key*>(keys), n, out, block, s, ws);
    }
    static hipError_t candidates(const void* keys, size_t n, size_t nb, uint64_t* i1, uint64_t* i2, uint32_t* fp,
                                 uint32_t block, hipStream_t s) {
        return E::candidates(static_cast<const Key*>(keys), n, nb, i1, i2, fp, block, s);
    }
    static hipError_t packedTags(const void* keys, size_t n, size_t nb, void* out, uint32_t block, hipStream_t s) {
        return E::packedTags(static_cast<const Key*>(keys), n, nb, static_cast<typename E::Packed*>(out), block, s);
    }
    static hipError_t sortPacked(void* packed, size_t n, size_t nb, hipStream_t s, cuckoo::Workspace& ws) {
        return E::sortPacked(static_cast<typename E::Packed*>(packed), n, nb, s, ws);
    }
    static hipError_t insertPacked(const cfabi::AnyView& a, const void* packed, size_t n, uint32_t block,
                                   hipStream_t s) {
        return E::insertPacked(view(a), static_cast<const typename E::Packed*>(packed), n, block, s);
    }

    static hipError_t batchExecute(void** state, const cuckoo::BatchOpDesc* ops, size_t n, hipStream_t s,
                                   uint64_t* occupiedOut) {
        if (*state == nullptr) *state = new cuckoo::BatchExecutor<Cfg>();
        return static_cast<cuckoo::BatchExecutor<Cfg>*>(*state)->execute(ops, n, s, occupiedOut);
    }
    static hipError_t batchLaunch(void** state, const cuckoo::BatchOpDesc* ops, size_t n, hipStream_t s,
                                  bool useGraph, uint32_t chunkKeys, bool phased, hipEvent_t beforeLast) {
        if (*state == nullptr) *state = new cuckoo::BatchExecutor<Cfg>();
        return static_cast<cuckoo::BatchExecutor<Cfg>*>(*state)->launch(ops, n, s, useGraph, chunkKeys, phased,
                                                                          beforeLast);
    }
    static const unsigned long long* batchResults(const void* state) {
        return static_cast<const cuckoo::BatchExecutor<Cfg>*>(state)->results();
    }
    static uint32_t batchLevels(const void* state) {
        return static_cast<const cuckoo::BatchExecutor<Cfg>*>(state)->levelsOfLastLaunch();
    }
    static void batchRelease(void* state) { delete static_cast<cuckoo::BatchExecutor<Cfg>*>(state); }

    static cfabi::Ops make() {
        cfabi::Ops o{};
        o.policy = CF_INST_POLICY;
        o.keyBytes = sizeof(Key);
        o.bitsPerTag = Cfg::bitsPerTag;
        o.bucketSize = Cfg::bucketSize;
        o.wordBytes = sizeof(typename Cfg::WordType);
        o.bucketBytes = cuckoo::Layout<Cfg>::bucketBytes;
        o.numBuckets = &numBuckets;
        o.hostCandidates = &hostCandidates;
        o.insert = &insert;
        o.insertWorkspaceBytes = &insertWorkspaceBytes;
        o.insertSorted = &insertSorted;
        o.contains = &contains;
        o.remove = &remove;
        o.candidates = &candidates;
        o.packedTags = &packedTags;
        o.sortPacked = &sortPacked;
        o.insertPacked = &insertPacked;
        o.batchExecute = &batchExecute;
        o.batchLaunch = &batchLaunch;
        o.batchResults = &batchResults;
        o.batchLevels = &batchLevels;
        o.batchRelease = &batchRelease;
        return o;
    }
};

template <size_t Bits, size_t Bucket, typename Word>
using Cfg = CuckooConfig<KeyT, Bits, 500, 256, Bucket, Pol, EvictionPolicy::DFS, Word>;

// (bitsPerTag, bucketSize, WordType) with bucketSize % tagsPerWord == 0.
#define CF_CONFIGS(X)                                                                               \
    X(8, 4, uint32_t) X(8, 8, uint32_t) X(8, 16, uint32_t) X(8, 32, uint32_t)                       \
    X(8, 8, uint64_t) X(8, 16, uint64_t) X(8, 32, uint64_t)                                         \
    X(16, 4, uint32_t) X(16, 8, uint32_t) X(16, 16, uint32_t) X(16, 32, uint32_t)                   \
    X(16, 4, uint64_t) X(16, 8, uint64_t) X(16, 16, uint64_t) X(16, 32, uint64_t)                   \
    X(32, 4, uint32_t) X(32, 8, uint32_t) X(32, 16, uint32_t) X(32, 32, uint32_t)                   \
    X(32, 4, uint64_t) X(32, 8, uint64_t) X(32, 16, uint64_t) X(32, 32, uint64_t)

#define CF_ENTRY(B, S, W) Entry<Cfg<B, S, W>>::make(),

const cfabi::Ops kTable[] = {CF_CONFIGS(CF_ENTRY)};

}  // namespace

#define CF_CAT2(a, b, c) cfabi_table_##a##_##b
#define CF_NAME(P, K) CF_CAT2(P, K, x)

const cfabi::Ops* CF_NAME(CF_INST_POLICY, CF_INST_KEY)(size_t* count) {
    *count = sizeof(kTable) / sizeof(kTable[0]);
    return kTable;
}
