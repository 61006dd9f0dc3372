// MI355X-native cuckoo filter: the C++ template drop-in for the reference's
// single-GPU API (include/CuckooFilter.cuh).
//
//   CuckooConfig<T, bitsPerTag, maxEvictions = 500, blockSize = 256,
//                bucketSize = 16, AltBucketPolicy = XorAltBucketPolicy,
//                evictionPolicy = EvictionPolicy::DFS, WordType = uint64_t>
//   CuckooFilter<Config>:
//       size_t insertMany(const T* d_keys, size_t n, hipStream_t = {});
//       size_t insertManySorted(const T* d_keys, size_t n, hipStream_t = {});
//       void   containsMany(const T* d_keys, size_t n, bool* d_output, hipStream_t = {});
//       size_t deleteMany(const T* d_keys, size_t n, bool* d_output = nullptr, hipStream_t = {});
//       clear, loadFactor, occupiedSlots, capacity, getNumBuckets, sizeInBytes,
//       countOccupiedSlots, evictionCount/resetEvictionCount
//
// Names, template defaults (DFS eviction), argument order and return values
// (insert/delete return the total occupied slots after the call) are the
// reference's; hipStream_t replaces cudaStream_t.  Every call is synchronous
// on its stream like the reference's.  Errors print file:line and exit
// (HIP_CALL), as the reference's CUDA_CALL does.
//
// The table image is the reference's: numBuckets buckets of
// bucketSize*sizeof(TagType) bytes, each a little-endian array of WordType
// words with tag slot s in word s/tagsPerWord at bits [(s%tpw)*bits, +bits).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "bucket_policies.hpp"
#include "device_ops.hpp"
#include "hashutil.hpp"
#include "helpers.hpp"
#include "engine.hpp"
#include "kernels.hpp"

#if !defined(CUCKOO_FILTER_NO_THRUST) && __has_include(<thrust/device_vector.h>)
#include <thrust/device_vector.h>
#define CUCKOO_FILTER_HAS_THRUST 1
#endif

/**
 * @brief Compile-time configuration (CuckooFilter.cuh:44-74).
 */
template <typename T,
          size_t bitsPerTag_,
          size_t maxEvictions_ = 500,
          size_t blockSize_ = 256,
          size_t bucketSize_ = 16,
          template <typename, typename, size_t, size_t> class AltBucketPolicy_ = XorAltBucketPolicy,
          EvictionPolicy evictionPolicy_ = EvictionPolicy::DFS,
          typename WordType_ = uint64_t>
struct CuckooConfig {
    using KeyType = T;
    static constexpr size_t bitsPerTag = bitsPerTag_;
    static constexpr size_t maxEvictions = maxEvictions_;
    static constexpr size_t blockSize = blockSize_;
    static constexpr size_t bucketSize = bucketSize_;
    static constexpr EvictionPolicy evictionPolicy = evictionPolicy_;

    using TagType = std::conditional_t<bitsPerTag <= 8, uint8_t,
                                       std::conditional_t<bitsPerTag <= 16, uint16_t, uint32_t>>;

    using WordType = WordType_;
    static_assert(std::is_same_v<WordType, uint32_t> || std::is_same_v<WordType, uint64_t>,
                  "WordType must be uint32_t or uint64_t");
    static_assert(sizeof(TagType) <= sizeof(WordType), "TagType must fit within WordType");
    static_assert(blockSize_ >= 1 && blockSize_ <= 1024, "blockSize must be in [1, 1024]");

    using AltBucketPolicy = AltBucketPolicy_<KeyType, TagType, bitsPerTag, bucketSize_>;
};

using cuckoo::computePackedTagsKernel;
using cuckoo::containsKernel;
using cuckoo::deleteKernel;
using cuckoo::insertKernel;
using cuckoo::insertKernelSorted;
using cuckoo::insertKernelSortedSmall;

/**
 * @brief GPU cuckoo filter over one HIP device (CuckooFilter.cuh:139-1117).
 */
template <typename Config>
struct CuckooFilter {
    using T = typename Config::KeyType;
    static constexpr size_t bitsPerTag = Config::bitsPerTag;

    using TagType = typename Config::TagType;
    using AltBucketPolicy = typename Config::AltBucketPolicy;

    static constexpr size_t tagEntryBytes = sizeof(TagType);
    static constexpr size_t bucketSize = Config::bucketSize;
    static constexpr size_t maxEvictions = Config::maxEvictions;
    static constexpr size_t blockSize = Config::blockSize;

    static_assert(bitsPerTag == 8 || bitsPerTag == 16 || bitsPerTag == 32, "The tag must be 8, 16 or 32 bits");
    static_assert(powerOfTwo(bucketSize), "Bucket size must be a power of 2");

    using PackedTagType = cuckoo::PackedTagOf<Config>;

    /// (primary bucket << bitsPerTag) | fp, sortable by bucket (CuckooFilter.cuh:161-207).
    struct PackedTag {
        PackedTagType value;

        static constexpr size_t fpBits = bitsPerTag;
        static constexpr size_t totalBits = sizeof(PackedTagType) * 8;
        static constexpr size_t bucketIdxBits = totalBits - fpBits;
        static_assert(fpBits < totalBits, "fpBits must leave at least some bits for bucketIdx");
        static constexpr PackedTagType fpMask = PackedTagType((1ULL << fpBits) - 1ULL);
        static constexpr PackedTagType bucketIdxMask =
            PackedTagType(((bucketIdxBits >= 64 ? ~0ULL : ((1ULL << bucketIdxBits) - 1ULL))) << fpBits);

        __host__ __device__ PackedTag() : value(0) {}
        __host__ __device__ PackedTag(TagType fp, uint64_t bucketIdx) : value(0) {
            setFingerprint(fp);
            setBucketIdx(bucketIdx);
        }
        __host__ __device__ TagType getFingerprint() const { return static_cast<TagType>(value & fpMask); }
        __host__ __device__ uint64_t getBucketIndex() const { return uint64_t((value & bucketIdxMask) >> fpBits); }
        __host__ __device__ void setFingerprint(TagType fp) {
            value = (value & ~fpMask) | (static_cast<PackedTagType>(fp) & fpMask);
        }
        __host__ __device__ void setBucketIdx(size_t bucketIdx) {
            value = (value & ~bucketIdxMask) | (static_cast<PackedTagType>(bucketIdx) << fpBits);
        }
    };

    static constexpr TagType EMPTY = 0;
    static constexpr size_t fpMask = (1ULL << bitsPerTag) - 1;

    using Ops = cuckoo::FilterOps<Config>;
    using View = cuckoo::DeviceView<Config>;

    /**
     * @brief One bucket: wordCount packed words (CuckooFilter.cuh:224-249).
     * Plain words in device memory; every device write goes through a
     * relaxed device-scope CAS.
     */
    struct Bucket {
        using WordType = typename Config::WordType;
        static constexpr size_t tagsPerWord = cuckoo::Layout<Config>::tagsPerWord;
        static constexpr size_t wordCount = cuckoo::Layout<Config>::wordCount;

        WordType packedTags[wordCount];

        __host__ __device__ TagType extractTag(WordType packed, size_t tagIdx) const {
            return Ops::extractTag(packed, tagIdx);
        }
        __host__ __device__ WordType replaceTag(WordType packed, size_t tagIdx, TagType newTag) const {
            return Ops::replaceTag(packed, tagIdx, newTag);
        }
    };
    static_assert(sizeof(Bucket) == bucketSize * sizeof(TagType), "bucket layout must be the reference's");

    size_t numBuckets;                               ///< Number of buckets in the filter
    Bucket* d_buckets = nullptr;                     ///< Device memory of the buckets
    unsigned long long* d_numOccupied = nullptr;     ///< Device occupancy counter
#ifdef CUCKOO_FILTER_COUNT_EVICTIONS
    unsigned long long* d_numEvictions = nullptr;    ///< Device eviction counter
#endif
    size_t h_numOccupied = 0;                        ///< Last value read back

    template <typename H>
    static __host__ __device__ uint64_t hash64(const H& key) {
        return AltBucketPolicy::hash64(key);
    }

    static __host__ __device__ cuckoo::CandidateBuckets<TagType> getCandidateBuckets(const T& key, size_t nb) {
        return AltBucketPolicy::getCandidateBuckets(key, nb);
    }

    static __host__ __device__ size_t getAlternateBucket(size_t bucket, TagType fp, size_t nb) {
        return Ops::altWithNewFp(bucket, fp, nb).bucket;
    }

    static __host__ __device__ cuckoo::AltBucket<TagType> getAlternateBucketWithNewFp(size_t bucket, TagType fp,
                                                                                     size_t nb) {
        return Ops::altWithNewFp(bucket, fp, nb);
    }

    static size_t calculateNumBuckets(size_t capacity) { return AltBucketPolicy::calculateNumBuckets(capacity); }

    CuckooFilter(const CuckooFilter&) = delete;
    CuckooFilter& operator=(const CuckooFilter&) = delete;

    explicit CuckooFilter(size_t capacity) : numBuckets(calculateNumBuckets(capacity)) {
        HIP_CALL(hipMalloc(&d_buckets, numBuckets * sizeof(Bucket)));
        HIP_CALL(hipMalloc(&d_numOccupied, sizeof(unsigned long long)));
#ifdef CUCKOO_FILTER_COUNT_EVICTIONS
        HIP_CALL(hipMalloc(&d_numEvictions, sizeof(unsigned long long)));
#endif
        clear();
    }

    ~CuckooFilter() {
        workspace_.release();
        if (d_buckets) HIP_CALL(hipFree(d_buckets));
        if (d_numOccupied) HIP_CALL(hipFree(d_numOccupied));
#ifdef CUCKOO_FILTER_COUNT_EVICTIONS
        if (d_numEvictions) HIP_CALL(hipFree(d_numEvictions));
#endif
    }

    /// Trivially-copyable device view passed to the kernels.
    View view() const {
        View v{};
        v.words = reinterpret_cast<typename View::Word*>(d_buckets);
        v.numBuckets = numBuckets;
        v.occupied = d_numOccupied;
#ifdef CUCKOO_FILTER_COUNT_EVICTIONS
        v.evictions = d_numEvictions;
#else
        v.evictions = nullptr;
#endif
        v.chainHist = nullptr;
        v.maxEvictions = static_cast<uint32_t>(maxEvictions);
        v.evictionPolicy = static_cast<uint32_t>(Config::evictionPolicy);
        return v;
    }

    /// Inserts n keys; returns the total number of occupied slots afterwards.
    size_t insertMany(const T* d_keys, const size_t n, hipStream_t stream = {}) {
        HIP_CALL((cuckoo::Engine<Config>::insert(view(), d_keys, n, blockSize, stream, workspace_)));
        HIP_CALL(hipStreamSynchronize(stream));
        return occupiedSlots();
    }

    /// Inserts n keys after ordering them by primary bucket.
    size_t insertManySorted(const T* d_keys, const size_t n, hipStream_t stream = {}) {
        HIP_CALL((cuckoo::Engine<Config>::insertSorted(view(), d_keys, n, blockSize, stream, workspace_)));
        HIP_CALL(hipStreamSynchronize(stream));
        return occupiedSlots();
    }

    /// d_output[i] = whether key i may be in the filter.
    void containsMany(const T* d_keys, const size_t n, bool* d_output, hipStream_t stream = {}) {
        HIP_CALL((cuckoo::Engine<Config>::contains(view(), d_keys, n, d_output, blockSize, stream)));
        HIP_CALL(hipStreamSynchronize(stream));
    }

    /// Removes one copy of each key's fingerprint; d_output (optional) gets per-key success.
    size_t deleteMany(const T* d_keys, const size_t n, bool* d_output = nullptr, hipStream_t stream = {}) {
        HIP_CALL((cuckoo::Engine<Config>::remove(view(), d_keys, n, d_output, blockSize, stream, workspace_)));
        HIP_CALL(hipStreamSynchronize(stream));
        return occupiedSlots();
    }

#ifdef CUCKOO_FILTER_HAS_THRUST
    size_t insertMany(const thrust::device_vector<T>& d_keys, hipStream_t stream = {}) {
        return insertMany(thrust::raw_pointer_cast(d_keys.data()), d_keys.size(), stream);
    }

    size_t insertManySorted(const thrust::device_vector<T>& d_keys, hipStream_t stream = {}) {
        return insertManySorted(thrust::raw_pointer_cast(d_keys.data()), d_keys.size(), stream);
    }

    void containsMany(const thrust::device_vector<T>& d_keys, thrust::device_vector<bool>& d_output,
                      hipStream_t stream = {}) {
        if (d_output.size() != d_keys.size()) d_output.resize(d_keys.size());
        containsMany(thrust::raw_pointer_cast(d_keys.data()), d_keys.size(),
                     thrust::raw_pointer_cast(d_output.data()), stream);
    }

    void containsMany(const thrust::device_vector<T>& d_keys, thrust::device_vector<uint8_t>& d_output,
                      hipStream_t stream = {}) {
        if (d_output.size() != d_keys.size()) d_output.resize(d_keys.size());
        containsMany(thrust::raw_pointer_cast(d_keys.data()), d_keys.size(),
                     reinterpret_cast<bool*>(thrust::raw_pointer_cast(d_output.data())), stream);
    }

    size_t deleteMany(const thrust::device_vector<T>& d_keys, thrust::device_vector<bool>& d_output,
                      hipStream_t stream = {}) {
        if (d_output.size() != d_keys.size()) d_output.resize(d_keys.size());
        return deleteMany(thrust::raw_pointer_cast(d_keys.data()), d_keys.size(),
                          thrust::raw_pointer_cast(d_output.data()), stream);
    }

    size_t deleteMany(const thrust::device_vector<T>& d_keys, thrust::device_vector<uint8_t>& d_output,
                      hipStream_t stream = {}) {
        if (d_output.size() != d_keys.size()) d_output.resize(d_keys.size());
        return deleteMany(thrust::raw_pointer_cast(d_keys.data()), d_keys.size(),
                          reinterpret_cast<bool*>(thrust::raw_pointer_cast(d_output.data())), stream);
    }

    size_t deleteMany(const thrust::device_vector<T>& d_keys, hipStream_t stream = {}) {
        return deleteMany(thrust::raw_pointer_cast(d_keys.data()), d_keys.size(), nullptr, stream);
    }
#endif

    /// Empties the filter (CuckooFilter.cuh:675-683).
    void clear() {
        HIP_CALL(hipMemset(d_buckets, 0, numBuckets * sizeof(Bucket)));
        HIP_CALL(hipMemset(d_numOccupied, 0, sizeof(unsigned long long)));
#ifdef CUCKOO_FILTER_COUNT_EVICTIONS
        HIP_CALL(hipMemset(d_numEvictions, 0, sizeof(unsigned long long)));
#endif
        h_numOccupied = 0;
    }

    [[nodiscard]] float loadFactor() {
        return static_cast<float>(occupiedSlots()) / (numBuckets * bucketSize);
    }

    size_t occupiedSlots() {
        unsigned long long v = 0;
        HIP_CALL(hipMemcpy(&v, d_numOccupied, sizeof(v), hipMemcpyDeviceToHost));
        h_numOccupied = static_cast<size_t>(v);
        return h_numOccupied;
    }

#ifdef CUCKOO_FILTER_COUNT_EVICTIONS
    size_t evictionCount() {
        unsigned long long v = 0;
        HIP_CALL(hipMemcpy(&v, d_numEvictions, sizeof(v), hipMemcpyDeviceToHost));
        return static_cast<size_t>(v);
    }

    void resetEvictionCount() { HIP_CALL(hipMemset(d_numEvictions, 0, sizeof(unsigned long long))); }
#endif

    size_t capacity() { return numBuckets * bucketSize; }

    [[nodiscard]] size_t getNumBuckets() const { return numBuckets; }

    [[nodiscard]] size_t sizeInBytes() const { return numBuckets * sizeof(Bucket); }

    /// Counts non-empty slots by copying the table to the host (CuckooFilter.cuh:759-785).
    size_t countOccupiedSlots() {
        std::vector<Bucket> h(numBuckets);
        HIP_CALL(hipMemcpy(h.data(), d_buckets, numBuckets * sizeof(Bucket), hipMemcpyDeviceToHost));
        size_t count = 0;
        for (const Bucket& b : h) {
            for (size_t w = 0; w < Bucket::wordCount; ++w) {
                for (size_t t = 0; t < Bucket::tagsPerWord; ++t) count += b.extractTag(b.packedTags[w], t) != EMPTY;
            }
        }
        return count;
    }

   private:
    cuckoo::Workspace workspace_;
};
