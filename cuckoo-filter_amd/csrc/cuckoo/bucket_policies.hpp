// Alternate-bucket policies: how a key maps to its fingerprint and its two
// candidate buckets.  Same template signature and static interface as the
// reference (include/bucket_policies.cuh), so a CuckooConfig naming a policy
// compiles unchanged:
//
//   template <KeyType, TagType, bitsPerTag, bucketSize> struct Policy {
//       static constexpr size_t fpMask; static constexpr bool usesChoiceBit;
//       static uint64_t hash64(const H&);
//       static CandidateBuckets<TagType> getCandidateBuckets(const KeyType&, size_t nb);
//       static size_t getAlternateBucket(size_t bucket, TagType fp, size_t nb);
//       static size_t calculateNumBuckets(size_t capacity);
//   };
//
// getCandidateBuckets returns an aggregate {i1, i2, fp}, so the reference's
// `auto [i1, i2, fp] = ...` structured bindings keep working.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "hashutil.hpp"
#include "helpers.hpp"

namespace cuckoo {

template <typename TagType>
struct CandidateBuckets {
    size_t i1;
    size_t i2;
    TagType fp;
};

template <typename TagType>
struct AltBucket {
    size_t bucket;
    TagType fp;
};

}  // namespace cuckoo

// Partial-key cuckoo hashing with XOR (bucket_policies.cuh:14-84).
// numBuckets is a power of two.
template <typename KeyType, typename TagType, size_t bitsPerTag, size_t bucketSize>
struct XorAltBucketPolicy {
    static constexpr size_t fpMask = (1ULL << bitsPerTag) - 1;
    static constexpr bool usesChoiceBit = false;
    static constexpr char name[] = "XorAltBucketPolicy";

    template <typename H>
    static __host__ __device__ __forceinline__ uint64_t hash64(const H& key) {
        return xxhash::xxhash64(key);
    }

    // fp = upper hash half masked to bitsPerTag (0 -> 1); i1 = lower half mod nb.
    static __host__ __device__ __forceinline__ cuckoo::CandidateBuckets<TagType>
    getCandidateBuckets(const KeyType& key, size_t numBuckets) {
        const uint64_t h = hash64(key);
        const uint32_t masked = static_cast<uint32_t>(h >> 32) & static_cast<uint32_t>(fpMask);
        const auto fp = static_cast<TagType>(masked == 0 ? 1 : masked);
        const size_t i1 = static_cast<uint32_t>(h) & (numBuckets - 1);
        return {i1, getAlternateBucket(i1, fp, numBuckets), fp};
    }

    static __host__ __device__ __forceinline__ size_t
    getAlternateBucket(size_t bucket, TagType fp, size_t numBuckets) {
        return bucket ^ (hash64(fp) & (numBuckets - 1));
    }

    static size_t calculateNumBuckets(size_t capacity) {
        const double required = std::ceil(static_cast<double>(capacity) / bucketSize);
        return nextPowerOfTwo(static_cast<size_t>(required));
    }
};

// Additive/subtractive cuckoo filter (bucket_policies.cuh:96-182): two halves
// of equal size; i1 in the lower half, i2 = i1 + H(fp) wrapped into the upper.
template <typename KeyType, typename TagType, size_t bitsPerTag, size_t bucketSize>
struct AddSubAltBucketPolicy {
    static constexpr size_t fpMask = (1ULL << bitsPerTag) - 1;
    static constexpr bool usesChoiceBit = false;
    static constexpr char name[] = "AddSubAltBucketPolicy";

    template <typename H>
    static __host__ __device__ __forceinline__ uint64_t hash64(const H& key) {
        return xxhash::xxhash64(key);
    }

    static __host__ __device__ __forceinline__ cuckoo::CandidateBuckets<TagType>
    getCandidateBuckets(const KeyType& key, size_t numBuckets) {
        const uint64_t h = hash64(key);
        const uint32_t masked = static_cast<uint32_t>(h >> 32) & static_cast<uint32_t>(fpMask);
        const auto fp = static_cast<TagType>(masked == 0 ? 1 : masked);
        const size_t half = numBuckets / 2;
        const size_t i1 = static_cast<uint32_t>(h) % half;
        return {i1, getAlternateBucket(i1, fp, numBuckets), fp};
    }

    static __host__ __device__ __forceinline__ size_t
    getAlternateBucket(size_t bucket, TagType fp, size_t numBuckets) {
        const size_t half = numBuckets / 2;
        const uint64_t fpHash = hash64(fp) % half;
        // both operands are below `half`, so one conditional subtraction is the modulo
        if (bucket < half) {
            size_t s = bucket + fpHash;
            return (s >= half ? s - half : s) + half;
        }
        size_t d = bucket - fpHash;  // in (0, 2*half)
        return d >= half ? d - half : d;
    }

    static size_t calculateNumBuckets(size_t capacity) {
        double required = std::ceil(static_cast<double>(capacity) / bucketSize);
        if (static_cast<size_t>(required) % 2 != 0) required += 1;
        return static_cast<size_t>(required);
    }
};

// Offset policy with a choice bit in the tag MSB (bucket_policies.cuh:198-350):
// b' = b + off(fp) mod nb from the primary, b = b' - off(fp) from the
// alternate; every move flips the choice bit.
template <typename KeyType, typename TagType, size_t bitsPerTag, size_t bucketSize>
struct OffsetAltBucketPolicy {
    static_assert(bitsPerTag >= 2, "bitsPerTag must be at least 2 for choice bit");

    static constexpr size_t choiceBitPos = bitsPerTag - 1;
    static constexpr TagType choiceBitMask = TagType(TagType(1) << choiceBitPos);
    static constexpr TagType pureFpMask = TagType((TagType(1) << choiceBitPos) - 1);
    static constexpr size_t fpMask = (1ULL << bitsPerTag) - 1;
    static constexpr bool usesChoiceBit = true;
    static constexpr char name[] = "OffsetAltBucketPolicy";

    template <typename H>
    static __host__ __device__ __forceinline__ uint64_t hash64(const H& key) {
        return xxhash::xxhash64(key);
    }

    static __host__ __device__ __forceinline__ TagType getPureFingerprint(TagType fp) {
        return fp & pureFpMask;
    }
    static __host__ __device__ __forceinline__ TagType getChoiceBit(TagType fp) {
        return (fp >> choiceBitPos) & 1;
    }
    static __host__ __device__ __forceinline__ TagType setChoiceBit(TagType fp, TagType choice) {
        return TagType((fp & pureFpMask) | TagType(choice << choiceBitPos));
    }

    // Non-zero offset in [1, nb-1] (bucket_policies.cuh:248-252).
    static __host__ __device__ __forceinline__ size_t computeOffset(TagType pureFp, size_t numBuckets) {
        const uint64_t offset = hash64(pureFp) % numBuckets;
        return offset == 0 ? 1 : offset;
    }

    static __host__ __device__ __forceinline__ cuckoo::CandidateBuckets<TagType>
    getCandidateBuckets(const KeyType& key, size_t numBuckets) {
        const uint64_t h = hash64(key);
        const uint32_t masked = static_cast<uint32_t>(h >> 32) & static_cast<uint32_t>(pureFpMask);
        const TagType pureFp = TagType(masked == 0 ? 1 : masked);
        const TagType fp = setChoiceBit(pureFp, 0);
        const size_t i1 = static_cast<uint32_t>(h) % numBuckets;
        size_t i2 = i1 + computeOffset(pureFp, numBuckets);
        if (i2 >= numBuckets) i2 -= numBuckets;
        return {i1, i2, fp};
    }

    // Flips the choice bit of fp in place (bucket_policies.cuh:302-321).
    static __host__ __device__ __forceinline__ size_t
    getAlternateBucket(size_t bucket, TagType& fp, size_t numBuckets) {
        const auto r = getAlternateBucketWithNewFp(bucket, fp, numBuckets);
        fp = r.fp;
        return r.bucket;
    }

    static __host__ __device__ __forceinline__ cuckoo::AltBucket<TagType>
    getAlternateBucketWithNewFp(size_t bucket, TagType fp, size_t numBuckets) {
        const TagType pureFp = getPureFingerprint(fp);
        const TagType choice = getChoiceBit(fp);
        const size_t offset = computeOffset(pureFp, numBuckets);
        size_t alt;
        if (choice == 0) {
            alt = bucket + offset;
            if (alt >= numBuckets) alt -= numBuckets;
        } else {
            alt = bucket + numBuckets - offset;
            if (alt >= numBuckets) alt -= numBuckets;
        }
        return {alt, setChoiceBit(pureFp, TagType(1 - choice))};
    }

    static size_t calculateNumBuckets(size_t capacity) {
        return static_cast<size_t>(std::ceil(static_cast<double>(capacity) / bucketSize));
    }
};
