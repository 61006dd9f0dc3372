// Device-side cuckoo-filter algorithm for gfx950 (wave64).
//
// This is the per-key method of the reference (include/CuckooFilter.cuh:
// tryRemoveAtBucket :805-848, tryInsertAtBucket :859-904, DFS eviction
// :917-954, BFS eviction :967-1018, insert :1029-1080, contains :1088-1098,
// remove :1106-1116), re-designed for CDNA:
//
//  * A bucket is read once as a register snapshot with 16-byte vector loads
//    (global_load_dwordx4), not word by word.  Slot choice then runs on the
//    snapshot: one SWAR mask per word, rotated so the scan still starts at the
//    word holding slot (tag & (bucketSize-1)) and wraps, and the lowest lane of
//    the first qualifying word wins -- the reference's order exactly.
//  * Every write is a relaxed device-scope compare-and-swap whose expected
//    value comes from the snapshot; a failed CAS returns the current word,
//    which replaces the snapshot word, and the scan continues.  A stale
//    snapshot therefore only costs a retry; it never writes a wrong value.
//    Reading the snapshot earlier than the reference's per-word loads is one
//    of the interleavings the reference's relaxed loads already allow.
//  * All operations take a trivially-copyable view by value, so kernels never
//    dereference a host object.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <utility>

#include "bucket_policies.hpp"
#include "hashutil.hpp"
#include "helpers.hpp"

/**
 * @brief Eviction policy for the cuckoo filter (CuckooFilter.cuh:19-22).
 */
enum class EvictionPolicy {
    BFS,  ///< Shallow breadth-first search with DFS fallback
    DFS   ///< Pure depth-first (greedy random-walk) search; the config default
};

namespace cuckoo {

/// Compile-time layout of a bucket for one configuration (CuckooFilter.cuh:226-249).
template <typename Config>
struct Layout {
    using Key = typename Config::KeyType;
    using Tag = typename Config::TagType;
    using Word = typename Config::WordType;
    static constexpr size_t bitsPerTag = Config::bitsPerTag;
    static constexpr size_t bucketSize = Config::bucketSize;
    static constexpr size_t tagsPerWord = sizeof(Word) / sizeof(Tag);
    static constexpr size_t wordCount = bucketSize / tagsPerWord;
    static constexpr size_t bucketBytes = wordCount * sizeof(Word);
    static constexpr Word fpMask = static_cast<Word>((1ULL << bitsPerTag) - 1);

    static_assert(bitsPerTag == 8 || bitsPerTag == 16 || bitsPerTag == 32,
                  "The tag must be 8, 16 or 32 bits");
    static_assert(powerOfTwo(bucketSize), "Bucket size must be a power of 2");
    static_assert(tagsPerWord >= 1, "TagType must fit within WordType");
    static_assert(bucketSize % tagsPerWord == 0, "bucketSize must be divisible by tagsPerWord");
    static_assert(powerOfTwo(wordCount), "wordCount must be a power of 2");
};

/// Everything a kernel needs about one filter, passed by value.
template <typename Config>
struct DeviceView {
    using Word = typename Config::WordType;
    Word* words;                       ///< numBuckets * wordCount words, bucket-major
    size_t numBuckets;                 ///< as CuckooFilter::numBuckets
    unsigned long long* occupied;      ///< occupancy counter (d_numOccupied)
    unsigned long long* evictions;     ///< eviction counter, nullptr = not counted
    unsigned int* chainHist;           ///< per-insert eviction-chain histogram, nullptr = off
    uint32_t maxEvictions;             ///< DFS chain bound
    uint32_t evictionPolicy;           ///< 0 = BFS, 1 = DFS

    DeviceView() = default;
    /// A filter's address converts to its view, so that launches written for
    /// the reference, which pass the filter object's address to the free
    /// kernels (CuckooFilter.cuh:459; benchmark/cuckoo_sorted_vs_unsorted.cu:74-76),
    /// compile unchanged and pass the view by value instead.
    template <typename F, typename = std::enable_if_t<std::is_same_v<decltype(std::declval<const F&>().view()),
                                                                    DeviceView>>>
    DeviceView(const F* filter) : DeviceView(filter->view()) {}
};

namespace detail {

template <typename Word>
__device__ __forceinline__ Word atomicLoadRelaxed(const Word* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/// Relaxed device-scope CAS; returns the word's value before the operation
/// (== expected exactly when the swap happened).
template <typename Word>
__device__ __forceinline__ Word casRelaxed(Word* p, Word expected, Word desired) {
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return expected;
}

/// Register copy of one bucket: a vector value, passed and returned by
/// value, so that the dynamically indexed word accesses below stay in
/// registers (v_cndmask selects).  A bucket kept in an addressable array (or a
/// struct passed by reference) gets its indexed reads rewritten into memory
/// loads with a computed offset, which moves the bucket to scratch or LDS.
template <typename Word, size_t WC>
using Snap = Word __attribute__((ext_vector_type(WC)));

/// Loads a whole bucket with the widest aligned vector loads it allows.
template <typename Word, size_t WC>
__device__ __forceinline__ Snap<Word, WC> loadBucket(const Word* p) {
    constexpr size_t bytes = WC * sizeof(Word);
    if constexpr (bytes % 16 == 0) {
        using V = unsigned int __attribute__((ext_vector_type(4)));
        using U = unsigned int __attribute__((ext_vector_type(bytes / 4)));
        U u;
#pragma unroll
        for (size_t i = 0; i < bytes / 16; ++i) {
            const V v = reinterpret_cast<const V*>(p)[i];
            u[4 * i + 0] = v.x;
            u[4 * i + 1] = v.y;
            u[4 * i + 2] = v.z;
            u[4 * i + 3] = v.w;
        }
        return __builtin_bit_cast(Snap<Word, WC>, u);
    } else if constexpr (bytes == 8) {
        using V = unsigned int __attribute__((ext_vector_type(2)));
        return __builtin_bit_cast(Snap<Word, WC>, *reinterpret_cast<const V*>(p));
    } else {
        static_assert(bytes == 4, "bucket of unexpected size");
        return __builtin_bit_cast(Snap<Word, WC>, *reinterpret_cast<const unsigned int*>(p));
    }
}

template <typename Word, size_t WC>
__device__ __forceinline__ Word pick(Snap<Word, WC> s, uint32_t idx) {
    return s[idx];
}

template <typename Word, size_t WC>
__device__ __forceinline__ Snap<Word, WC> put(Snap<Word, WC> s, uint32_t idx, Word v) {
    s[idx] = v;
    return s;
}

/// First word index at or after `start` (cyclically) whose bit is set in `mask`.
template <size_t WC>
__device__ __forceinline__ uint32_t firstFrom(uint64_t mask, uint32_t start) {
    if constexpr (WC == 1) {
        return 0;
    } else {
        constexpr uint64_t full = (WC >= 64) ? ~0ull : ((1ull << WC) - 1);
        const uint64_t rot = start == 0 ? mask : (((mask >> start) | (mask << (WC - start))) & full);
        return (start + lowestBit(rot)) & (WC - 1);
    }
}

}  // namespace detail

/// The per-key operations of one configuration.
template <typename Config>
struct FilterOps {
    using L = Layout<Config>;
    using Key = typename L::Key;
    using Tag = typename L::Tag;
    using Word = typename L::Word;
    using Policy = typename Config::AltBucketPolicy;
    using View = DeviceView<Config>;
    static constexpr size_t WC = L::wordCount;
    static constexpr size_t TPW = L::tagsPerWord;
    static constexpr size_t BITS = L::bitsPerTag;
    static constexpr size_t BS = L::bucketSize;
    static constexpr Tag EMPTY = 0;
    /// Buckets up to 128 B are handled as register snapshots; larger ones word by word.
    static constexpr bool kSnapshot = L::bucketBytes <= 128;
    using SnapT = detail::Snap<Word, WC>;
    /// Lanes that read one bucket together in the lookup kernel: one per 32
    /// bytes of 64- and 128-byte buckets.  Over tables of 4 GiB and more the
    /// random reads are bound by address translation, once per load
    /// instruction and lane, so a bucket read by one lane as four or eight
    /// 16-byte loads costs two to three times what it costs spread over lanes.
    static constexpr uint32_t kLookupLanes =
        (kSnapshot && L::bucketBytes >= 64) ? static_cast<uint32_t>(L::bucketBytes / 32) : 1u;
    /// Whether a lookup kernel of `block` threads reads buckets with
    /// kLookupLanes lanes per key: always for 128-byte buckets (equal or
    /// faster at every table size measured), for 64-byte buckets only past
    /// 2 GiB of table (below that one lane per key is 1-6% faster).
    static __host__ __device__ __forceinline__ bool groupedLookup(size_t numBuckets, uint32_t block) {
        if constexpr (kLookupLanes == 1) {
            return false;
        } else {
            return block % kLookupLanes == 0 &&
                   (L::bucketBytes >= 128 || numBuckets * L::bucketBytes > (static_cast<size_t>(2) << 30));
        }
    }

    static __host__ __device__ __forceinline__ Tag extractTag(Word packed, size_t lane) {
        return static_cast<Tag>((packed >> (lane * BITS)) & L::fpMask);
    }

    static __host__ __device__ __forceinline__ Word replaceTag(Word packed, size_t lane, Tag t) {
        const size_t shift = lane * BITS;
        return (packed & ~(static_cast<Word>(L::fpMask) << shift)) | (static_cast<Word>(t) << shift);
    }

    /// getAlternateBucketWithNewFp (CuckooFilter.cuh:388-400).
    static __host__ __device__ __forceinline__ cuckoo::AltBucket<Tag>
    altWithNewFp(size_t bucket, Tag fp, size_t nb) {
        if constexpr (Policy::usesChoiceBit) {
            return Policy::getAlternateBucketWithNewFp(bucket, fp, nb);
        } else {
            return {Policy::getAlternateBucket(bucket, fp, nb), fp};
        }
    }

    static __device__ __forceinline__ Word* bucketPtr(const View& v, size_t b) {
        return v.words + b * WC;
    }

    static __device__ __forceinline__ uint32_t startWord(Tag tag) {
        return static_cast<uint32_t>((tag & (BS - 1)) / TPW);
    }

    // ---------------------------------------------------------------------
    // snapshot primitives
    // ---------------------------------------------------------------------

    static __device__ __forceinline__ uint64_t emptyWords(SnapT s) {
        uint64_t m = 0;
#pragma unroll
        for (uint32_t j = 0; j < WC; ++j) m |= uint64_t(getZeroMask<Tag, Word>(s[j]) != 0) << j;
        return m;
    }

    static __device__ __forceinline__ uint64_t matchWords(SnapT s, Tag tag) {
        const Word rep = replicateTag<Tag, Word>(tag);
        uint64_t m = 0;
#pragma unroll
        for (uint32_t j = 0; j < WC; ++j) m |= uint64_t(getZeroMask<Tag, Word>(s[j] ^ rep) != 0) << j;
        return m;
    }

    static __device__ __forceinline__ bool snapHas(SnapT s, Tag tag) {
        const Word rep = replicateTag<Tag, Word>(tag);
        Word acc = 0;
#pragma unroll
        for (uint32_t j = 0; j < WC; ++j) acc |= getZeroMask<Tag, Word>(s[j] ^ rep);
        return acc != 0;
    }

    /// tryInsertAtBucket on a snapshot the caller already holds (CuckooFilter.cuh:859-904).
    static __device__ __forceinline__ bool tryInsertSnap(Word* bw, SnapT s, Tag tag) {
        const uint32_t start = startWord(tag);
        for (;;) {
            const uint64_t m = emptyWords(s);
            if (m == 0) return false;
            const uint32_t w = detail::firstFrom<WC>(m, start);
            const Word cur = detail::pick(s, w);
            const uint32_t lane = lowestBit(getZeroMask<Tag, Word>(cur)) / BITS;
            const Word old = detail::casRelaxed(bw + w, cur, replaceTag(cur, lane, tag));
            if (old == cur) return true;
            s = detail::put(s, w, old);
        }
    }

    /// The first step of tryInsertAtBucket on a bucket in LDS: one attempt at
    /// the tag's start word, where the reference's scan starts
    /// (CuckooFilter.cuh:866-872).  False if that word is full or the CAS
    /// lost; the caller then continues with the whole scan (tryInsertWordsLds).
    static __device__ __forceinline__ bool tryStartWordLds(Word* bw, Tag tag) {
        const uint32_t start = startWord(tag);
        const Word cur = bw[start];
        const Word zm = getZeroMask<Tag, Word>(cur);
        if (zm == 0) return false;
        const uint32_t lane = lowestBit(zm) / BITS;
        return detail::casRelaxed(bw + start, cur, replaceTag(cur, lane, tag)) == cur;
    }

    // Buckets of an even number of 32-bit words are scanned in LDS two words
    // at a time.  Two adjacent little-endian 32-bit words hold their slots
    // exactly where one 64-bit word would, so one 64-bit access covers two of
    // the reference's words.  The lane masks are still taken per 32-bit word
    // (the reference's SWAR, whose lowest flag per word is exact), and the
    // reference's word order from the start word is kept: when the start
    // word is the upper half of its pair, the first visit allows that half
    // only and the scan ends with the pair's lower half.  A 64-bit CAS that
    // fails because the other half changed rereads and continues, as a lost
    // CAS does.
    static constexpr bool kPairScan = sizeof(Word) == 4 && WC % 2 == 0;

    /// Per-lane flags of a pair of 32-bit words: empty lanes, or (Match)
    /// lanes equal to the tag whose broadcast is `rep`.
    template <bool Match>
    static __device__ __forceinline__ uint64_t pairFlags(uint64_t pair, Word rep) {
        const Word lo = static_cast<Word>(pair);
        const Word hi = static_cast<Word>(pair >> 32);
        const Word fl = getZeroMask<Tag, Word>(Match ? (lo ^ rep) : lo);
        const Word fh = getZeroMask<Tag, Word>(Match ? (hi ^ rep) : hi);
        return static_cast<uint64_t>(fl) | (static_cast<uint64_t>(fh) << 32);
    }

    /// tryInsertWordsLds / tryRemoveWordsLds (Remove) over word pairs.
    template <bool Remove>
    static __device__ __forceinline__ bool scanPairsLds(Word* bw32, Tag tag) {
        constexpr uint32_t PC = static_cast<uint32_t>(WC / 2);
        auto* bw = reinterpret_cast<uint64_t*>(bw32);  // a bucket of an even word count is 8-byte aligned
        const Word rep = replicateTag<Tag, Word>(tag);
        const uint32_t start = startWord(tag);
        const uint32_t first = start / 2;
        const uint64_t firstAllowed = (start & 1u) ? (~uint64_t(0) << 32) : ~uint64_t(0);
        const uint32_t visits = PC + (start & 1u);
        for (uint32_t i = 0; i < visits; ++i) {
            const uint32_t w = (first + i) & (PC - 1);
            const uint64_t allowed = i == 0 ? firstAllowed : (i == PC ? ~firstAllowed : ~uint64_t(0));
            uint64_t cur = bw[w];
            for (;;) {
                const uint64_t fl = pairFlags<Remove>(cur, rep) & allowed;
                if (fl == 0) break;
                const uint32_t shift = (lowestBit(fl) / static_cast<uint32_t>(BITS)) * static_cast<uint32_t>(BITS);
                const uint64_t put = Remove ? uint64_t(EMPTY) : static_cast<uint64_t>(tag);
                const uint64_t desired = (cur & ~(static_cast<uint64_t>(L::fpMask) << shift)) | (put << shift);
                const uint64_t old = detail::casRelaxed(bw + w, cur, desired);
                if (old == cur) return true;
                cur = old;
            }
        }
        return false;
    }

    /// tryInsertAtBucket on a bucket in LDS as the reference scans it
    /// (CuckooFilter.cuh:859-904): word by word from the tag's start word,
    /// reading a word only when the ones before it are full.  In LDS the
    /// reads cost bank cycles, not round trips, so the early exit beats one
    /// whole-bucket read.
    static __device__ __forceinline__ bool tryInsertWordsLds(Word* bw, Tag tag) {
        if constexpr (kPairScan) return scanPairsLds<false>(bw, tag);
        const uint32_t start = startWord(tag);
#pragma unroll
        for (uint32_t i = 0; i < WC; ++i) {
            const uint32_t w = (start + i) & (WC - 1);
            Word cur = bw[w];
            for (;;) {
                const Word zm = getZeroMask<Tag, Word>(cur);
                if (zm == 0) break;
                const uint32_t lane = lowestBit(zm) / BITS;
                const Word old = detail::casRelaxed(bw + w, cur, replaceTag(cur, lane, tag));
                if (old == cur) return true;
                cur = old;
            }
        }
        return false;
    }

    /// tryRemoveAtBucket on a bucket in LDS, word by word from the start word
    /// (CuckooFilter.cuh:805-848); a tag is most often found in its start word.
    static __device__ __forceinline__ bool tryRemoveWordsLds(Word* bw, Tag tag) {
        if constexpr (kPairScan) return scanPairsLds<true>(bw, tag);
        const uint32_t start = startWord(tag);
        const Word rep = replicateTag<Tag, Word>(tag);
#pragma unroll
        for (uint32_t i = 0; i < WC; ++i) {
            const uint32_t w = (start + i) & (WC - 1);
            Word cur = bw[w];
            for (;;) {
                const Word mm = getZeroMask<Tag, Word>(cur ^ rep);
                if (mm == 0) break;
                const uint32_t lane = lowestBit(mm) / BITS;
                const Word old = detail::casRelaxed(bw + w, cur, replaceTag(cur, lane, EMPTY));
                if (old == cur) return true;
                cur = old;
            }
        }
        return false;
    }

    /// tryRemoveAtBucket on a snapshot (CuckooFilter.cuh:805-848).
    static __device__ __forceinline__ bool tryRemoveSnap(Word* bw, SnapT s, Tag tag) {
        const uint32_t start = startWord(tag);
        const Word rep = replicateTag<Tag, Word>(tag);
        for (;;) {
            const uint64_t m = matchWords(s, tag);
            if (m == 0) return false;
            const uint32_t w = detail::firstFrom<WC>(m, start);
            const Word cur = detail::pick(s, w);
            const uint32_t lane = lowestBit(getZeroMask<Tag, Word>(cur ^ rep)) / BITS;
            const Word old = detail::casRelaxed(bw + w, cur, replaceTag(cur, lane, EMPTY));
            if (old == cur) return true;
            s = detail::put(s, w, old);
        }
    }

    // ---------------------------------------------------------------------
    // bucket-level operations (snapshot for <=128 B buckets, word loop above)
    // ---------------------------------------------------------------------

    static __device__ __forceinline__ bool tryInsertAtBucket(const View& v, size_t b, Tag tag) {
        Word* bw = bucketPtr(v, b);
        if constexpr (kSnapshot) {
            return tryInsertSnap(bw, detail::loadBucket<Word, WC>(bw), tag);
        } else {
            const uint32_t start = startWord(tag);
            for (uint32_t i = 0; i < WC; ++i) {
                const uint32_t w = (start + i) & (WC - 1);
                Word cur = detail::atomicLoadRelaxed(bw + w);
                for (;;) {
                    const Word zm = getZeroMask<Tag, Word>(cur);
                    if (zm == 0) break;
                    const uint32_t lane = lowestBit(zm) / BITS;
                    const Word old = detail::casRelaxed(bw + w, cur, replaceTag(cur, lane, tag));
                    if (old == cur) return true;
                    cur = old;
                }
            }
            return false;
        }
    }

    static __device__ __forceinline__ bool tryRemoveAtBucket(const View& v, size_t b, Tag tag) {
        Word* bw = bucketPtr(v, b);
        if constexpr (kSnapshot) {
            return tryRemoveSnap(bw, detail::loadBucket<Word, WC>(bw), tag);
        } else {
            const uint32_t start = startWord(tag);
            const Word rep = replicateTag<Tag, Word>(tag);
            for (uint32_t i = 0; i < WC; ++i) {
                const uint32_t w = (start + i) & (WC - 1);
                Word cur = detail::atomicLoadRelaxed(bw + w);
                for (;;) {
                    const Word mm = getZeroMask<Tag, Word>(cur ^ rep);
                    if (mm == 0) break;
                    const uint32_t lane = lowestBit(mm) / BITS;
                    const Word old = detail::casRelaxed(bw + w, cur, replaceTag(cur, lane, EMPTY));
                    if (old == cur) return true;
                    cur = old;
                }
            }
            return false;
        }
    }

    /// Bucket::contains (CuckooFilter.cuh:290-346): true iff some slot == tag.
    static __device__ __forceinline__ bool bucketContains(const View& v, size_t b, Tag tag) {
        const Word* bw = bucketPtr(v, b);
        if constexpr (kSnapshot) {
            return snapHas(detail::loadBucket<Word, WC>(bw), tag);
        } else {
            // 128-byte chunks starting at the chunk holding the start slot, with early exit.
            constexpr uint32_t CW = 128 / sizeof(Word);
            constexpr uint32_t NCH = WC / CW;
            const uint32_t c0 = startWord(tag) / CW;
            for (uint32_t i = 0; i < NCH; ++i) {
                const uint32_t c = (c0 + i) & (NCH - 1);
                if (FilterOps::chunkHas<CW>(detail::loadBucket<Word, CW>(bw + c * CW), tag)) {
                    return true;
                }
            }
            return false;
        }
    }

    /// Whether the 32-byte part `part` of bucket b holds the tag (one lane's
    /// share of a bucket read by kLookupLanes lanes).
    static __device__ __forceinline__ bool bucketPartHas(const View& v, size_t b, uint32_t part, Tag tag) {
        constexpr size_t PW = 32 / sizeof(Word);
        return chunkHas<PW>(detail::loadBucket<Word, PW>(bucketPtr(v, b) + part * PW), tag);
    }

    template <size_t N>
    static __device__ __forceinline__ bool chunkHas(detail::Snap<Word, N> s, Tag tag) {
        const Word rep = replicateTag<Tag, Word>(tag);
        Word acc = 0;
#pragma unroll
        for (uint32_t j = 0; j < N; ++j) acc |= getZeroMask<Tag, Word>(s[j] ^ rep);
        return acc != 0;
    }

    // ---------------------------------------------------------------------
    // eviction (CuckooFilter.cuh:917-1018)
    // ---------------------------------------------------------------------

    /// How a run of DFS swaps ended.
    enum DfsState : uint32_t { kDfsPlaced = 0, kDfsExhausted = 1, kDfsPaused = 2 };

    /// Swaps e, e+1, ... of insertWithEvictionDFS carrying curFp into curB:
    /// ends when the carried tag lands (kDfsPlaced), when maxEvictions swaps
    /// have been made (kDfsExhausted), or before swap number `limit`
    /// (kDfsPaused).  curFp, curB and e then hold the state to resume from; e
    /// counts the swaps made.
    static __device__ uint32_t dfsSteps(const View& v, Tag& curFp, size_t& curB, uint32_t& e, uint32_t limit) {
        for (; e < v.maxEvictions; ++e) {
            if (e == limit) return kDfsPaused;
            const size_t slot = (static_cast<size_t>(curFp) + static_cast<size_t>(e) * 7) & (BS - 1);
            const uint32_t w = static_cast<uint32_t>(slot / TPW);
            const uint32_t lane = static_cast<uint32_t>(slot & (TPW - 1));
            Word* p = bucketPtr(v, curB) + w;
            Word cur = detail::atomicLoadRelaxed(p);  // the CAS loop settles the real value
            Tag evicted;
            for (;;) {
                evicted = extractTag(cur, lane);
                const Word old = detail::casRelaxed(p, cur, replaceTag(cur, lane, curFp));
                if (old == cur) break;
                cur = old;
            }
            const auto alt = altWithNewFp(curB, evicted, v.numBuckets);
            curB = alt.bucket;
            curFp = alt.fp;
            if (tryInsertAtBucket(v, curB, curFp)) {
                ++e;
                return kDfsPlaced;
            }
        }
        return kDfsExhausted;
    }

    /// insertWithEvictionDFS (CuckooFilter.cuh:917-958).  `chain` receives
    /// the number of swaps made.
    static __device__ bool insertDFS(const View& v, Tag fp, size_t startBucket, uint32_t& chain) {
        Tag curFp = fp;
        size_t curB = startBucket;
        uint32_t e = 0;
        const bool placed = dfsSteps(v, curFp, curB, e, ~0u) == kDfsPlaced;
        if (e != 0) chain = e;
        return placed;
    }

    /// The BFS stage of insertWithEvictionBFS (CuckooFilter.cuh:968-1014): up to
    /// min(8, bucketSize/2) candidates at slots (fp + i*7) & (bucketSize-1);
    /// false means the caller continues with DFS from the same bucket (:1017).
    static constexpr uint32_t kBfsCandidates = (BS / 2) < 8 ? uint32_t(BS / 2) : 8u;

    static __device__ bool bfsCandidates(const View& v, Tag fp, size_t startBucket, uint32_t& chain) {
        for (uint32_t i = 0; i < kBfsCandidates; ++i) {
            if (bfsCandidate(v, fp, startBucket, i, chain)) return true;
        }
        return false;
    }

    /// Candidate i of the BFS stage: true if the tag was placed (chain = 1
    /// when a candidate moved to its alternate bucket to make room).
    static __device__ bool bfsCandidate(const View& v, Tag fp, size_t startBucket, uint32_t i, uint32_t& chain) {
        Word* sb = bucketPtr(v, startBucket);
        const size_t slot = (static_cast<size_t>(fp) + static_cast<size_t>(i) * 7) & (BS - 1);
        const uint32_t w = static_cast<uint32_t>(slot / TPW);
        const uint32_t lane = static_cast<uint32_t>(slot & (TPW - 1));
        Word cur = detail::atomicLoadRelaxed(sb + w);
        const Tag cand = extractTag(cur, lane);
        if (cand == EMPTY) return tryInsertAtBucket(v, startBucket, fp);
        const auto alt = altWithNewFp(startBucket, cand, v.numBuckets);
        if (tryInsertAtBucket(v, alt.bucket, alt.fp)) {
            // swap our tag in only while the candidate still sits there
            while (extractTag(cur, lane) == cand) {
                const Word old = detail::casRelaxed(sb + w, cur, replaceTag(cur, lane, fp));
                if (old == cur) {
                    chain = 1;
                    return true;
                }
                cur = old;
            }
            // :1011-1012 -- undo the copy placed in the alternate bucket
            tryRemoveAtBucket(v, alt.bucket, cand);
        }
        return false;
    }

    /// insertWithEvictionBFS: the BFS stage, then DFS (CuckooFilter.cuh:967-1018).
    static __device__ bool insertBFS(const View& v, Tag fp, size_t startBucket, uint32_t& chain) {
        if (bfsCandidates(v, fp, startBucket, chain)) return true;
        return insertDFS(v, fp, startBucket, chain);
    }

    /// CuckooFilter::insert from its primary bucket (CuckooFilter.cuh:1029-1080,
    /// insertKernelSorted :1228-1258).
    static __device__ __forceinline__ bool insertFrom(const View& v, size_t i1, Tag fp, uint32_t& chain) {
        chain = 0;
        if (tryInsertAtBucket(v, i1, fp)) return true;
        return insertAfterPrimary(v, i1, fp, chain);
    }

    /// The rest of an insert whose attempt at the primary bucket i1 has
    /// already failed: the secondary bucket, then eviction from the bucket
    /// the fingerprint's low bit selects (CuckooFilter.cuh:1062-1076).
    static __device__ __forceinline__ bool insertAfterPrimary(const View& v, size_t i1, Tag fp, uint32_t& chain) {
        chain = 0;
        const auto second = altWithNewFp(i1, fp, v.numBuckets);
        if (tryInsertAtBucket(v, second.bucket, second.fp)) return true;
        const EvictStart e = evictStart(i1, fp, second);
        if (v.evictionPolicy == static_cast<uint32_t>(EvictionPolicy::BFS)) {
            return insertBFS(v, e.fp, e.bucket, chain);
        }
        return insertDFS(v, e.fp, e.bucket, chain);
    }

    /// Where eviction starts after both candidate buckets were full: the
    /// bucket selected by the fingerprint's low bit, with the fingerprint as
    /// stored there (CuckooFilter.cuh:1062-1066, 1243-1244).
    struct EvictStart {
        size_t bucket;
        Tag fp;
    };
    static __device__ __forceinline__ EvictStart evictStart(size_t i1, Tag fp, cuckoo::AltBucket<Tag> second) {
        const bool first = (fp & 1) == 0;
        return {first ? i1 : second.bucket, first ? fp : second.fp};
    }

    static __device__ __forceinline__ bool insert(const View& v, const Key& key, uint32_t& chain) {
        const auto c = Policy::getCandidateBuckets(key, v.numBuckets);
        return insertFrom(v, c.i1, c.fp, chain);
    }

    /// CuckooFilter::contains (CuckooFilter.cuh:1088-1098).
    static __device__ __forceinline__ bool contains(const View& v, const Key& key) {
        const auto c = Policy::getCandidateBuckets(key, v.numBuckets);
        if (bucketContains(v, c.i1, c.fp)) return true;
        Tag fp2 = c.fp;
        if constexpr (Policy::usesChoiceBit) fp2 = altWithNewFp(c.i1, c.fp, v.numBuckets).fp;
        if (c.i2 == c.i1 && fp2 == c.fp) return false;
        return bucketContains(v, c.i2, fp2);
    }

    /// CuckooFilter::remove (CuckooFilter.cuh:1106-1116).
    static __device__ __forceinline__ bool remove(const View& v, const Key& key) {
        const auto c = Policy::getCandidateBuckets(key, v.numBuckets);
        if (tryRemoveAtBucket(v, c.i1, c.fp)) return true;
        Tag fp2 = c.fp;
        if constexpr (Policy::usesChoiceBit) fp2 = altWithNewFp(c.i1, c.fp, v.numBuckets).fp;
        return tryRemoveAtBucket(v, c.i2, fp2);
    }
};

}  // namespace cuckoo
