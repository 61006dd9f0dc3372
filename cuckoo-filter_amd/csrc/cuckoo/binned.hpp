// Bucket-binned insert: the MI355X path for large insert batches.
//
// A device-wide CAS on MI355X is executed at the memory side and moves a
// 64-byte line; random CAS tops out near 22 G/s, and the reference's
// one-thread-per-key insert (CuckooFilter.cuh:1119-1141) issues at least one
// per key.  This path removes the CAS from the common case:
//
//  1. entries (primary bucket << 32 | fingerprint) are computed from the keys
//     inside the first pass of an LSD radix partition that groups them by
//     table slice (a slice is the run of buckets that fills kSliceBytes);
//  2. one workgroup per slice copies the slice into LDS, performs every
//     primary-bucket attempt of its entries there (the same
//     tryInsertAtBucket slot choice, with LDS compare-and-swap between the
//     workgroup's own lanes) and writes the slice back with plain vector
//     stores; entries whose primary bucket is full are appended to an
//     overflow list;
//  3. the overflow entries continue exactly where the reference's insert
//     continues after a full primary bucket -- the secondary bucket, then
//     BFS/DFS eviction -- with device-scope CAS.
//
// Why this is the reference's method: its insert is concurrent, so every
// interleaving of the per-key steps is a possible outcome.  Running all
// primary-bucket attempts first, then all the rest, is one such
// interleaving, and during step 2 no other workgroup touches a slice while it
// sits in LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "device_ops.hpp"
#include "kernels.hpp"
#include "radix.hpp"

namespace cuckoo {

/// Table bytes one workgroup stages in LDS.
inline constexpr size_t kSliceBytes = 32 * 1024;
inline constexpr int kBinBlock = 1024;

/// Packed (primary bucket, fingerprint) entry of key i.  With SortedSemantics
/// the primary bucket is the one insertManySorted derives from the reference's
/// PackedTagType (CuckooFilter.cuh:159,1205-1206), which keeps only the bucket
/// bits that fit beside the tag.
template <typename Config, bool SortedSemantics>
struct EntrySource {
    using Key = typename Config::KeyType;
    using Raw = Key;
    const Key* keys;
    size_t numBuckets;

    __device__ __forceinline__ Raw load(size_t i) const { return keys[i]; }
    __device__ __forceinline__ Raw loadRel(size_t tileBase, uint32_t off) const { return (keys + tileBase)[off]; }
    __device__ __forceinline__ uint64_t operator()(size_t i) const { return cook(keys[i]); }
    __device__ __forceinline__ uint64_t cook(Key key) const {
        const auto c = Config::AltBucketPolicy::getCandidateBuckets(key, numBuckets);
        uint64_t i1 = c.i1;
        if constexpr (SortedSemantics) {
            using P = std::conditional_t<Config::bitsPerTag <= 8, uint32_t, uint64_t>;
            i1 = static_cast<uint64_t>(static_cast<P>(static_cast<P>(c.i1) << Config::bitsPerTag) >>
                                       Config::bitsPerTag);
        }
        return (i1 << 32) | static_cast<uint64_t>(c.fp);
    }
};

/// First pass of the two-level partition: the entry's bucket bits below the
/// first digit, beside the fingerprint, in 32 bits.
struct PackBelowDigit {
    using Out = uint32_t;
    uint64_t bucketMask;
    int tagBits;
    __device__ __forceinline__ uint32_t operator()(uint64_t e) const {
        const uint64_t fp = e & ((uint64_t(1) << tagBits) - 1);
        return static_cast<uint32_t>((((e >> 32) & bucketMask) << tagBits) | fp);
    }
};

/// Second pass: the bucket's position inside its slice, beside the fingerprint.
struct KeepLowBits {
    using Out = uint32_t;
    uint32_t mask;
    __device__ __forceinline__ uint32_t operator()(uint32_t v) const { return v & mask; }
};

/// Zeroes the three list counters of the insert (overflow, eviction starts,
/// DFS list).
template <typename C = unsigned long long>
__global__ __launch_bounds__(64) void zeroCountersKernel(C* counters) {
    if (threadIdx.x < 3) counters[threadIdx.x] = 0;
}

template <typename Src>
__global__ __launch_bounds__(1024) void packEntriesKernel(Src src, size_t n, uint64_t* __restrict__ out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = src(i);
}

/// start[s] = first entry whose slice is >= s (entries sorted by slice), for
/// s in [0, numSlices]; also zeroes the three list counters of the insert.
template <typename T = uint64_t>
__global__ __launch_bounds__(256) void sliceStartKernel(const T* __restrict__ entries, uint32_t n, int sliceShift,
                                                        uint32_t numSlices, uint32_t* __restrict__ start,
                                                        unsigned long long* overflowCount) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < 3) overflowCount[s] = 0;  // overflow, eviction-start and DFS list counters
    if (s > numSlices) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if ((entries[mid] >> sliceShift) < s) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    start[s] = lo;
}

/// Entries each thread of binInsertKernel loads at once, and the entries a
/// workgroup takes per round.
inline constexpr uint32_t kBinDepth = 4;
inline constexpr uint32_t kBinRound = kBinBlock * kBinDepth;

/// Dynamic LDS one slice workgroup may take so that two stay resident on a
/// CU (160 KiB): half of it, less room for the kernels' static counters.
inline constexpr size_t kBinLdsBudget = 80 * 1024 - 256;

/// Overflow entries a workgroup buffers in LDS before reserving their global
/// space with a single atomic: what the budget leaves beside the slice (and,
/// in the insert, beside a round's retry list of `retryBytes` entries).
inline constexpr uint32_t binOverflowCapacity(size_t retryBytes) {
    return static_cast<uint32_t>((kBinLdsBudget - kSliceBytes - kBinRound * retryBytes) / sizeof(uint64_t));
}

/// Dynamic LDS of binInsertKernel: the table slice, the overflow buffer, then
/// the round's retry list (kBinRound entries of retryBytes each).
inline constexpr size_t binOverflowOffset(size_t sliceBytes) { return (sliceBytes + 15) / 16 * 16; }
inline constexpr size_t binRetryOffset(size_t sliceBytes, size_t retryBytes) {
    return binOverflowOffset(sliceBytes) + binOverflowCapacity(retryBytes) * sizeof(uint64_t);
}
inline constexpr size_t binLdsBytes(size_t sliceBytes, size_t retryBytes) {
    return binRetryOffset(sliceBytes, retryBytes) + kBinRound * retryBytes;
}
/// Dynamic LDS of binDeleteKernel: the table slice, then its overflow buffer.
inline constexpr size_t binDeleteLdsBytes(size_t sliceBytes) {
    return binOverflowOffset(sliceBytes) + binOverflowCapacity(0) * sizeof(uint64_t);
}
static_assert(binLdsBytes(kSliceBytes, 8) <= kBinLdsBudget && binDeleteLdsBytes(kSliceBytes) <= kBinLdsBudget);
static_assert(binOverflowCapacity(8) >= 2000);

/// Appends the lanes' entries with `emit` to the workgroup's LDS overflow
/// buffer of `capacity` entries (one LDS atomic per wave).  Entries past it
/// go straight to the global list, with one global atomic per wave for the
/// wave's whole run of them.
__device__ __forceinline__ void appendOverflow(bool emit, uint64_t entry, uint64_t* ovBuf, uint32_t* ovCount,
                                               uint32_t capacity, uint64_t* overflow,
                                               unsigned long long* overflowCount) {
    const uint64_t mask = __ballot(emit);
    if (mask == 0) return;
    const int lane = __lane_id();
    const uint64_t ltMask = (1ull << lane) - 1;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(ovCount, static_cast<uint32_t>(__popcll(mask)));
    base = __shfl(base, leader, 64);
    const uint32_t mine = base + static_cast<uint32_t>(__popcll(mask & ltMask));
    const bool spill = emit && mine >= capacity;
    if (emit && !spill) ovBuf[mine] = entry;
    const uint64_t spills = __ballot(spill);
    if (spills == 0) return;
    const int sl = __ffsll(static_cast<long long>(spills)) - 1;
    unsigned long long gbase = 0;
    if (lane == sl) gbase = atomicAdd(overflowCount, static_cast<unsigned long long>(__popcll(spills)));
    gbase = __shfl(gbase, sl, 64);
    if (spill) overflow[gbase + __popcll(spills & ltMask)] = entry;
}

/// Between two rounds of a slice kernel (every thread, after a barrier):
/// once the LDS overflow buffer is half full, moves its entries to the global
/// list with one global atomic.  Slices of small buckets near full load
/// overflow by thousands of entries; reserving their space lane by lane
/// serialised the whole grid on the list counter's address, and flushing
/// after every round costs three barriers a round where few entries fail.
__device__ __forceinline__ void flushOverflowBetweenRounds(uint64_t* ovBuf, uint32_t* ovCount, uint32_t capacity,
                                                           uint64_t* overflow, unsigned long long* overflowCount,
                                                           unsigned long long* ovBase) {
    const uint32_t total = *ovCount;  // buffered entries plus those appendOverflow spilled past the buffer
    if (total < capacity / 2) return;  // uniform: every thread reads the same count
    const uint32_t count = total < capacity ? total : capacity;  // the spilled ones are on the list already
    if (threadIdx.x == 0) *ovBase = atomicAdd(overflowCount, static_cast<unsigned long long>(count));
    __syncthreads();
    const unsigned long long base = *ovBase;
    for (uint32_t i = threadIdx.x; i < count; i += kBinBlock) overflow[base + i] = ovBuf[i];
    __syncthreads();
    if (threadIdx.x == 0) *ovCount = 0;
    __syncthreads();
}

/// Retry-list entry of binInsertKernel: (bucket inside the slice << tag bits
/// | fingerprint), in 32 bits where that fits.
template <typename Config>
using BinRetryOf = std::conditional_t<(sizeof(typename Config::TagType) <= 2), uint32_t, uint64_t>;

/// Step 2: primary-bucket attempts of one slice's entries inside LDS.  An
/// entry E is either (bucket << 32 | fingerprint) (uint64_t) or, from the
/// two-level partition, (bucket inside the slice << tag bits | fingerprint)
/// (uint32_t).
///
/// Each round of kBinRound entries runs in two phases, so that the lanes of
/// a wave do the same work:
///  1. every entry tries its tag's start word (one LDS word read, one CAS),
///     which is where the reference's slot scan starts; entries whose start
///     word is full, or whose CAS lost, go to an LDS retry list;
///  2. the retry list, compacted into full waves, runs the rest of the
///     reference's scan on the whole bucket from the start word on; entries
///     that find the bucket full go to the overflow list.
/// Deferring a key between its start word and the rest of its bucket is an
/// interleaving the concurrent reference allows; a full word stays full while
/// only inserts run, so the later scan sees what the reference's thread would
/// see after a delay.
template <typename Config, typename E = uint64_t>
__global__ __launch_bounds__(kBinBlock) void binInsertKernel(const E* __restrict__ entries,
                                                             const uint32_t* __restrict__ start, int log2Slice,
                                                             DeviceView<Config> v, uint64_t* __restrict__ overflow,
                                                             unsigned long long* overflowCount) {
    using Ops = FilterOps<Config>;
    using Word = typename Config::WordType;
    using Tag = typename Config::TagType;
    using R = BinRetryOf<Config>;
    constexpr size_t WC = Layout<Config>::wordCount;
    constexpr size_t kBucketBytes = Layout<Config>::bucketBytes;
    constexpr int kTagBits = 8 * sizeof(Tag);
    extern __shared__ __align__(16) unsigned char ldsRaw[];
    __shared__ uint32_t ovCount;
    __shared__ uint32_t retryCount[2];  // by round parity
    __shared__ unsigned long long ovBase;

    const uint32_t s = blockIdx.x;
    const uint32_t lo = start[s];
    const uint32_t hi = start[s + 1];
    if (lo == hi) return;  // no entry for this slice: the table is left untouched

    const size_t sliceBuckets = static_cast<size_t>(1) << log2Slice;
    Word* slice = reinterpret_cast<Word*>(ldsRaw);
    uint64_t* ovBuf = reinterpret_cast<uint64_t*>(ldsRaw + binOverflowOffset(sliceBuckets * kBucketBytes));
    R* retry = reinterpret_cast<R*>(ldsRaw + binRetryOffset(sliceBuckets * kBucketBytes, sizeof(R)));
    constexpr uint32_t kOverflowCap = binOverflowCapacity(sizeof(R));
    const size_t b0 = static_cast<size_t>(s) << log2Slice;
    const size_t nbs = v.numBuckets - b0 < sliceBuckets ? v.numBuckets - b0 : sliceBuckets;
    const size_t words = nbs * WC;
    const Word* g = v.words + b0 * WC;
    Word* gw = v.words + b0 * WC;

    if (threadIdx.x == 0) {
        ovCount = 0;
        retryCount[0] = 0;
        retryCount[1] = 0;
    }
    // A round's entries are loaded one round ahead (kBinDepth per thread, each
    // load coalesced across the wave; past-the-end lanes reread the last), the
    // first round's together with the slice.
    const uint32_t last = hi - 1;
    E next[kBinDepth];
    auto loadRound = [&](uint32_t r0) {
#pragma unroll
        for (uint32_t k = 0; k < kBinDepth; ++k) {
            const uint32_t e = r0 + k * kBinBlock + threadIdx.x;
            next[k] = entries[e < last ? e : last];
        }
    };
    loadRound(lo);
    using V4 = unsigned int __attribute__((ext_vector_type(4)));
    const size_t bytes = words * sizeof(Word);
    const size_t vecs = bytes / 16;
    for (size_t i = threadIdx.x; i < vecs; i += kBinBlock) {
        reinterpret_cast<V4*>(slice)[i] = reinterpret_cast<const V4*>(g)[i];
    }
    for (size_t i = vecs * 16 / sizeof(Word) + threadIdx.x; i < words; i += kBinBlock) slice[i] = g[i];
    __syncthreads();

    uint32_t ok = 0;
    const int lane = __lane_id();
    const uint64_t ltMask = (1ull << lane) - 1;
    uint32_t parity = 0;
    for (uint32_t r0 = lo; r0 < hi; r0 += kBinRound, parity ^= 1u) {
        // 1. start words
        E ent[kBinDepth];
#pragma unroll
        for (uint32_t k = 0; k < kBinDepth; ++k) ent[k] = next[k];
        loadRound(r0 + kBinRound < hi ? r0 + kBinRound : r0);  // (past the last round: a harmless reload)
#pragma unroll
        for (uint32_t k = 0; k < kBinDepth; ++k) {
            const bool valid = r0 + k * kBinBlock + threadIdx.x < hi;
            const Tag fp = static_cast<Tag>(ent[k]);
            uint32_t local;
            if constexpr (sizeof(E) == 4) {
                local = static_cast<uint32_t>(ent[k] >> kTagBits);
            } else {
                local = static_cast<uint32_t>(static_cast<size_t>(ent[k] >> 32) - b0);
            }
            bool again = false;
            if (valid) {
                const bool placed = Ops::tryStartWordLds(slice + static_cast<size_t>(local) * WC, fp);
                ok += placed;
                again = !placed;
            }
            const uint64_t mask = __ballot(again);
            if (mask != 0) {
                // one LDS atomic per wave reserves this wave's run of the retry list
                const int leader = __ffsll(static_cast<long long>(mask)) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(&retryCount[parity], static_cast<uint32_t>(__popcll(mask)));
                base = __shfl(base, leader, 64);
                if (again) retry[base + __popcll(mask & ltMask)] = (static_cast<R>(local) << kTagBits) | static_cast<R>(fp);
            }
        }
        __syncthreads();
        // 2. the rest of the bucket for the round's retries, in full waves
        const uint32_t nr = retryCount[parity];
        if (threadIdx.x == 0) retryCount[parity ^ 1u] = 0;  // the next round's list (last read a round ago)
        for (uint32_t j0 = 0; j0 < nr; j0 += kBinBlock) {
            const uint32_t j = j0 + threadIdx.x;
            const bool valid = j < nr;
            const R rv = valid ? retry[j] : R(0);
            const uint32_t local = static_cast<uint32_t>(rv >> kTagBits);
            const Tag fp = static_cast<Tag>(rv);
            bool placed = true;
            if (valid) {
                Word* bw = slice + static_cast<size_t>(local) * WC;
                placed = Ops::tryInsertWordsLds(bw, fp);
                ok += placed;
            }
            appendOverflow(!placed, ((b0 + local) << 32) | static_cast<uint64_t>(fp), ovBuf, &ovCount,
                           kOverflowCap, overflow, overflowCount);
        }
        __syncthreads();  // the retry list is refilled by the next round
        if (r0 + kBinRound < hi) {
            flushOverflowBetweenRounds(ovBuf, &ovCount, kOverflowCap, overflow, overflowCount, &ovBase);
        }
    }
    const uint32_t buffered = ovCount < kOverflowCap ? ovCount : kOverflowCap;
    if (threadIdx.x == 0 && buffered != 0) ovBase = atomicAdd(overflowCount, static_cast<unsigned long long>(buffered));
    for (size_t i = threadIdx.x; i < vecs; i += kBinBlock) {
        reinterpret_cast<V4*>(gw)[i] = reinterpret_cast<const V4*>(slice)[i];
    }
    for (size_t i = vecs * 16 / sizeof(Word) + threadIdx.x; i < words; i += kBinBlock) gw[i] = slice[i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < buffered; i += kBinBlock) overflow[ovBase + i] = ovBuf[i];
    detail::blockCountAdd(ok, v.occupied, false);
}

// Step 3 runs as three kernels over compacted lists, so that the lanes of a
// wave do the same kind of work: the secondary-bucket attempt of every
// overflow entry; the BFS stage of the entries whose secondary bucket was
// full; DFS chains for the entries BFS could not place.  Each list entry is
// (bucket << 32 | fingerprint as stored in that bucket).  Deferring a key
// between these steps is again an interleaving the concurrent reference
// allows.

inline constexpr int kStageBlock = 1024;
inline constexpr uint32_t kStageChunk = 2048;  // entries per workgroup round (LDS list buffer)

/// Appends the values of the lanes with `emit` to the workgroup's LDS list
/// (one LDS atomic per wave).
__device__ __forceinline__ void waveAppend(bool emit, uint64_t value, uint64_t* buf, uint32_t* bufCount) {
    const uint64_t mask = __ballot(emit);
    if (mask == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(bufCount, static_cast<uint32_t>(__popcll(mask)));
    base = __shfl(base, leader, 64);
    if (emit) buf[base + __popcll(mask & ((1ull << lane) - 1))] = value;
}

/// Secondary-bucket attempt of each overflow entry (i1 << 32 | fp); entries
/// whose secondary bucket is full go to `next` as their eviction start.
template <typename Config>
__global__ __launch_bounds__(kStageBlock) void secondaryKernel(const uint64_t* __restrict__ in,
                                                               const unsigned long long* __restrict__ inCount,
                                                               uint64_t* __restrict__ next,
                                                               unsigned long long* nextCount, DeviceView<Config> v) {
    using Ops = FilterOps<Config>;
    using Tag = typename Config::TagType;
    __shared__ uint64_t buf[kStageChunk];
    __shared__ uint32_t bufCount;
    __shared__ unsigned long long base;
    const size_t n = static_cast<size_t>(*inCount);
    uint32_t ok = 0;
    for (size_t c0 = static_cast<size_t>(blockIdx.x) * kStageChunk; c0 < n;
         c0 += static_cast<size_t>(gridDim.x) * kStageChunk) {
        if (threadIdx.x == 0) bufCount = 0;
        __syncthreads();
        const size_t c1 = n - c0 < kStageChunk ? n : c0 + kStageChunk;
        for (size_t i = c0 + threadIdx.x; i < c1; i += kStageBlock) {
            const uint64_t entry = in[i];
            const size_t i1 = static_cast<size_t>(entry >> 32);
            const Tag fp = static_cast<Tag>(entry);
            const auto second = Ops::altWithNewFp(i1, fp, v.numBuckets);
            const bool placed = Ops::tryInsertAtBucket(v, second.bucket, second.fp);
            ok += placed;
            const auto e = Ops::evictStart(i1, fp, second);
            waveAppend(!placed, (static_cast<uint64_t>(e.bucket) << 32) | static_cast<uint64_t>(e.fp), buf,
                       &bufCount);
        }
        __syncthreads();
        if (threadIdx.x == 0 && bufCount != 0) base = atomicAdd(nextCount, static_cast<unsigned long long>(bufCount));
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < bufCount; j += kStageBlock) next[base + j] = buf[j];
        __syncthreads();
    }
    detail::blockCountAdd(ok, v.occupied, false);
}

/// BFS stage of each eviction start (bucket << 32 | fp); unplaced entries go
/// to `next` for DFS.  With the DFS policy everything goes straight on.
template <typename Config>
__global__ __launch_bounds__(kStageBlock) void bfsStageKernel(const uint64_t* __restrict__ in,
                                                              const unsigned long long* __restrict__ inCount,
                                                              uint64_t* __restrict__ next,
                                                              unsigned long long* nextCount, DeviceView<Config> v) {
    using Ops = FilterOps<Config>;
    using Tag = typename Config::TagType;
    __shared__ uint64_t buf[kStageChunk];
    __shared__ uint32_t bufCount;
    __shared__ unsigned long long base;
    const size_t n = static_cast<size_t>(*inCount);
    uint32_t ok = 0;
    uint32_t ev = 0;
    for (size_t c0 = static_cast<size_t>(blockIdx.x) * kStageChunk; c0 < n;
         c0 += static_cast<size_t>(gridDim.x) * kStageChunk) {
        if (threadIdx.x == 0) bufCount = 0;
        __syncthreads();
        const size_t c1 = n - c0 < kStageChunk ? n : c0 + kStageChunk;
        for (size_t i = c0 + threadIdx.x; i < c1; i += kStageBlock) {
            const uint64_t entry = in[i];
            uint32_t chain = 0;
            const bool placed = Ops::bfsCandidates(v, static_cast<Tag>(entry), static_cast<size_t>(entry >> 32), chain);
            ok += placed;
            ev += chain;
            if (placed && v.chainHist != nullptr) atomicAdd(&v.chainHist[1], 1u);
            waveAppend(!placed, entry, buf, &bufCount);
        }
        __syncthreads();
        if (threadIdx.x == 0 && bufCount != 0) base = atomicAdd(nextCount, static_cast<unsigned long long>(bufCount));
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < bufCount; j += kStageBlock) next[base + j] = buf[j];
        __syncthreads();
    }
    detail::blockCountAdd(ok, v.occupied, false);
    if (v.evictions != nullptr) detail::blockCountAdd(ev, v.evictions, false);
}

/// DFS chains (insertWithEvictionDFS) from each eviction start.
template <typename Config>
__global__ __launch_bounds__(kMaxBlock) void dfsStageKernel(const uint64_t* __restrict__ in,
                                                            const unsigned long long* __restrict__ inCount,
                                                            DeviceView<Config> v) {
    using Ops = FilterOps<Config>;
    using Tag = typename Config::TagType;
    const size_t n = static_cast<size_t>(*inCount);
    uint32_t ok = 0;
    uint32_t ev = 0;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t entry = in[i];
        uint32_t chain = 0;
        const bool s = Ops::insertDFS(v, static_cast<Tag>(entry), static_cast<size_t>(entry >> 32), chain);
        ok += s;
        ev += chain;
        if (v.chainHist != nullptr && (chain != 0 || !s)) {
            atomicAdd(&v.chainHist[s ? chain : v.maxEvictions + 1], 1u);
        }
    }
    detail::blockCountAdd(ok, v.occupied, false);
    if (v.evictions != nullptr) detail::blockCountAdd(ev, v.evictions, false);
}

// ---------------------------------------------------------------------------
// Bucket-binned delete (deleteMany, CuckooFilter.cuh:536-548, 1106-1116,
// 1157-1183).  The same partition groups the keys' (primary bucket,
// fingerprint) entries by table slice; one workgroup per slice removes every
// entry's tag from its primary bucket inside LDS, the entries not found there
// try their secondary bucket with device-scope CAS, and the entries not found
// in either are the failed deletes.  Running all first removals, then all
// second ones, is an interleaving the concurrent deleteKernel allows.
//
// Results: the entries carry no key index, so a delete's result cannot be
// stored at its key's place by the slice that decides it (and random byte
// stores would cost more than the delete itself).  Instead the failed
// deletes' (primary bucket, fingerprint) pairs go into a hash table, and one
// streaming pass over the keys writes true for every key except those that
// claim one of the table's entries.  Keys with equal pairs are
// interchangeable -- their deletes remove the same tag from the same two
// buckets -- so which of them reports the failure is one of the orders the
// concurrent reference allows.  With no failures the pass only writes true.
// ---------------------------------------------------------------------------

/// Step 2 of the delete: first-bucket removals of one slice's entries in LDS;
/// entries not found go to `overflow` as (bucket << 32 | fingerprint).
template <typename Config, typename E = uint64_t>
__global__ __launch_bounds__(kBinBlock) void binDeleteKernel(const E* __restrict__ entries,
                                                             const uint32_t* __restrict__ start, int log2Slice,
                                                             DeviceView<Config> v, uint64_t* __restrict__ overflow,
                                                             unsigned long long* overflowCount) {
    using Ops = FilterOps<Config>;
    using Word = typename Config::WordType;
    using Tag = typename Config::TagType;
    constexpr size_t WC = Layout<Config>::wordCount;
    constexpr size_t kBucketBytes = Layout<Config>::bucketBytes;
    constexpr int kTagBits = 8 * sizeof(Tag);
    extern __shared__ __align__(16) unsigned char ldsRaw[];
    __shared__ uint32_t ovCount;
    __shared__ unsigned long long ovBase;

    const uint32_t s = blockIdx.x;
    const uint32_t lo = start[s];
    const uint32_t hi = start[s + 1];
    if (lo == hi) return;

    const size_t sliceBuckets = static_cast<size_t>(1) << log2Slice;
    Word* slice = reinterpret_cast<Word*>(ldsRaw);
    uint64_t* ovBuf = reinterpret_cast<uint64_t*>(ldsRaw + binOverflowOffset(sliceBuckets * kBucketBytes));
    constexpr uint32_t kOverflowCap = binOverflowCapacity(0);
    const size_t b0 = static_cast<size_t>(s) << log2Slice;
    const size_t nbs = v.numBuckets - b0 < sliceBuckets ? v.numBuckets - b0 : sliceBuckets;
    const size_t words = nbs * WC;
    Word* gw = v.words + b0 * WC;

    if (threadIdx.x == 0) ovCount = 0;
    const uint32_t last = hi - 1;
    E next[kBinDepth];
    auto loadRound = [&](uint32_t r0) {
#pragma unroll
        for (uint32_t k = 0; k < kBinDepth; ++k) {
            const uint32_t e = r0 + k * kBinBlock + threadIdx.x;
            next[k] = entries[e < last ? e : last];
        }
    };
    loadRound(lo);
    using V4 = unsigned int __attribute__((ext_vector_type(4)));
    const size_t vecs = words * sizeof(Word) / 16;
    for (size_t i = threadIdx.x; i < vecs; i += kBinBlock) {
        reinterpret_cast<V4*>(slice)[i] = reinterpret_cast<const V4*>(gw)[i];
    }
    for (size_t i = vecs * 16 / sizeof(Word) + threadIdx.x; i < words; i += kBinBlock) slice[i] = gw[i];
    __syncthreads();

    uint32_t ok = 0;
    for (uint32_t r0 = lo; r0 < hi; r0 += kBinRound) {
        E ent[kBinDepth];
#pragma unroll
        for (uint32_t k = 0; k < kBinDepth; ++k) ent[k] = next[k];
        loadRound(r0 + kBinRound < hi ? r0 + kBinRound : r0);
#pragma unroll
        for (uint32_t k = 0; k < kBinDepth; ++k) {
            const bool valid = r0 + k * kBinBlock + threadIdx.x < hi;
            const Tag fp = static_cast<Tag>(ent[k]);
            uint32_t local;
            if constexpr (sizeof(E) == 4) {
                local = static_cast<uint32_t>(ent[k] >> kTagBits);
            } else {
                local = static_cast<uint32_t>(static_cast<size_t>(ent[k] >> 32) - b0);
            }
            bool removed = true;
            if (valid) {
                Word* bw = slice + static_cast<size_t>(local) * WC;
                removed = Ops::tryRemoveWordsLds(bw, fp);
                ok += removed;
            }
            appendOverflow(!removed, ((b0 + local) << 32) | static_cast<uint64_t>(fp), ovBuf, &ovCount,
                           kOverflowCap, overflow, overflowCount);
        }
        if (r0 + kBinRound < hi) {
            __syncthreads();
            flushOverflowBetweenRounds(ovBuf, &ovCount, kOverflowCap, overflow, overflowCount, &ovBase);
        }
    }
    __syncthreads();
    const uint32_t buffered = ovCount < kOverflowCap ? ovCount : kOverflowCap;
    if (threadIdx.x == 0 && buffered != 0) ovBase = atomicAdd(overflowCount, static_cast<unsigned long long>(buffered));
    for (size_t i = threadIdx.x; i < vecs; i += kBinBlock) {
        reinterpret_cast<V4*>(gw)[i] = reinterpret_cast<const V4*>(slice)[i];
    }
    for (size_t i = vecs * 16 / sizeof(Word) + threadIdx.x; i < words; i += kBinBlock) gw[i] = slice[i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < buffered; i += kBinBlock) overflow[ovBase + i] = ovBuf[i];
    detail::blockCountAdd(ok, v.occupied, true);
}

/// Step 3 of the delete: the secondary-bucket removal of each overflow entry
/// (i1 << 32 | fp); entries found in neither bucket go to `failed` unchanged.
template <typename Config>
__global__ __launch_bounds__(kStageBlock) void secondaryDeleteKernel(const uint64_t* __restrict__ in,
                                                                     const unsigned long long* __restrict__ inCount,
                                                                     uint64_t* __restrict__ failed,
                                                                     unsigned long long* failedCount,
                                                                     DeviceView<Config> v) {
    using Ops = FilterOps<Config>;
    using Tag = typename Config::TagType;
    __shared__ uint64_t buf[kStageChunk];
    __shared__ uint32_t bufCount;
    __shared__ unsigned long long base;
    const size_t n = static_cast<size_t>(*inCount);
    uint32_t ok = 0;
    for (size_t c0 = static_cast<size_t>(blockIdx.x) * kStageChunk; c0 < n;
         c0 += static_cast<size_t>(gridDim.x) * kStageChunk) {
        if (threadIdx.x == 0) bufCount = 0;
        __syncthreads();
        const size_t c1 = n - c0 < kStageChunk ? n : c0 + kStageChunk;
        for (size_t i = c0 + threadIdx.x; i < c1; i += kStageBlock) {
            const uint64_t entry = in[i];
            const size_t i1 = static_cast<size_t>(entry >> 32);
            const Tag fp = static_cast<Tag>(entry);
            const auto second = Ops::altWithNewFp(i1, fp, v.numBuckets);
            const bool removed = Ops::tryRemoveAtBucket(v, second.bucket, second.fp);
            ok += removed;
            waveAppend(!removed, entry, buf, &bufCount);
        }
        __syncthreads();
        if (threadIdx.x == 0 && bufCount != 0) base = atomicAdd(failedCount, static_cast<unsigned long long>(bufCount));
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < bufCount; j += kStageBlock) failed[base + j] = buf[j];
        __syncthreads();
    }
    detail::blockCountAdd(ok, v.occupied, true);
}

/// Hash table of failed deletes: open addressing over 64-bit (i1 << 32 | fp)
/// entries, one slot per failure (equal pairs take separate slots).  A
/// fingerprint is never 0, so 0 marks an empty slot and a value with a zero
/// fingerprint field marks a slot a key has claimed.
inline constexpr uint64_t kFailEmpty = 0;
inline constexpr uint64_t kFailClaimed = 0xFFFFFFFF00000000ull;

/// Table slots for f failures: the power of two at or above 2f (at least 64).
__device__ __forceinline__ uint64_t failTableSlots(unsigned long long f) {
    uint64_t c = 64;
    while (c < 2 * static_cast<uint64_t>(f)) c <<= 1;
    return c;
}

__device__ __forceinline__ uint64_t failHash(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    return x;
}

/// (Templates only so that every translation unit may instantiate them.)
template <typename C = unsigned long long>
__global__ __launch_bounds__(256) void failTableClearKernel(uint64_t* __restrict__ table,
                                                            const C* __restrict__ failedCount) {
    const unsigned long long f = *failedCount;
    if (f == 0) return;
    const uint64_t slots = failTableSlots(f);
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < slots;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        table[i] = kFailEmpty;
    }
}

template <typename C = unsigned long long>
__global__ __launch_bounds__(256) void failTableBuildKernel(uint64_t* __restrict__ table,
                                                            const uint64_t* __restrict__ failed,
                                                            const C* __restrict__ failedCount) {
    const unsigned long long f = *failedCount;
    if (f == 0) return;
    const uint64_t mask = failTableSlots(f) - 1;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < f;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t key = failed[i];
        for (uint64_t h = failHash(key) & mask;; h = (h + 1) & mask) {
            unsigned long long expected = kFailEmpty;
            if (__hip_atomic_compare_exchange_strong(reinterpret_cast<unsigned long long*>(table + h), &expected,
                                                     static_cast<unsigned long long>(key), __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                break;
            }
        }
    }
}

/// deleteMany's per-key results: true, except for keys that claim a failed
/// entry with their (i1, fp) -- as many keys per pair as deletes of that
/// pair failed.
template <typename Config>
__global__ __launch_bounds__(256) void deleteResultKernel(const typename Config::KeyType* __restrict__ keys,
                                                          size_t n, size_t numBuckets, bool* __restrict__ out,
                                                          uint64_t* __restrict__ table,
                                                          const unsigned long long* __restrict__ failedCount) {
    const unsigned long long f = *failedCount;
    const uint64_t mask = f != 0 ? failTableSlots(f) - 1 : 0;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        bool ok = true;
        if (f != 0) {
            const auto c = Config::AltBucketPolicy::getCandidateBuckets(keys[i], numBuckets);
            const uint64_t key = (static_cast<uint64_t>(c.i1) << 32) | static_cast<uint64_t>(c.fp);
            for (uint64_t h = failHash(key) & mask;; h = (h + 1) & mask) {
                unsigned long long cur = __hip_atomic_load(reinterpret_cast<unsigned long long*>(table + h),
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cur == kFailEmpty) break;
                if (cur == key &&
                    __hip_atomic_compare_exchange_strong(reinterpret_cast<unsigned long long*>(table + h), &cur,
                                                         kFailClaimed, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)) {
                    ok = false;
                    break;
                }
            }
        }
        out[i] = ok;
    }
}

namespace detail {

inline int log2Floor(size_t v) {
    int b = 0;
    while ((size_t(2) << b) <= v) ++b;
    return b;
}

/// Insert path selection: CUCKOO_INSERT_PATH = "direct" | "binned" |
/// "presorted" | "auto" (default auto: binned once the batch holds at least
/// kBinnedMinInsertKeys keys and one key per kBinnedMinBucketsPerKey
/// buckets).  "presorted" sends
/// insertManySorted down the reference's own route -- packed tags, a full
/// ascending sort, insertKernelSorted -- at every table size.
enum class InsertPath { Auto, Direct, Binned, Presorted };

inline InsertPath insertPathFromEnv() {
    const char* e = std::getenv("CUCKOO_INSERT_PATH");
    if (e == nullptr) return InsertPath::Auto;
    if (std::strcmp(e, "direct") == 0) return InsertPath::Direct;
    if (std::strcmp(e, "binned") == 0) return InsertPath::Binned;
    if (std::strcmp(e, "presorted") == 0) return InsertPath::Presorted;
    return InsertPath::Auto;
}

/// Delete path selection: CUCKOO_DELETE_PATH = "direct" | "binned" | "auto"
/// (default auto: binned from kBinnedMinDeleteKeys keys and one key per
/// kBinnedMinBucketsPerKey buckets).
inline InsertPath deletePathFromEnv() {
    const char* e = std::getenv("CUCKOO_DELETE_PATH");
    if (e == nullptr) return InsertPath::Auto;
    if (std::strcmp(e, "direct") == 0) return InsertPath::Direct;
    if (std::strcmp(e, "binned") == 0) return InsertPath::Binned;
    return InsertPath::Auto;
}

/// The binned path reads and writes every table slice it has entries for, so
/// it pays off once a batch is large against the table: measured at 2^28
/// slots, the two paths cost the same at about one key per two buckets for
/// fills of 0.3 to 0.8 (the binned path wins from about one per three at 0.9).
inline constexpr size_t kBinnedMinBucketsPerKey = 2;

/// Smallest batches the binned paths take by default: below them the direct
/// kernels are faster at any table size, the binned pipeline's dozen launches
/// outweighing its savings.  Measured on the reference's benchmark matrix
/// (n = 95% of 2^16..2^28 slots): the binned insert breaks even at about 2^20
/// keys (5.2 Gkeys/s either way at 2^20 slots; direct 1.9x faster at 2^16),
/// the binned delete between 1M and 2M keys (direct 16% faster at 996k keys,
/// 2.4x at 62k; binned 6% faster at 1.99M).
inline constexpr size_t kBinnedMinInsertKeys = size_t(1) << 20;
inline constexpr size_t kBinnedMinDeleteKeys = size_t(3) << 19;

/// CUCKOO_PARTITION=lsd keeps the one-level-per-pass LSD partition even where
/// the two-level one applies (for comparisons).
inline bool forceOnePassFromEnv() {
    const char* e = std::getenv("CUCKOO_PARTITION");
    return e != nullptr && std::strcmp(e, "lsd") == 0;
}

}  // namespace detail

/// Host orchestration of the binned insert for one configuration.
template <typename Config>
struct BinnedInsert {
    using Key = typename Config::KeyType;
    using View = DeviceView<Config>;
    static constexpr size_t kBucketBytes = Layout<Config>::bucketBytes;

    /// Whether the table can be processed by this path (32-bit bucket field).
    static bool applicable(size_t numBuckets) {
        return numBuckets <= (size_t(1) << 32) && kBucketBytes <= kSliceBytes;
    }

    static int log2Slice(size_t numBuckets) {
        int l = detail::log2Floor(kSliceBytes / kBucketBytes);
        const int nbBits = detail::bitsFor(numBuckets);
        return l < nbBits ? l : nbBits;
    }

    static size_t numSlices(size_t numBuckets) {
        const size_t S = size_t(1) << log2Slice(numBuckets);
        return (numBuckets + S - 1) / S;
    }

    static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

    static size_t workspaceBytes(size_t n, size_t numBuckets) {
        return 2 * align256(n * sizeof(uint64_t)) + align256(radix::twoLevelScratchBytes(n)) +
               align256((numSlices(numBuckets) + 1) * sizeof(uint32_t)) + align256(3 * sizeof(unsigned long long));
    }

    /// The delete's workspace with per-key results: the insert's, plus the
    /// failed-delete table (a power of two of slots at or above 2x the
    /// failures, at most 4n slots of 8 bytes).
    static size_t deleteWorkspaceBytes(size_t n, size_t numBuckets, bool results) {
        return workspaceBytes(n, numBuckets) + (results ? align256((4 * n + 64) * sizeof(uint64_t)) : 0);
    }

    /// Whether the slice index splits into two digits: then the partition is
    /// two-level (MSD).  Where they fit in 32 bits its values are (bucket bits
    /// below the first digit, fingerprint) after the first pass and (bucket
    /// inside the slice, fingerprint) after the second; otherwise they stay
    /// whole (bucket << 32 | fingerprint) entries.
    static constexpr int kTagBits = 8 * sizeof(typename Config::TagType);
    static bool twoLevel(int sliceBits) {
        return sliceBits > radix::kDigitBits && sliceBits <= 2 * radix::kDigitBits && !detail::forceOnePassFromEnv();
    }
    /// Whether the two-level partition's values fit in 32 bits (else they stay
    /// whole 64-bit entries).
    static bool twoLevelPacked(int l2s, int sliceBits) {
        return l2s + (sliceBits - radix::kDigitBits) + kTagBits <= 32;
    }

    /// The slice kernel of step 2: primary-bucket insert attempts, or
    /// first-bucket removals for the delete.
    template <bool Delete, typename E>
    static void launchSlices(const E* entries, const uint32_t* start, int l2s, const View& v, uint64_t* overflow,
                             unsigned long long* counters, size_t slices, hipStream_t s) {
        const size_t sliceBytes = (size_t(1) << l2s) * kBucketBytes;
        if constexpr (Delete) {
            binDeleteKernel<Config, E><<<static_cast<uint32_t>(slices), kBinBlock, binDeleteLdsBytes(sliceBytes), s>>>(
                entries, start, l2s, v, overflow, counters);
        } else {
            const size_t lds = binLdsBytes(sliceBytes, sizeof(BinRetryOf<Config>));
            binInsertKernel<Config, E><<<static_cast<uint32_t>(slices), kBinBlock, lds, s>>>(entries, start, l2s, v,
                                                                                           overflow, counters);
        }
    }

    /// Bucket-binned insertMany / insertManySorted (Delete = false), or
    /// deleteMany (Delete = true; `out` may be nullptr).
    template <bool SortedSemantics, bool Delete = false>
    static hipError_t run(const View& v, const Key* keys, size_t n, uint32_t block, hipStream_t s, Workspace& ws,
                          bool* out = nullptr) {
        constexpr size_t kChunk = size_t(1) << 31;
        const int l2s = log2Slice(v.numBuckets);
        const size_t slices = numSlices(v.numBuckets);
        const int sliceBits = detail::bitsFor(slices);
        for (size_t off = 0; off < n; off += kChunk) {
            const size_t m = n - off < kChunk ? n - off : kChunk;
            const size_t need = Delete ? deleteWorkspaceBytes(m, v.numBuckets, out != nullptr)
                                       : workspaceBytes(m, v.numBuckets);
            hipError_t e = ws.reserve(need);
            if (e != hipSuccess) return e;
            char* p = static_cast<char*>(ws.ptr);
            uint64_t* a = reinterpret_cast<uint64_t*>(p);
            p += align256(m * sizeof(uint64_t));
            uint64_t* b = reinterpret_cast<uint64_t*>(p);
            p += align256(m * sizeof(uint64_t));
            void* scratch = p;
            p += align256(radix::twoLevelScratchBytes(m));
            uint32_t* start = reinterpret_cast<uint32_t*>(p);
            p += align256((slices + 1) * sizeof(uint32_t));
            unsigned long long* counters = reinterpret_cast<unsigned long long*>(p);  // overflow, evict/fail, dfs
            p += align256(3 * sizeof(unsigned long long));
            uint64_t* failTable = reinterpret_cast<uint64_t*>(p);  // delete with results only

            EntrySource<Config, SortedSemantics> src{keys + off, v.numBuckets};
            // after step 2 the overflow list is in `ov`; `other` is free
            uint64_t* ov = nullptr;
            uint64_t* other = nullptr;
            if (twoLevel(sliceBits) && !twoLevelPacked(l2s, sliceBits)) {
                // the first pass keeps whole 64-bit entries in a; the second
                // writes b, in 32 bits (the bucket's place in its slice beside
                // the tag) where they fit, else whole entries
                const int d2 = sliceBits - radix::kDigitBits;
                zeroCountersKernel<unsigned long long><<<1, 64, 0, s>>>(counters);
                bool packedSecond = false;
                if constexpr (kTagBits < 32) {  // 32-bit tags never fit beside a bucket index in 32 bits
                    if (l2s + kTagBits <= 32) {
                        const PackBelowDigit pack2{(uint64_t(1) << l2s) - 1, kTagBits};
                        uint32_t* b32 = reinterpret_cast<uint32_t*>(b);
                        e = radix::partitionTwoLevel<uint64_t>(src, m, 32 + l2s + d2, radix::Identity<uint64_t>{},
                                                               32 + l2s, uint32_t(1) << d2, pack2, a, b32, scratch,
                                                               start, static_cast<uint32_t>(slices), s);
                        if (e != hipSuccess) return e;
                        launchSlices<Delete>(b32, start, l2s, v, a, counters, slices, s);
                        packedSecond = true;
                    }
                }
                if (!packedSecond) {
                    e = radix::partitionTwoLevel<uint64_t>(src, m, 32 + l2s + d2, radix::Identity<uint64_t>{},
                                                           32 + l2s, uint32_t(1) << d2, radix::Identity<uint64_t>{},
                                                           a, b, scratch, start, static_cast<uint32_t>(slices), s);
                    if (e != hipSuccess) return e;
                    launchSlices<Delete>(b, start, l2s, v, a, counters, slices, s);
                }
                ov = a;
                other = b;
            } else if (kTagBits < 32 && twoLevel(sliceBits)) {
                // a holds the first pass's 32-bit values, b the second's; after
                // the slices the overflow list goes to a
                if constexpr (kTagBits < 32) {
                    const int d2 = sliceBits - radix::kDigitBits;
                    const PackBelowDigit pack1{(uint64_t(1) << (l2s + d2)) - 1, kTagBits};
                    const KeepLowBits pack2{(uint32_t(1) << (l2s + kTagBits)) - 1};
                    uint32_t* a32 = reinterpret_cast<uint32_t*>(a);
                    uint32_t* b32 = reinterpret_cast<uint32_t*>(b);
                    zeroCountersKernel<unsigned long long><<<1, 64, 0, s>>>(counters);
                    e = radix::partitionTwoLevel<uint64_t>(src, m, 32 + l2s + d2, pack1, l2s + kTagBits,
                                                           uint32_t(1) << d2, pack2, a32, b32, scratch, start,
                                                           static_cast<uint32_t>(slices), s);
                    if (e != hipSuccess) return e;
                    launchSlices<Delete>(b32, start, l2s, v, a, counters, slices, s);
                    ov = a;
                    other = b;
                }
            } else {
                uint64_t* sorted = a;
                if (sliceBits > 0) {
                    e = radix::sortBits<uint64_t>(src, m, 32 + l2s, 32 + l2s + sliceBits, a, b, scratch, s, &sorted);
                    if (e != hipSuccess) return e;
                } else {
                    auto k = packEntriesKernel<decltype(src)>;
                    k<<<detail::gridFor(k, m, 256), 256, 0, s>>>(src, m, a);
                }
                uint64_t* spare = sorted == a ? b : a;
                sliceStartKernel<uint64_t><<<static_cast<uint32_t>((slices + 1 + 255) / 256), 256, 0, s>>>(
                    sorted, static_cast<uint32_t>(m), 32 + l2s, static_cast<uint32_t>(slices), start, counters);
                launchSlices<Delete>(sorted, start, l2s, v, spare, counters, slices, s);
                ov = spare;
                other = sorted;  // free once the slices have read it
            }
            if constexpr (Delete) {
                e = runDeleteStages(v, keys + off, m, s, counters, ov, other, failTable,
                                    out != nullptr ? out + off : nullptr);
            } else {
                e = runStages(v, m, block, s, counters, ov, other);
            }
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }

    /// The delete after its slices: secondary removals of the overflow list,
    /// then (with `out`) the per-key results from the failed entries.
    static hipError_t runDeleteStages(const View& v, const Key* keys, size_t m, hipStream_t s,
                                      unsigned long long* counters, uint64_t* overflow, uint64_t* failed,
                                      uint64_t* failTable, bool* out) {
        const uint32_t stageGrid = detail::gridFor(secondaryDeleteKernel<Config>,
                                                   (m + kStageChunk - 1) / kStageChunk * kStageBlock, kStageBlock);
        secondaryDeleteKernel<Config><<<stageGrid, kStageBlock, 0, s>>>(overflow, counters, failed, counters + 1, v);
        if (out == nullptr) return hipGetLastError();
        auto clear = failTableClearKernel<unsigned long long>;
        auto build = failTableBuildKernel<unsigned long long>;
        clear<<<detail::gridFor(clear, 4 * m + 64, 256), 256, 0, s>>>(failTable, counters + 1);
        build<<<detail::gridFor(build, m, 256), 256, 0, s>>>(failTable, failed, counters + 1);
        auto k = deleteResultKernel<Config>;
        k<<<detail::gridFor(k, m, 256), 256, 0, s>>>(keys, m, v.numBuckets, out, failTable, counters + 1);
        return hipGetLastError();
    }

    /// Step 3 over the overflow list in `overflow`; `other` receives the
    /// eviction starts, then `overflow` the DFS list.
    static hipError_t runStages(const View& v, size_t m, uint32_t block, hipStream_t s, unsigned long long* counters,
                                uint64_t* overflow, uint64_t* other) {
        const uint32_t stageGrid = detail::gridFor(secondaryKernel<Config>,
                                                   (m + kStageChunk - 1) / kStageChunk * kStageBlock, kStageBlock);
        secondaryKernel<Config><<<stageGrid, kStageBlock, 0, s>>>(overflow, counters, other, counters + 1, v);
        if (v.evictionPolicy == static_cast<uint32_t>(EvictionPolicy::BFS)) {
            bfsStageKernel<Config><<<stageGrid, kStageBlock, 0, s>>>(other, counters + 1, overflow, counters + 2, v);
            auto k = dfsStageKernel<Config>;
            k<<<detail::gridFor(k, m, block), block, 0, s>>>(overflow, counters + 2, v);
        } else {
            auto k = dfsStageKernel<Config>;
            k<<<detail::gridFor(k, m, block), block, 0, s>>>(other, counters + 1, v);
        }
        return hipGetLastError();
    }
};

}  // namespace cuckoo
