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
#include <type_traits>
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

namespace detail {

/// Orders the LDS accesses of one wave's lanes around it (a compiler barrier
/// with wavefront-scope release and acquire): LDS executes a wave's
/// operations in program order, so a lane reads what another lane of its
/// wave wrote before the sync.
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/// How an insert's attempt at its primary bucket ended: stored; the bucket
/// was full (the insert goes on at the secondary bucket, then eviction); or
/// the bucket changed under the attempt (the insert starts over).  A pending
/// insert may also be a BFS search paused between candidates (kRunBfs) or a
/// DFS chain paused between swaps (kRunResume).
enum RunOutcome : uint32_t { kRunPlaced = 0, kRunFull = 1, kRunRetry = 2, kRunResume = 3, kRunBfs = 4 };

/// DFS swaps a pending insert makes per round of its wave's queue (at 2^28
/// slots, 95% load, DFS: 1, 2, 4 and 8 swaps per round took 5.26, 5.36, 5.60
/// and 6.08 ms for the presorted insert of 255M keys).
inline constexpr uint32_t kSwapsPerRound = 1;

/// LDS slots of a workgroup's pending inserts: 64 per wave, at the wave's
/// slots of the workgroup (Slots >= blockDim.x).
template <uint32_t Slots>
struct PendingLds {
    unsigned long long bucket[Slots];
    uint32_t tag[Slots];
    uint32_t swaps[Slots];  // a paused chain's swaps so far
    uint8_t kind[Slots];
};

/// One wave's inserts whose primary attempt did not finish.  They continue
/// together, one per lane, once 64 have gathered (and at the end), instead of
/// a few lanes at a time while the rest of the wave waits for them.  An
/// insert that goes on to eviction tries one BFS candidate (BFS policy) or
/// makes at most kSwapsPerRound DFS swaps per round and, if not done, goes
/// back into the queue with its tag, bucket and candidate index or swap
/// count: a round costs the wave one step rather than the longest of 64
/// searches.  Continuing an insert later, or
/// pausing its chain between swaps, is an order the reference's concurrent
/// kernel allows (its threads run at independent speeds); every chain still
/// makes the reference's swaps in the reference's order, and finish() runs
/// rounds until none is left.  push() and finish() are called by every lane
/// of the wave at the same points.
template <typename Config, uint32_t Slots>
struct PendingInserts {
    using Ops = FilterOps<Config>;
    using Tag = typename Config::TagType;
    PendingLds<Slots>& q;
    uint32_t lane;
    uint32_t waveSlots;  // the wave's first slot
    uint32_t waveLanes;  // lanes present (fewer than 64 in the last wave of an odd-sized workgroup)
    uint32_t queued = 0;  // wave-uniform

    __device__ explicit PendingInserts(PendingLds<Slots>& lds)
        : q(lds),
          lane(__lane_id()),
          waveSlots(threadIdx.x - __lane_id()),
          waveLanes(static_cast<uint32_t>(__popcll(__ballot(1)))) {}

    template <typename Account>
    __device__ void push(const DeviceView<Config>& v, bool defer, size_t bucket, Tag tag, uint32_t kind,
                         Account& account) {
        const uint64_t deferred = __ballot(defer);
        const uint32_t count = static_cast<uint32_t>(__popcll(deferred));
        while (queued + count > 64) round(v, account);
        if (defer) {
            const uint32_t slot = waveSlots + queued + static_cast<uint32_t>(__popcll(deferred & ((1ull << lane) - 1)));
            q.bucket[slot] = bucket;
            q.tag[slot] = tag;
            q.swaps[slot] = 0;
            q.kind[slot] = static_cast<uint8_t>(kind);
        }
        queued += count;
    }

    template <typename Account>
    __device__ void finish(const DeviceView<Config>& v, Account& account) {
        while (queued != 0) round(v, account);
    }

   private:
    /// One round over the queue: each entry's insert continues as insertFrom
    /// and insertAfterPrimary go on (the primary bucket again for kRunRetry,
    /// then the secondary bucket), then takes one step of its eviction: the
    /// next BFS candidate (kRunBfs, the candidate index in `swaps`; DFS from
    /// the same start once all have failed) or up to kSwapsPerRound DFS swaps
    /// (kRunResume).  Entries not done are compacted to the front of the
    /// queue.
    template <typename Account>
    __device__ void round(const DeviceView<Config>& v, Account& account) {
        waveSync();
        uint32_t kept = 0;  // wave-uniform
        for (uint32_t first = 0; first < queued; first += waveLanes) {
            const uint32_t k = first + lane;
            bool keep = false;
            size_t b = 0;
            Tag t = 0;
            uint32_t e = 0;
            uint32_t kind = kRunResume;
            if (k < queued) {
                b = static_cast<size_t>(q.bucket[waveSlots + k]);
                t = static_cast<Tag>(q.tag[waveSlots + k]);
                e = q.swaps[waveSlots + k];
                kind = q.kind[waveSlots + k];
                uint32_t chain = 0;
                bool placed = false;
                if (kind == kRunRetry) placed = Ops::tryInsertAtBucket(v, b, t);
                if (!placed && (kind == kRunRetry || kind == kRunFull)) {
                    const auto second = Ops::altWithNewFp(b, t, v.numBuckets);
                    placed = Ops::tryInsertAtBucket(v, second.bucket, second.fp);
                    if (!placed) {
                        const auto start = Ops::evictStart(b, t, second);
                        b = start.bucket;
                        t = start.fp;
                        e = 0;
                        kind = v.evictionPolicy == static_cast<uint32_t>(EvictionPolicy::BFS) ? kRunBfs : kRunResume;
                    }
                }
                if (!placed && kind == kRunBfs) {
                    if (e < Ops::kBfsCandidates) placed = Ops::bfsCandidate(v, t, b, e++, chain);
                    if (!placed) {
                        keep = true;
                        if (e >= Ops::kBfsCandidates) {  // DFS from the same start next round
                            kind = kRunResume;
                            e = 0;
                        }
                    }
                } else if (!placed) {
                    const uint32_t limit = v.maxEvictions - e > kSwapsPerRound ? e + kSwapsPerRound : v.maxEvictions;
                    const uint32_t st = Ops::dfsSteps(v, t, b, e, limit);
                    if (st == Ops::kDfsPaused) {
                        keep = true;
                    } else {
                        account(st == Ops::kDfsPlaced, e);
                    }
                }
                if (placed) account(true, chain);
            }
            const uint64_t keeps = __ballot(keep);
            waveSync();  // this pass's entries are read before paused chains are written back
            if (keep) {
                const uint32_t slot = waveSlots + kept + static_cast<uint32_t>(__popcll(keeps & ((1ull << lane) - 1)));
                q.bucket[slot] = b;
                q.tag[slot] = t;
                q.swaps[slot] = e;
                q.kind[slot] = static_cast<uint8_t>(kind);
            }
            kept += static_cast<uint32_t>(__popcll(keeps));
        }
        waveSync();  // the queue holds the paused chains
        queued = kept;
    }
};

}  // namespace detail

/// insertKernel (CuckooFilter.cuh:1119-1141).  One insert per thread:
/// unlike the sorted kernel's runs, random keys spread a wave's work over
/// unrelated buckets, and queueing the inserts that go on to eviction
/// (detail::PendingInserts) made this kernel slower (14.4 -> 13.8 Gkeys/s at
/// 2^28 slots, 95%): other waves already hide one wave's eviction chains.
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

namespace detail {

/// LDS through which the lanes of a wave hand each other their bucket, tag,
/// slot and the run's changed words: one slot per thread of the workgroup
/// (Slots >= blockDim.x).
template <uint32_t Slots>
struct RunLds {
    unsigned long long bucket[Slots];   // i1, then the old value of the word this lane writes
    unsigned long long desired[Slots];  // the new value of the word this lane writes
    uint32_t tag[Slots];                // the lane's tag, then (at a run's first lane) its changed words
    uint8_t out[Slots];                 // the bucket slot the lane's tag takes
    uint8_t word[Slots];                // the word this lane writes
};

/// Bit s set iff slot s of the bucket snapshot is empty (slot s is lane
/// s % TPW of word s / TPW).  Exact in every lane, unlike getZeroMask.
template <typename Config>
__device__ __forceinline__ uint64_t emptySlots(typename FilterOps<Config>::SnapT s) {
    using Ops = FilterOps<Config>;
    using Word = typename Config::WordType;
    using Tag = typename Config::TagType;
    static_assert(Ops::BITS == 8 * sizeof(Tag), "tags fill their lanes");
    constexpr Word ones = cuckoo::detail::laneOnes<Tag, Word>();
    constexpr Word highs = ones << (Ops::BITS - 1);
    constexpr Word low = ~highs;
    uint64_t e = 0;
#pragma unroll
    for (uint32_t j = 0; j < Ops::WC; ++j) {
        const Word x = s[j];
        const Word zero = ~(((x & low) + low) | x | low);  // high bit of each zero lane, exactly
#pragma unroll
        for (uint32_t i = 0; i < Ops::TPW; ++i) {
            e |= static_cast<uint64_t>((zero >> (i * Ops::BITS + Ops::BITS - 1)) & 1) << (j * Ops::TPW + i);
        }
    }
    return e;
}

/// Primary-bucket attempts of one wave's inserts, shared by the lanes that
/// insert into the same bucket.  Sorted packed tags put runs of consecutive
/// lanes on one bucket (about 15 at 95% load with 16-slot buckets); one device
/// CAS per lane on the same words serialises at the memory side.  Here the
/// first lane of each run (consecutive active lanes with equal i1) reads the
/// bucket once and places the run's tags one after another in lane order as
/// tryInsertAtBucket would (CuckooFilter.cuh:859-904: the first word from the
/// tag's start word with an empty slot, its lowest empty slot), which over a
/// mask of the bucket's empty slots is the lowest set bit after rotating the
/// mask by the start word: a few integer operations per tag.  Each lane then
/// ORs its tag into the new value of its word (LDS atomics), and each changed
/// word is written with one CAS against the value read, the c-th changed word
/// by the run's c-th lane (a run of m tags changes at most m words): the
/// run's CASes are one wave instruction.  That is the run's inserts done back
/// to back, an order the reference's concurrent kernel allows.  Lanes
/// exchange data through LDS and ballots, not cross-lane permutes.  Returns
/// for the calling lane: kRunPlaced if its tag was stored, kRunFull if the
/// bucket had no room for it (the rest of its insert starts at the secondary
/// bucket), kRunRetry if a word it depended on changed under the run (its
/// insert starts over from the primary bucket).
template <typename Config, uint32_t Slots>
__device__ __forceinline__ uint32_t primaryRunInsert(const DeviceView<Config>& v, bool active, size_t i1,
                                                     typename Config::TagType fp, RunLds<Slots>& lds) {
    using Ops = FilterOps<Config>;
    using Word = typename Config::WordType;
    constexpr uint32_t WC = Ops::WC;
    constexpr uint32_t TPW = Ops::TPW;
    constexpr uint32_t BS = Ops::BS;
    static_assert(BS <= 64, "one bit per slot in 64 bits");
    // slot masks in 32 bits where the bucket allows: half the shift and bit-scan work
    using Mask = std::conditional_t<(BS <= 32), uint32_t, uint64_t>;
    constexpr uint32_t kMaskBits = 8 * sizeof(Mask);
    constexpr Mask allSlots = BS == kMaskBits ? static_cast<Mask>(~Mask(0)) : static_cast<Mask>((Mask(1) << BS) - 1);
    constexpr uint64_t wordSlots = TPW == 64 ? ~0ull : ((1ull << TPW) - 1);
    constexpr uint8_t kNone = 0xff;  // no slot (the bucket was full) / no word to write
    const uint32_t lane = __lane_id();
    const uint32_t tid = threadIdx.x;
    lds.bucket[tid] = i1;
    lds.tag[tid] = fp;
    waveSync();
    // Active lanes are a prefix of the wave: the caller's index is the wave's
    // uniform base plus the lane, compared against n.
    const bool head = active && (lane == 0 || lds.bucket[tid - 1] != static_cast<unsigned long long>(i1));
    const uint64_t act = __ballot(active);
    const uint64_t heads = __ballot(head);
    const uint64_t upTo = lane == 63 ? ~0ull : ((2ull << lane) - 1);  // lanes 0..lane
    waveSync();  // every lane has compared its neighbour's bucket before the slots take words
    if (head) {
        const uint64_t later = heads & ~upTo;
        const uint32_t runEnd = later != 0 ? static_cast<uint32_t>(__builtin_ctzll(later))
                                           : static_cast<uint32_t>(__popcll(act));
        const uint32_t len = runEnd - lane;
        const typename Ops::SnapT orig = loadBucket<Word, WC>(Ops::bucketPtr(v, i1));
        const Mask empty0 = static_cast<Mask>(emptySlots<Config>(orig));
        Mask empty = empty0;
#pragma unroll 2
        for (uint32_t k = 0; k < len; ++k) {
            uint8_t slot = kNone;
            if (empty != 0) {
                const uint32_t sh = Ops::startWord(static_cast<typename Config::TagType>(lds.tag[tid + k])) * TPW;
                Mask rot;
                if constexpr (BS == kMaskBits) {
                    rot = static_cast<Mask>((empty >> sh) | (empty << ((kMaskBits - sh) & (kMaskBits - 1))));
                } else {
                    rot = static_cast<Mask>(((empty >> sh) | (empty << (BS - sh))) & allSlots);
                }
                uint32_t low;
                if constexpr (sizeof(Mask) == 4) {
                    low = static_cast<uint32_t>(__builtin_ctz(rot));
                } else {
                    low = static_cast<uint32_t>(__builtin_ctzll(rot));
                }
                const uint32_t sl = (sh + low) & (BS - 1);
                empty &= static_cast<Mask>(~(Mask(1) << sl));
                slot = static_cast<uint8_t>(sl);
            }
            lds.out[tid + k] = slot;
        }
        const uint64_t filled = static_cast<uint64_t>(empty0 & ~empty);
        uint32_t dirty = 0;
        uint32_t c = 0;
#pragma unroll
        for (uint32_t j = 0; j < WC; ++j) {
            if ((filled >> (j * TPW)) & wordSlots) {
                dirty |= 1u << j;
                const Word o = pick<Word, WC>(orig, j);
                lds.bucket[tid + c] = static_cast<unsigned long long>(o);
                lds.desired[tid + c] = static_cast<unsigned long long>(o);
                lds.word[tid + c] = static_cast<uint8_t>(j);
                ++c;
            }
        }
        for (; c < len; ++c) lds.word[tid + c] = kNone;
        lds.tag[tid] = dirty;
    }
    waveSync();
    // every lane of a run: its slot, the run's changed words, and the lane
    // that writes its word (the run's first lane plus the word's rank)
    uint32_t slot = kNone;
    uint32_t headLane = 0;
    uint32_t dirty = 0;
    if (active) {
        headLane = 63u - static_cast<uint32_t>(__builtin_clzll(heads & upTo));
        dirty = lds.tag[tid - lane + headLane];
        slot = lds.out[tid];
        if (slot != kNone) {
            const uint32_t w = slot / TPW;
            const uint32_t rank = static_cast<uint32_t>(__popc(dirty & ((1u << w) - 1)));
            const unsigned long long bits = static_cast<unsigned long long>(static_cast<Word>(fp))
                                            << ((slot % TPW) * Ops::BITS);
            __hip_atomic_fetch_or(&lds.desired[tid - lane + headLane + rank], bits, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
    }
    waveSync();
    const uint8_t myWord = active ? lds.word[tid] : kNone;
    bool casFailed = false;
    if (myWord != kNone) {
        const Word o = static_cast<Word>(lds.bucket[tid]);
        casFailed = casRelaxed(Ops::bucketPtr(v, i1) + myWord, o, static_cast<Word>(lds.desired[tid])) != o;
    }
    const uint64_t failedLanes = __ballot(casFailed);
    if (!active) return kRunPlaced;  // not used
    if (slot != kNone) {
        const uint32_t rank = static_cast<uint32_t>(__popc(dirty & ((1u << (slot / TPW)) - 1)));
        return ((failedLanes >> (headLane + rank)) & 1) ? kRunRetry : kRunPlaced;
    }
    const uint32_t writers = static_cast<uint32_t>(__popc(dirty));
    const uint64_t runWriters = writers == 0 ? 0 : (((1ull << writers) - 1) << headLane);
    return (failedLanes & runWriters) != 0 ? kRunRetry : kRunFull;
}

/// The body of insertKernelSorted with LDS for workgroups of up to Slots
/// threads.  Buckets of up to 64 bytes take their primary attempts per run of
/// lanes on one bucket (primaryRunInsert); the secondary bucket and eviction
/// follow per lane, as in the reference.  128-byte buckets keep one insert per
/// lane: two register copies of such a bucket spill to scratch.
template <typename Config, uint32_t Slots>
__device__ __forceinline__ void insertSortedBody(const PackedTagOf<Config>* __restrict__ packedTags, size_t n,
                                                 const DeviceView<Config>& v) {
    using Ops = FilterOps<Config>;
    using Tag = typename Config::TagType;
    using P = PackedTagOf<Config>;
    constexpr P fpMask = static_cast<P>((1ULL << Config::bitsPerTag) - 1);
    uint32_t ok = 0;
    uint32_t ev = 0;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    auto account = [&](bool s, uint32_t chain) {
        ok += s;
        ev += chain;
        if (v.chainHist != nullptr && (chain != 0 || !s)) {
            atomicAdd(&v.chainHist[s ? chain : v.maxEvictions + 1], 1u);
        }
    };
    if constexpr (Ops::kSnapshot && Ops::L::bucketBytes <= 64) {
        __shared__ RunLds<Slots> runLds;
        __shared__ PendingLds<Slots> pendingLds;
        PendingInserts<Config, Slots> pending(pendingLds);
        // every lane of a wave runs the same iterations (the base is uniform
        // per workgroup), so that the run's lanes meet in the wave operations;
        // the next iteration's packed tag is read ahead
        size_t base = static_cast<size_t>(blockIdx.x) * blockDim.x;
        P packed = base + threadIdx.x < n ? packedTags[base + threadIdx.x] : P(0);
        for (; base < n; base += stride) {
            const bool active = base + threadIdx.x < n;
            const size_t nextI = base + stride + threadIdx.x;
            const P next = nextI < n ? packedTags[nextI] : P(0);
            const size_t i1 = static_cast<size_t>(packed >> Config::bitsPerTag);
            const Tag fp = static_cast<Tag>(packed & fpMask);
            const uint32_t r = primaryRunInsert<Config, Slots>(v, active, i1, fp, runLds);
            if (active && r == kRunPlaced) account(true, 0);
            pending.push(v, active && r != kRunPlaced, i1, fp, r, account);
            packed = next;
        }
        pending.finish(v, account);
    } else {
        for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
            const P packed = packedTags[i];
            uint32_t chain;
            const bool s = Ops::insertFrom(v, static_cast<size_t>(packed >> Config::bitsPerTag),
                                           static_cast<Tag>(packed & fpMask), chain);
            account(s, chain);
        }
    }
    blockCountAdd(ok, v.occupied, false);
    if (v.evictions != nullptr) blockCountAdd(ev, v.evictions, false);
}

}  // namespace detail

/// insertKernelSorted (CuckooFilter.cuh:1209-1266); `keys` is unused, as there.
/// Any workgroup of up to 1024 threads.
template <typename Config>
__global__ __launch_bounds__(kMaxBlock) void insertKernelSorted(const typename Config::KeyType* /*keys*/,
                                                                const PackedTagOf<Config>* __restrict__ packedTags,
                                                                size_t n, DeviceView<Config> v) {
    detail::insertSortedBody<Config, kMaxBlock>(packedTags, n, v);
}

/// insertKernelSorted for workgroups of at most kSortedSmallBlock threads: the
/// same work with LDS sized for them (10 KiB per workgroup instead of 39 KiB),
/// so that LDS does not bound how many workgroups a CU holds.
inline constexpr uint32_t kSortedSmallBlock = 256;
template <typename Config>
__global__ __launch_bounds__(kSortedSmallBlock) void insertKernelSortedSmall(
    const PackedTagOf<Config>* __restrict__ packedTags, size_t n, DeviceView<Config> v) {
    detail::insertSortedBody<Config, kSortedSmallBlock>(packedTags, n, v);
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
