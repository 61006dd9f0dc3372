// Batched many-filter execution: many independent filters (one
// configuration, any capacities) and many uneven insert / lookup / delete
// batches, executed with one hipGraph launch per window instead of one
// launch + synchronisation + counter read-back per batch.
//
// Ordering: results must equal running the batches one by one through the
// plain API.  Batches on different filters commute; on one filter, lookups
// commute with each other but not with inserts or deletes.  Every batch gets
// a *level*: a lookup runs one level after the last earlier insert/delete on
// its filter, an insert/delete one level after every earlier batch on its
// filter.  A phased window (the serving loop's phased ordering) instead
// treats its batches as concurrent: each filter's inserts run as one level
// (one concurrent insert), then its lookups and deletes together in the next.
// A delete never moves a tag, so a lookup running beside deletes differs
// from one before them only for keys those deletes remove; inserts move tags
// (evictions) and so stay in a level of their own.  The levels run in order,
// each as one grid-stride kernel over
// fixed-size key chunks of all of its batches' keys (512 keys, two per thread,
// for throughput; the serving loop uses 256, one per thread, for latency).
// In a level with an insert or delete, the last workgroup to finish records
// every batch's filter occupancy, which is exactly the value the plain
// insertMany / deleteMany call would have returned; a level of lookups only
// changes no counter, so its first workgroup records them at once, and also
// those of the level before it when that level mutated: the mutating level
// then keeps no completion count.
//
// A window is one device buffer: a header, the batch descriptors, the
// level tables and per-level completion counters, uploaded with one copy
// from pinned memory.
// The kernels find everything through the header, so a captured graph
// depends only on the level count and stays valid for every later window
// with that many levels; graphs are cached per level count.  Callers whose
// windows all differ (the serving loop) launch level by level instead, each
// level with a grid sized to its own work.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "device_ops.hpp"
#include "kernels.hpp"

namespace cuckoo {

enum class BatchOpKind : uint32_t { Insert = 0, Lookup = 1, Delete = 2 };

/// One batch, as the device sees it.
struct BatchOpDesc {
    void* words;                    // filter table
    size_t numBuckets;
    unsigned long long* occupied;   // filter counters
    unsigned long long* evictions;  // may be nullptr
    unsigned int* chainHist;        // may be nullptr
    const void* keys;
    bool* out;                      // lookup: results; delete: optional
    uint64_t n;
    uint32_t kind;
    uint32_t maxEvictions;
    uint32_t evictionPolicy;
    uint32_t pad;
};

inline constexpr uint32_t kBatchChunk = 512;        // default keys per workgroup step: two per thread
inline constexpr uint32_t kBatchChunkLatency = 256;  // one key per thread (serving windows)
inline constexpr int kBatchBlock = 256;
inline constexpr uint32_t kMaxGraphLevels = 64;  // deeper windows launch level by level

/// Start of a window buffer; offsets are bytes from the start of the buffer.
struct alignas(64) WindowHeader {
    uint32_t numOps;
    uint32_t numLevels;
    uint32_t pad[2];
    uint64_t opsOff;       // BatchOpDesc[numOps]
    uint64_t levelOpsOff;  // uint32_t[numOps]: op indices grouped by level, window order inside a level
    uint64_t chunkCumOff;  // uint32_t[numOps + 1]: chunks before each levelOps position
    uint64_t opStartOff;   // uint32_t[numLevels + 1]: first levelOps position of each level
    uint64_t doneOff;      // uint32_t[numLevels]: workgroups finished per level (uploaded as zero)
    uint64_t levelModeOff;  // uint32_t[numLevels]: a LevelMode
};

namespace detail {
template <typename T>
__device__ __forceinline__ T* windowPart(unsigned char* w, uint64_t off) {
    return reinterpret_cast<T*>(w + off);
}
}  // namespace detail

/// Who records a level's occupancies.
enum LevelMode : uint32_t {
    kLookupsOnly = 0,     // workgroup 0, as the level starts (and the previous level's, if deferred)
    kMutatesRecords = 1,  // the level mutates; its last workgroup records them
    kMutatesDeferred = 2  // the level mutates; the next level (lookups only) records them
};

/// A level's place in the window, passed by a level-by-level launch so that
/// the kernel starts on its keys without first reading the window's tables
/// (a chain of dependent loads of a buffer just copied in).  Graph launches
/// pass valid = 0 and the kernel reads the same values from the window.
struct LevelArgs {
    const BatchOpDesc* ops;
    const uint32_t* levelOps;
    const uint32_t* chunkCum;
    uint32_t* done;
    uint32_t p0, p1;    // the level's levelOps positions [p0, p1)
    uint32_t g0, g1;    // its chunks [g0, g1)
    uint32_t op0;       // levelOps[p0]
    uint32_t mode;      // LevelMode
    uint32_t r0;        // first levelOps position workgroup 0 records (p0, or the previous level's start)
    uint32_t valid;
};

/// One level: a grid-stride loop over the level's chunks (each workgroup
/// finds its batch by binary search over the chunk counts), then the last
/// workgroup to finish records results[op] = the op's filter occupancy for
/// every op of the level.  A level of lookups only leaves every counter as
/// the previous level (an earlier kernel on the stream) left it, so
/// workgroup 0 records them first, together with the previous level's when
/// that level deferred them, and no completion count is kept.
template <typename Config, uint32_t Chunk>
__global__ __launch_bounds__(kBatchBlock) void batchLevelKernel(unsigned char* __restrict__ window, uint32_t level,
                                                                unsigned long long* __restrict__ results,
                                                                LevelArgs la) {
    using Ops = FilterOps<Config>;
    using Key = typename Config::KeyType;
    const BatchOpDesc* ops;
    const uint32_t* levelOps;
    const uint32_t* chunkCum;
    uint32_t* done;
    uint32_t p0, p1, g0, g1, mode, r0;
    if (la.valid) {
        ops = la.ops;
        levelOps = la.levelOps;
        chunkCum = la.chunkCum;
        done = la.done;
        p0 = la.p0;
        p1 = la.p1;
        g0 = la.g0;
        g1 = la.g1;
        mode = la.mode;
        r0 = la.r0;
    } else {
        const auto* hdr = reinterpret_cast<const WindowHeader*>(window);
        ops = detail::windowPart<const BatchOpDesc>(window, hdr->opsOff);
        levelOps = detail::windowPart<const uint32_t>(window, hdr->levelOpsOff);
        chunkCum = detail::windowPart<const uint32_t>(window, hdr->chunkCumOff);
        done = detail::windowPart<uint32_t>(window, hdr->doneOff);
        const auto* opStart = detail::windowPart<const uint32_t>(window, hdr->opStartOff);
        p0 = opStart[level];
        p1 = opStart[level + 1];
        g0 = chunkCum[p0];
        g1 = chunkCum[p1];
        const auto* levelMode = detail::windowPart<const uint32_t>(window, hdr->levelModeOff);
        mode = levelMode[level];
        r0 = level > 0 && levelMode[level - 1] == kMutatesDeferred ? opStart[level - 1] : p0;
    }
    // a level of one batch (most levels of a server's windows) needs no search
    const bool single = p1 - p0 == 1;
    const bool op0Known = single && la.valid;
    if (mode == kLookupsOnly && blockIdx.x == 0) {
        for (uint32_t j = r0 + threadIdx.x; j < p1; j += kBatchBlock) {
            const uint32_t op = op0Known && j == p0 ? la.op0 : levelOps[j];
            results[op] = __hip_atomic_load(ops[op].occupied, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    for (uint32_t g = g0 + blockIdx.x; g < g1; g += gridDim.x) {
        // last position j in [p0, p1) with chunkCum[j] <= g (skips empty batches)
        uint32_t lo = p0;
        uint32_t base = g0;
        if (!single) {
            uint32_t hi = p1 - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (chunkCum[mid] <= g) {
                    lo = mid;
                } else {
                    hi = mid - 1;
                }
            }
            base = chunkCum[lo];
        }
        const BatchOpDesc& o = ops[op0Known ? la.op0 : levelOps[lo]];
        const uint64_t offset = static_cast<uint64_t>(g - base) * Chunk;
        const uint32_t count = static_cast<uint32_t>(o.n - offset < Chunk ? o.n - offset : Chunk);
        DeviceView<Config> v;
        v.words = static_cast<typename DeviceView<Config>::Word*>(o.words);
        v.numBuckets = o.numBuckets;
        v.occupied = o.occupied;
        v.evictions = o.evictions;
        v.chainHist = o.chainHist;
        v.maxEvictions = o.maxEvictions;
        v.evictionPolicy = o.evictionPolicy;
        const Key* keys = static_cast<const Key*>(o.keys) + offset;
        uint32_t ok = 0;
        uint32_t ev = 0;
        if (o.kind == static_cast<uint32_t>(BatchOpKind::Lookup)) {
            for (uint32_t i = threadIdx.x; i < count; i += kBatchBlock) o.out[offset + i] = Ops::contains(v, keys[i]);
        } else if (o.kind == static_cast<uint32_t>(BatchOpKind::Insert)) {
            for (uint32_t i = threadIdx.x; i < count; i += kBatchBlock) {
                uint32_t chain;
                const bool s = Ops::insert(v, keys[i], chain);
                ok += s;
                ev += chain;
                if (v.chainHist != nullptr && (chain != 0 || !s)) {
                    atomicAdd(&v.chainHist[s ? chain : v.maxEvictions + 1], 1u);
                }
            }
        } else {
            for (uint32_t i = threadIdx.x; i < count; i += kBatchBlock) {
                const bool s = Ops::remove(v, keys[i]);
                ok += s;
                if (o.out != nullptr) o.out[offset + i] = s;
            }
        }
        // uniform across the block: every thread takes the same branch per chunk
        if (o.kind != static_cast<uint32_t>(BatchOpKind::Lookup)) {
            detail::blockCountAdd(ok, v.occupied, o.kind == static_cast<uint32_t>(BatchOpKind::Delete));
            if (o.kind == static_cast<uint32_t>(BatchOpKind::Insert) && v.evictions != nullptr) {
                detail::blockCountAdd(ev, v.evictions, false);
            }
        }
    }
    // The last workgroup of the level snapshots the counters: every other
    // workgroup's counter updates precede its increment of done[level].  Only
    // workgroups that own a chunk (block b < the level's chunk count) take
    // part; with no chunk at all, workgroup 0 snapshots alone.
    if (mode != kMutatesRecords) return;
    const uint32_t levelChunks = g1 - g0;
    const uint32_t parts = levelChunks < gridDim.x ? levelChunks : gridDim.x;
    if (parts == 0 ? blockIdx.x != 0 : blockIdx.x >= parts) return;
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = parts == 0 || atomicAdd(&done[level], 1u) == parts - 1;
    }
    __syncthreads();
    if (last) {
        __threadfence();
        for (uint32_t j = p0 + threadIdx.x; j < p1; j += kBatchBlock) {
            const uint32_t op = op0Known ? la.op0 : levelOps[j];
            results[op] = __hip_atomic_load(ops[op].occupied, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

/// Host side: plans windows into levels and launches them.
///
/// launch() is asynchronous; its occupancies (results()) are valid once the
/// stream has passed the launch.  The executor's staging buffers are reused
/// by the next launch(), so a caller must not call launch() again before the
/// previous launch has completed on the device (execute() synchronises;
/// callers that keep windows in flight use one executor per window slot).
template <typename Config>
class BatchExecutor {
   public:
    BatchExecutor() = default;
    BatchExecutor(const BatchExecutor&) = delete;
    BatchExecutor& operator=(const BatchExecutor&) = delete;
    ~BatchExecutor() { release(); }

    /// Executes `ops` (host descriptors) on `stream` and synchronises.
    /// occupiedOut (optional, host) receives each op's post-op occupancy.
    hipError_t execute(const BatchOpDesc* ops, size_t count, hipStream_t stream, uint64_t* occupiedOut) {
        if (count == 0) return hipSuccess;
        hipError_t e = launch(ops, count, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return e;
        if (occupiedOut != nullptr) {
            for (size_t i = 0; i < count; ++i) occupiedOut[i] = results()[i];
        }
        return hipSuccess;
    }

    /// Plans `ops`, uploads the window and launches it on `stream` without
    /// waiting; results()[i] holds op i's post-op occupancy once the stream
    /// has passed this launch.  With useGraph the levels run as the graph
    /// cached for this level count (cheapest when windows of one shape
    /// repeat); otherwise level by level, each with a grid no larger than its
    /// chunk count (cheapest when every window differs, as in a server).
    /// chunkKeys is the keys per workgroup step: kBatchChunk (two per
    /// thread) or kBatchChunkLatency (one per thread); graphs are kept for
    /// kBatchChunk only, other windows run level by level.  phased plans the
    /// window in two phases per filter (its inserts, then its lookups and
    /// deletes) instead of in window order; every op then reports its
    /// filter's occupancy after its phase.  beforeLast (optional, level-by-level launches) is
    /// recorded on the stream just before the window's last level.
    hipError_t launch(const BatchOpDesc* ops, size_t count, hipStream_t stream, bool useGraph = true,
                      uint32_t chunkKeys = kBatchChunk, bool phased = false, hipEvent_t beforeLast = nullptr) {
        if (count == 0) return hipSuccess;
        if (chunkKeys != kBatchChunk && chunkKeys != kBatchChunkLatency) return hipErrorInvalidValue;
        const uint32_t numLevels = plan(ops, count, chunkKeys, phased);
        if (stageError_ != hipSuccess) return stageError_;
        hipError_t e;
        bool moved = false;
        if ((e = growDevice(&dWin_, &capWin_, staged_, &moved)) != hipSuccess) return e;
        if ((e = growDevice(&dResults_, &capResults_, count * sizeof(unsigned long long), &moved)) != hipSuccess)
            return e;
        if ((e = growPinned(&hResults_, &capHResults_, count * sizeof(unsigned long long))) != hipSuccess) return e;
        if (moved) dropGraphs();
        if ((e = hipMemcpyAsync(dWin_, hStage_, staged_, hipMemcpyHostToDevice, stream)) != hipSuccess) return e;
        auto* win = static_cast<unsigned char*>(dWin_);
        auto* results = static_cast<unsigned long long*>(dResults_);
        if (useGraph && chunkKeys == kBatchChunk && numLevels <= kMaxGraphLevels) {
            if (graphs_[numLevels] == nullptr && (e = capture(numLevels)) != hipSuccess) return e;
            if ((e = hipGraphLaunch(graphs_[numLevels], stream)) != hipSuccess) return e;
        } else {
            // level by level, each with a grid no larger than its chunk count
            const uint32_t grid = levelGrid(phased);
            LevelArgs la{};
            la.ops = reinterpret_cast<const BatchOpDesc*>(win + hdr_.opsOff);
            la.levelOps = reinterpret_cast<const uint32_t*>(win + hdr_.levelOpsOff);
            la.chunkCum = reinterpret_cast<const uint32_t*>(win + hdr_.chunkCumOff);
            la.done = reinterpret_cast<uint32_t*>(win + hdr_.doneOff);
            la.valid = 1;
            for (uint32_t l = 0; l < numLevels; ++l) {
                la.p0 = opStart_[l];
                la.p1 = opStart_[l + 1];
                la.g0 = chunkCum_[la.p0];
                la.g1 = chunkCum_[la.p1];
                la.op0 = levelOps_[la.p0];
                la.mode = levelMode_[l];
                la.r0 = l > 0 && levelMode_[l - 1] == kMutatesDeferred ? opStart_[l - 1] : la.p0;
                const uint32_t chunks = la.g1 - la.g0;
                const uint32_t g = chunks == 0 ? 1u : (chunks < grid ? chunks : grid);
                if (beforeLast != nullptr && l + 1 == numLevels &&
                    (e = hipEventRecord(beforeLast, stream)) != hipSuccess)
                    return e;
                if (chunkKeys == kBatchChunk) {
                    batchLevelKernel<Config, kBatchChunk><<<g, kBatchBlock, 0, stream>>>(win, l, results, la);
                } else {
                    batchLevelKernel<Config, kBatchChunkLatency><<<g, kBatchBlock, 0, stream>>>(win, l, results, la);
                }
            }
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        lastLevels_ = numLevels;
        return hipMemcpyAsync(hResults_, dResults_, count * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream);
    }

    /// Post-op occupancies of the last launch (pinned host memory).
    const unsigned long long* results() const { return static_cast<const unsigned long long*>(hResults_); }
    uint32_t levelsOfLastLaunch() const { return lastLevels_; }

    void release() {
        dropGraphs();
        for (void** p : {&dWin_, &dResults_}) {
            if (*p != nullptr) (void)hipFree(*p);
            *p = nullptr;
        }
        for (void** p : {&hStage_, &hResults_}) {
            if (*p != nullptr) (void)hipHostFree(*p);
            *p = nullptr;
        }
        capWin_ = capResults_ = capStage_ = capHResults_ = 0;
        if (captureStream_ != nullptr) (void)hipStreamDestroy(captureStream_);
        captureStream_ = nullptr;
    }

   private:
    static constexpr size_t align64(size_t x) { return (x + 63) & ~static_cast<size_t>(63); }

    /// Writes the window (header, descriptors, chunks, level tables) into the
    /// pinned staging buffer; returns the level count.
    uint32_t plan(const BatchOpDesc* ops, size_t count, uint32_t chunk, bool phased) {
        seen_.clear();
        level_.resize(count);
        uint32_t numLevels = 0;
        if (phased) {
            // level 1 for a lookup or delete on a filter that has inserts in
            // the window, else level 0
            hasInserts_.clear();
            filterOf_.resize(count);
            for (size_t i = 0; i < count; ++i) {
                auto [it, fresh] = seen_.try_emplace(ops[i].occupied, static_cast<uint32_t>(hasInserts_.size()));
                if (fresh) hasInserts_.push_back(0);
                filterOf_[i] = it->second;
                if (ops[i].kind == static_cast<uint32_t>(BatchOpKind::Insert)) hasInserts_[it->second] = 1;
            }
            for (size_t i = 0; i < count; ++i) {
                const bool insert = ops[i].kind == static_cast<uint32_t>(BatchOpKind::Insert);
                level_[i] = !insert && hasInserts_[filterOf_[i]] != 0 ? 1u : 0u;
                numLevels = std::max(numLevels, level_[i] + 1);
            }
        }
        for (size_t i = 0; i < count && !phased; ++i) {
            auto [it, fresh] = seen_.try_emplace(ops[i].occupied, static_cast<uint32_t>(state_.size()));
            if (fresh) state_.push_back(FilterLevelsPod{-1, -1});
            FilterLevelsPod& st = state_[it->second];
            int32_t l;
            if (ops[i].kind == static_cast<uint32_t>(BatchOpKind::Lookup)) {
                l = st.lastMutation + 1;
                st.maxLevel = std::max(st.maxLevel, l);
            } else {
                l = st.maxLevel + 1;
                st.lastMutation = l;
                st.maxLevel = l;
            }
            level_[i] = static_cast<uint32_t>(l);
            numLevels = std::max(numLevels, static_cast<uint32_t>(l) + 1);
        }
        state_.clear();
        // counting sort of ops by level, stable in window order
        opStart_.assign(numLevels + 1, 0);
        for (size_t i = 0; i < count; ++i) ++opStart_[level_[i] + 1];
        for (uint32_t l = 0; l < numLevels; ++l) opStart_[l + 1] += opStart_[l];
        levelOps_.resize(count);
        {
            std::vector<uint32_t>& cursor = cursor_;
            cursor.assign(opStart_.begin(), opStart_.end() - 1);
            for (size_t i = 0; i < count; ++i) levelOps_[cursor[level_[i]]++] = static_cast<uint32_t>(i);
        }
        // chunk counts in level order
        chunkCum_.resize(count + 1);
        chunkCum_[0] = 0;
        uint64_t chunks = 0;
        for (size_t j = 0; j < count; ++j) {
            chunks += (ops[levelOps_[j]].n + chunk - 1) / chunk;
            if (chunks > UINT32_MAX) {
                stageError_ = hipErrorInvalidValue;  // more chunks than the 32-bit tables index
                staged_ = 0;
                return 0;
            }
            chunkCum_[j + 1] = static_cast<uint32_t>(chunks);
        }
        levelMode_.assign(numLevels, kLookupsOnly);
        for (size_t i = 0; i < count; ++i) {
            if (ops[i].kind != static_cast<uint32_t>(BatchOpKind::Lookup)) levelMode_[level_[i]] = kMutatesRecords;
        }
        for (uint32_t l = 0; l + 1 < numLevels; ++l) {
            if (levelMode_[l] == kMutatesRecords && levelMode_[l + 1] == kLookupsOnly) levelMode_[l] = kMutatesDeferred;
        }

        WindowHeader& h = hdr_;
        h = WindowHeader{};
        h.numOps = static_cast<uint32_t>(count);
        h.numLevels = numLevels;
        size_t off = align64(sizeof(WindowHeader));
        h.opsOff = off;
        off = align64(off + count * sizeof(BatchOpDesc));
        h.levelOpsOff = off;
        off = align64(off + count * sizeof(uint32_t));
        h.chunkCumOff = off;
        off = align64(off + (count + 1) * sizeof(uint32_t));
        h.opStartOff = off;
        off = align64(off + (numLevels + 1) * sizeof(uint32_t));
        h.doneOff = off;
        off = align64(off + numLevels * sizeof(uint32_t));
        h.levelModeOff = off;
        off = align64(off + numLevels * sizeof(uint32_t));
        staged_ = off;
        stageError_ = growPinned(&hStage_, &capStage_, staged_);
        if (stageError_ != hipSuccess) {
            staged_ = 0;
            return 0;
        }
        auto* b = static_cast<unsigned char*>(hStage_);
        std::memcpy(b, &h, sizeof(h));
        std::memcpy(b + h.opsOff, ops, count * sizeof(BatchOpDesc));
        std::memcpy(b + h.levelOpsOff, levelOps_.data(), count * sizeof(uint32_t));
        std::memcpy(b + h.chunkCumOff, chunkCum_.data(), (count + 1) * sizeof(uint32_t));
        std::memcpy(b + h.opStartOff, opStart_.data(), (numLevels + 1) * sizeof(uint32_t));
        std::memset(b + h.doneOff, 0, numLevels * sizeof(uint32_t));
        std::memcpy(b + h.levelModeOff, levelMode_.data(), numLevels * sizeof(uint32_t));
        return numLevels;
    }

    static hipError_t growDevice(void** p, size_t* cap, size_t want, bool* moved) {
        want = std::max<size_t>(want, 4096);
        if (want <= *cap) return hipSuccess;
        want = want + want / 2;
        if (*p != nullptr) {
            const hipError_t e = hipFree(*p);
            *p = nullptr;
            *cap = 0;
            if (e != hipSuccess) return e;
        }
        const hipError_t e = hipMalloc(p, want);
        if (e == hipSuccess) {
            *cap = want;
            *moved = true;
        }
        return e;
    }

    static hipError_t growPinned(void** p, size_t* cap, size_t want) {
        want = std::max<size_t>(want, 4096);
        if (want <= *cap) return hipSuccess;
        want = want + want / 2;
        if (*p != nullptr) {
            const hipError_t e = hipHostFree(*p);
            *p = nullptr;
            *cap = 0;
            if (e != hipSuccess) return e;
        }
        const hipError_t e = hipHostMalloc(p, want, hipHostMallocDefault);
        if (e == hipSuccess) *cap = want;
        return e;
    }

    /// Level grid: the resident workgroups of the two-keys-per-thread kernel
    /// (graph windows, and the serving loop's arrival-ordered windows: 6 per
    /// CU), and three workgroups per CU for phased windows.  Measured on the
    /// serving workload (2^30-slot k-mer filter + 64 tenants), phased windows
    /// of one key per thread: 512 / 768 / 1024 / 1536 / 2048 workgroups gave
    /// 15.3 / 15.7 / 15.1 / 13.3 / 11.9 Gkeys/s saturated at 2^22-key windows
    /// (fewer device-scope CAS in flight queue less at the memory side); of
    /// two keys per thread, 3 / 4 / 6 per CU gave p99 0.73 / 0.83 / 1.07 ms
    /// at 16.6 Gkeys/s offered and 23.4 / 24.1 / 23.4 saturated.
    /// Arrival-ordered windows hold many small levels, whose latency rises
    /// when fewer workgroups share a level (p99 at half load 0.25-0.39 ms at
    /// 1536, 0.75 ms at 768); graph windows of the batched workload ran
    /// fastest at full residency (23.8 Gkeys/s at 1536 against 20.0 at 768).
    static uint32_t levelGrid(bool phased) {
        constexpr size_t kMany = static_cast<size_t>(1) << 30;
        const uint32_t resident = detail::gridFor(batchLevelKernel<Config, kBatchChunk>, kMany, kBatchBlock);
        if (!phased) return resident;
        int dev = 0;
        (void)hipGetDevice(&dev);
        return std::min(resident, 3u * detail::computeUnits(dev));
    }

    void dropGraphs() {
        for (auto& g : graphs_) {
            if (g != nullptr) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }

    hipError_t capture(uint32_t numLevels) {
        hipError_t e;
        if (captureStream_ == nullptr &&
            (e = hipStreamCreateWithFlags(&captureStream_, hipStreamNonBlocking)) != hipSuccess)
            return e;
        auto* win = static_cast<unsigned char*>(dWin_);
        auto* results = static_cast<unsigned long long*>(dResults_);
        const uint32_t grid = levelGrid(false);
        if ((e = hipStreamBeginCapture(captureStream_, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
        for (uint32_t l = 0; l < numLevels; ++l) {
            batchLevelKernel<Config, kBatchChunk><<<grid, kBatchBlock, 0, captureStream_>>>(win, l, results,
                                                                                           LevelArgs{});
        }
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(captureStream_, &graph);
        if (e != hipSuccess) return e;
        e = hipGraphInstantiate(&graphs_[numLevels], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) graphs_[numLevels] = nullptr;
        return e;
    }

    struct FilterLevelsPod {
        int32_t lastMutation;
        int32_t maxLevel;
    };

    // device: the window buffer and the occupancy results
    void* dWin_ = nullptr;
    void* dResults_ = nullptr;
    size_t capWin_ = 0, capResults_ = 0;
    // pinned host: the staged window and the results read back
    void* hStage_ = nullptr;
    void* hResults_ = nullptr;
    size_t capStage_ = 0, capHResults_ = 0, staged_ = 0;
    hipError_t stageError_ = hipSuccess;
    // planning scratch (host)
    std::unordered_map<const void*, uint32_t> seen_;
    std::vector<FilterLevelsPod> state_;
    std::vector<uint32_t> level_, opStart_, levelOps_, chunkCum_, cursor_, levelMode_;
    std::vector<uint32_t> hasInserts_, filterOf_;  // phased planning: per filter, whether it has inserts; each op's filter
    WindowHeader hdr_{};  // the last planned window's header (its offsets into the device window)
    // graphs by level count
    hipStream_t captureStream_ = nullptr;
    hipGraphExec_t graphs_[kMaxGraphLevels + 1] = {};
    uint32_t lastLevels_ = 0;
};

}  // namespace cuckoo
