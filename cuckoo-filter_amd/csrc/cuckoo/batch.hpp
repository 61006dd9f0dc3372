// Batched many-filter execution: many independent filters (one
// configuration, any capacities) and many uneven insert / lookup / delete
// batches, executed with one hipGraph launch per window instead of one
// launch + synchronisation + counter read-back per batch.
//
// Ordering: results must equal running the batches one by one through the
// plain API.  Batches on different filters commute; batches on the same
// filter do not.  A batch's *level* is the number of earlier batches in the
// window that target the same filter, and the levels run in order, each as
// one grid-stride kernel over fixed-size chunks of all of its batches' keys
// (the batches of one level all target different filters).  After each level
// a small kernel records every batch's filter occupancy, which is exactly the
// value the plain insertMany / deleteMany call would have returned.
//
// All descriptors live in device buffers that the kernels index by level, so
// the captured graph stays valid from one window to the next as long as the
// level count and buffer capacities do not change.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
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

/// A run of at most kBatchChunk keys of one batch.
struct BatchChunk {
    uint32_t op;
    uint32_t count;
    uint64_t offset;
};

inline constexpr uint32_t kBatchChunk = 8192;
inline constexpr int kBatchBlock = 256;

template <typename Config>
__global__ __launch_bounds__(kBatchBlock) void batchLevelKernel(const BatchOpDesc* __restrict__ ops,
                                                                const BatchChunk* __restrict__ chunks,
                                                                const uint32_t* __restrict__ levelChunkStart,
                                                                uint32_t level) {
    using Ops = FilterOps<Config>;
    using Key = typename Config::KeyType;
    const uint32_t c0 = levelChunkStart[level];
    const uint32_t c1 = levelChunkStart[level + 1];
    for (uint32_t c = c0 + blockIdx.x; c < c1; c += gridDim.x) {
        const BatchChunk ch = chunks[c];
        const BatchOpDesc& o = ops[ch.op];
        DeviceView<Config> v;
        v.words = static_cast<typename DeviceView<Config>::Word*>(o.words);
        v.numBuckets = o.numBuckets;
        v.occupied = o.occupied;
        v.evictions = o.evictions;
        v.chainHist = o.chainHist;
        v.maxEvictions = o.maxEvictions;
        v.evictionPolicy = o.evictionPolicy;
        const Key* keys = static_cast<const Key*>(o.keys) + ch.offset;
        uint32_t ok = 0;
        uint32_t ev = 0;
        if (o.kind == static_cast<uint32_t>(BatchOpKind::Lookup)) {
            for (uint32_t i = threadIdx.x; i < ch.count; i += kBatchBlock) o.out[ch.offset + i] = Ops::contains(v, keys[i]);
        } else if (o.kind == static_cast<uint32_t>(BatchOpKind::Insert)) {
            for (uint32_t i = threadIdx.x; i < ch.count; i += kBatchBlock) {
                uint32_t chain;
                const bool s = Ops::insert(v, keys[i], chain);
                ok += s;
                ev += chain;
                if (v.chainHist != nullptr && (chain != 0 || !s)) {
                    atomicAdd(&v.chainHist[s ? chain : v.maxEvictions + 1], 1u);
                }
            }
        } else {
            for (uint32_t i = threadIdx.x; i < ch.count; i += kBatchBlock) {
                const bool s = Ops::remove(v, keys[i]);
                ok += s;
                if (o.out != nullptr) o.out[ch.offset + i] = s;
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
}

/// results[op] = occupancy of the op's filter, for the ops of one level.
template <typename T = uint32_t>
__global__ __launch_bounds__(256) void batchSnapshotKernel(const BatchOpDesc* __restrict__ ops,
                                                           const T* __restrict__ levelOps,
                                                           const uint32_t* __restrict__ levelOpStart,
                                                           uint32_t level, unsigned long long* __restrict__ results) {
    const uint32_t a = levelOpStart[level];
    const uint32_t b = levelOpStart[level + 1];
    for (uint32_t i = a + blockIdx.x * blockDim.x + threadIdx.x; i < b; i += gridDim.x * blockDim.x) {
        const uint32_t op = levelOps[i];
        results[op] = *ops[op].occupied;
    }
}

/// Host side: plans a window into levels and chunks, keeps the device
/// buffers and the captured graph between windows.
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
        // --- plan: levels per filter, chunks per level ---------------------
        std::unordered_map<const void*, uint32_t> seen;
        std::vector<uint32_t> level(count);
        uint32_t numLevels = 0;
        for (size_t i = 0; i < count; ++i) {
            uint32_t& k = seen[ops[i].occupied];
            level[i] = k++;
            numLevels = std::max(numLevels, k);
        }
        std::vector<std::vector<uint32_t>> byLevel(numLevels);
        for (size_t i = 0; i < count; ++i) byLevel[level[i]].push_back(static_cast<uint32_t>(i));
        hChunks_.clear();
        hLevelOps_.clear();
        std::vector<uint32_t> chunkStart(numLevels + 1), opStart(numLevels + 1);
        for (uint32_t l = 0; l < numLevels; ++l) {
            chunkStart[l] = static_cast<uint32_t>(hChunks_.size());
            opStart[l] = static_cast<uint32_t>(hLevelOps_.size());
            for (uint32_t op : byLevel[l]) {
                hLevelOps_.push_back(op);
                for (uint64_t off = 0; off < ops[op].n; off += kBatchChunk) {
                    const uint64_t m = std::min<uint64_t>(kBatchChunk, ops[op].n - off);
                    hChunks_.push_back(BatchChunk{op, static_cast<uint32_t>(m), off});
                }
            }
        }
        chunkStart[numLevels] = static_cast<uint32_t>(hChunks_.size());
        opStart[numLevels] = static_cast<uint32_t>(hLevelOps_.size());

        // --- device buffers (grown only) ------------------------------------
        hipError_t e;
        bool moved = false;
        if ((e = grow(&dOps_, &capOps_, count * sizeof(BatchOpDesc), &moved)) != hipSuccess) return e;
        if ((e = grow(&dChunks_, &capChunks_, std::max<size_t>(1, hChunks_.size()) * sizeof(BatchChunk), &moved)) !=
            hipSuccess)
            return e;
        if ((e = grow(&dLevelOps_, &capLevelOps_, hLevelOps_.size() * sizeof(uint32_t), &moved)) != hipSuccess) return e;
        if ((e = grow(&dStarts_, &capStarts_, 2 * (numLevels + 1) * sizeof(uint32_t), &moved)) != hipSuccess) return e;
        if ((e = grow(&dResults_, &capResults_, count * sizeof(unsigned long long), &moved)) != hipSuccess) return e;
        uint32_t* dChunkStart = static_cast<uint32_t*>(dStarts_);
        uint32_t* dOpStart = dChunkStart + (numLevels + 1);

        // one host->device upload of everything the graph reads
        if ((e = hipMemcpyAsync(dOps_, ops, count * sizeof(BatchOpDesc), hipMemcpyHostToDevice, stream)) != hipSuccess ||
            (e = hipMemcpyAsync(dChunks_, hChunks_.data(), hChunks_.size() * sizeof(BatchChunk),
                                hipMemcpyHostToDevice, stream)) != hipSuccess ||
            (e = hipMemcpyAsync(dLevelOps_, hLevelOps_.data(), hLevelOps_.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, stream)) != hipSuccess ||
            (e = hipMemcpyAsync(dChunkStart, chunkStart.data(), chunkStart.size() * sizeof(uint32_t),
                                hipMemcpyHostToDevice, stream)) != hipSuccess ||
            (e = hipMemcpyAsync(dOpStart, opStart.data(), opStart.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                stream)) != hipSuccess)
            return e;

        // --- the graph: numLevels x (level kernel, occupancy snapshot) -----
        if (moved || graphExec_ == nullptr || graphLevels_ != numLevels) {
            if ((e = capture(numLevels, dChunkStart, dOpStart)) != hipSuccess) return e;
        }
        if ((e = hipGraphLaunch(graphExec_, stream)) != hipSuccess) return e;
        if (occupiedOut != nullptr) {
            results_.resize(count);
            if ((e = hipMemcpyAsync(results_.data(), dResults_, count * sizeof(unsigned long long),
                                    hipMemcpyDeviceToHost, stream)) != hipSuccess)
                return e;
        }
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
        if (occupiedOut != nullptr) {
            for (size_t i = 0; i < count; ++i) occupiedOut[i] = results_[i];
        }
        return hipSuccess;
    }

    uint32_t levelsOfLastGraph() const { return graphLevels_; }

    void release() {
        if (graphExec_ != nullptr) (void)hipGraphExecDestroy(graphExec_);
        graphExec_ = nullptr;
        for (void** p : {&dOps_, &dChunks_, &dLevelOps_, &dStarts_, &dResults_}) {
            if (*p != nullptr) (void)hipFree(*p);
            *p = nullptr;
        }
        capOps_ = capChunks_ = capLevelOps_ = capStarts_ = capResults_ = 0;
        if (captureStream_ != nullptr) (void)hipStreamDestroy(captureStream_);
        captureStream_ = nullptr;
    }

   private:
    static hipError_t grow(void** p, size_t* cap, size_t want, bool* moved) {
        want = std::max<size_t>(want, 256);
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

    hipError_t capture(uint32_t numLevels, const uint32_t* dChunkStart, const uint32_t* dOpStart) {
        hipError_t e;
        if (graphExec_ != nullptr) {
            (void)hipGraphExecDestroy(graphExec_);
            graphExec_ = nullptr;
        }
        if (captureStream_ == nullptr && (e = hipStreamCreateWithFlags(&captureStream_, hipStreamNonBlocking)) != hipSuccess)
            return e;
        auto k = batchLevelKernel<Config>;
        const uint32_t grid = detail::gridFor(k, static_cast<size_t>(1) << 30, kBatchBlock);
        const auto* ops = static_cast<const BatchOpDesc*>(dOps_);
        const auto* chunks = static_cast<const BatchChunk*>(dChunks_);
        const auto* levelOps = static_cast<const uint32_t*>(dLevelOps_);
        auto* results = static_cast<unsigned long long*>(dResults_);
        if ((e = hipStreamBeginCapture(captureStream_, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
        for (uint32_t l = 0; l < numLevels; ++l) {
            k<<<grid, kBatchBlock, 0, captureStream_>>>(ops, chunks, dChunkStart, l);
            batchSnapshotKernel<uint32_t><<<4, 256, 0, captureStream_>>>(ops, levelOps, dOpStart, l, results);
        }
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(captureStream_, &graph);
        if (e != hipSuccess) return e;
        e = hipGraphInstantiate(&graphExec_, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            graphExec_ = nullptr;
            return e;
        }
        graphLevels_ = numLevels;
        return hipSuccess;
    }

    void* dOps_ = nullptr;
    void* dChunks_ = nullptr;
    void* dLevelOps_ = nullptr;
    void* dStarts_ = nullptr;
    void* dResults_ = nullptr;
    size_t capOps_ = 0, capChunks_ = 0, capLevelOps_ = 0, capStarts_ = 0, capResults_ = 0;
    std::vector<BatchChunk> hChunks_;
    std::vector<uint32_t> hLevelOps_;
    std::vector<unsigned long long> results_;
    hipStream_t captureStream_ = nullptr;
    hipGraphExec_t graphExec_ = nullptr;
    uint32_t graphLevels_ = 0;
};

}  // namespace cuckoo
