// Error-returning launch layer shared by the C++ template API
// (CuckooFilter.hpp) and the C ABI: every operation is asynchronous on the
// given stream and returns the first HIP error it meets.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "binned.hpp"
#include "kernels.hpp"
#include "radix.hpp"

namespace cuckoo {

/// Packed tag of key i as computePackedTagsKernel computes it (CuckooFilter.cuh:1185-1207).
template <typename Config>
struct PackedTagSource {
    using P = PackedTagOf<Config>;
    using Raw = typename Config::KeyType;
    const typename Config::KeyType* keys;
    size_t numBuckets;
    __device__ __forceinline__ Raw load(size_t i) const { return keys[i]; }
    __device__ __forceinline__ Raw loadRel(size_t tileBase, uint32_t off) const { return (keys + tileBase)[off]; }
    __device__ __forceinline__ P operator()(size_t i) const { return cook(keys[i]); }
    __device__ __forceinline__ P cook(Raw key) const {
        const auto c = Config::AltBucketPolicy::getCandidateBuckets(key, numBuckets);
        return (static_cast<P>(c.i1) << Config::bitsPerTag) | static_cast<P>(c.fp);
    }
};

template <typename Config>
struct Engine {
    using Key = typename Config::KeyType;
    using View = DeviceView<Config>;
    using Packed = PackedTagOf<Config>;
    using Binned = BinnedInsert<Config>;

    static bool useBinned(size_t n, size_t numBuckets) {
        if (!Binned::applicable(numBuckets) || n == 0) return false;
        switch (detail::insertPathFromEnv()) {
            case detail::InsertPath::Direct:
                return false;
            case detail::InsertPath::Binned:
                return true;
            default:
                return n >= detail::kBinnedMinInsertKeys && n * detail::kBinnedMinBucketsPerKey >= numBuckets;
        }
    }

    /// insertMany (CuckooFilter.cuh:457-464): the bucket-binned path for
    /// batches dense enough to pay for staging table slices, one thread per
    /// key with device-scope CAS otherwise.
    static hipError_t insert(const View& v, const Key* keys, size_t n, uint32_t block, hipStream_t s,
                             Workspace& ws) {
        if (n == 0) return hipSuccess;
        if (useBinned(n, v.numBuckets)) return Binned::template run<false>(v, keys, n, block, s, ws);
        auto k = insertKernel<Config>;
        k<<<detail::gridFor(k, n, block), block, 0, s>>>(keys, n, v);
        return hipGetLastError();
    }

    /// Device workspace insert() reserves for n keys (0 on the direct path).
    static size_t insertWorkspaceBytes(size_t n, size_t numBuckets) {
        if (!useBinned(n, numBuckets)) return 0;
        return Binned::workspaceBytes(std::min(n, size_t(1) << 31), numBuckets);
    }

    static hipError_t contains(const View& v, const Key* keys, size_t n, bool* out, uint32_t block,
                               hipStream_t s) {
        if (n == 0) return hipSuccess;
        auto k = containsKernel<Config>;
        using Ops = FilterOps<Config>;
        const size_t threads = Ops::groupedLookup(v.numBuckets, block) ? n * Ops::kLookupLanes : n;
        k<<<detail::gridFor(k, threads, block), block, 0, s>>>(keys, out, n, v);
        return hipGetLastError();
    }

    static bool useBinnedDelete(size_t n, size_t numBuckets) {
        if (!Binned::applicable(numBuckets) || n == 0) return false;
        switch (detail::deletePathFromEnv()) {
            case detail::InsertPath::Direct:
                return false;
            case detail::InsertPath::Binned:
                return true;
            default:
                return n >= detail::kBinnedMinDeleteKeys && n * detail::kBinnedMinBucketsPerKey >= numBuckets;
        }
    }

    /// deleteMany (CuckooFilter.cuh:536-548): the bucket-binned path for
    /// batches dense enough to pay for staging table slices, one thread per
    /// key with device-scope CAS otherwise.
    static hipError_t remove(const View& v, const Key* keys, size_t n, bool* out, uint32_t block, hipStream_t s,
                             Workspace& ws) {
        if (n == 0) return hipSuccess;
        if (useBinnedDelete(n, v.numBuckets)) return Binned::template run<false, true>(v, keys, n, block, s, ws, out);
        auto k = deleteKernel<Config>;
        k<<<detail::gridFor(k, n, block), block, 0, s>>>(keys, out, n, v);
        return hipGetLastError();
    }

    static hipError_t packedTags(const Key* keys, size_t n, size_t nb, Packed* out, uint32_t block,
                                 hipStream_t s) {
        if (n == 0) return hipSuccess;
        auto k = computePackedTagsKernel<Config>;
        k<<<detail::gridFor(k, n, block), block, 0, s>>>(keys, out, n, nb);
        return hipGetLastError();
    }

    static hipError_t candidates(const Key* keys, size_t n, size_t nb, uint64_t* i1, uint64_t* i2,
                                 uint32_t* fp, uint32_t block, hipStream_t s) {
        if (n == 0) return hipSuccess;
        auto k = candidatesKernel<Config>;
        k<<<detail::gridFor(k, n, block), block, 0, s>>>(keys, n, nb, i1, i2, fp);
        return hipGetLastError();
    }

    /// Bits that order packed tags of a table of nb buckets: the fingerprint
    /// and the primary bucket, capped at PackedTagType's width (the
    /// reference's u32 truncation for 8-bit tags, CuckooFilter.cuh:159).
    static int packedBits(size_t nb) {
        const int hi = static_cast<int>(Config::bitsPerTag) + detail::bitsFor(nb);
        return std::min<int>(hi, static_cast<int>(sizeof(Packed) * 8));
    }

    static size_t sortWorkspaceBytes(size_t m) {
        return ((m * sizeof(Packed) + 255) / 256) * 256 + radix::scratchBytes(m);
    }

    /// The cub::DeviceRadixSort::SortKeys step of insertManySorted
    /// (CuckooFilter.cuh:485-498), in place: n packed tags of a table of nb
    /// buckets in ascending order.  Every bit a packed tag of this table can
    /// hold is sorted, so the order is the reference's full-value order.  The
    /// LSD passes ping-pong between the workspace and `data`.  n < 2^32.
    static hipError_t sortPacked(Packed* data, size_t n, size_t nb, hipStream_t s, Workspace& ws) {
        if (n < 2) return hipSuccess;
        const size_t vals = ((n * sizeof(Packed) + 255) / 256) * 256;
        hipError_t e = ws.reserve(sortWorkspaceBytes(n));
        if (e != hipSuccess) return e;
        Packed* a = static_cast<Packed*>(ws.ptr);
        void* scratch = static_cast<char*>(ws.ptr) + vals;
        Packed* sorted = data;
        e = radix::sortBits<Packed>(radix::ArraySource<Packed>{data}, n, 0, packedBits(nb), a, data, scratch, s,
                                    &sorted);
        if (e != hipSuccess) return e;
        if (sorted != data) e = hipMemcpyAsync(data, sorted, n * sizeof(Packed), hipMemcpyDeviceToDevice, s);
        return e;
    }

    /// insertKernelSorted over n packed tags (CuckooFilter.cuh:1209-1266), as
    /// the reference's presorted benchmark launches it
    /// (benchmark/cuckoo_sorted_vs_unsorted.cu:74-76).
    /// Workgroups of up to kSortedSmallBlock threads take the variant whose
    /// LDS is sized for them.
    static hipError_t insertPacked(const View& v, const Packed* packed, size_t n, uint32_t block, hipStream_t s) {
        if (n == 0) return hipSuccess;
        if (block <= kSortedSmallBlock) {
            auto k = insertKernelSortedSmall<Config>;
            k<<<detail::gridFor(k, n, block), block, 0, s>>>(packed, n, v);
        } else {
            auto k = insertKernelSorted<Config>;
            k<<<detail::gridFor(k, n, block), block, 0, s>>>(nullptr, packed, n, v);
        }
        return hipGetLastError();
    }

    /// insertManySorted (CuckooFilter.cuh:475-508): keys are radix-partitioned
    /// by primary bucket (as packed into PackedTagType) and inserted slice by
    /// slice through LDS; batches too small for that (useBinned) go one
    /// thread per key, as insertMany's do.  Tables beyond 2^32 buckets, and
    /// every table when CUCKOO_INSERT_PATH=presorted, take the reference's own
    /// route instead: packed tags, sorted ascending over all their bits, then
    /// insertKernelSorted.
    static hipError_t insertSorted(const View& v, const Key* keys, size_t n, uint32_t block, hipStream_t s,
                                   Workspace& ws) {
        if (n == 0) return hipSuccess;
        const bool presorted = detail::insertPathFromEnv() == detail::InsertPath::Presorted;
        if (!presorted && Binned::applicable(v.numBuckets)) {
            if (useBinned(n, v.numBuckets)) return Binned::template run<true>(v, keys, n, block, s, ws);
            auto k = insertKernel<Config>;
            k<<<detail::gridFor(k, n, block), block, 0, s>>>(keys, n, v);
            return hipGetLastError();
        }
        constexpr size_t kChunk = size_t(1) << 31;
        for (size_t off = 0; off < n; off += kChunk) {
            const size_t m = std::min(kChunk, n - off);
            const size_t vals = ((m * sizeof(Packed) + 255) / 256) * 256;
            hipError_t e = ws.reserve(2 * vals + radix::scratchBytes(m));
            if (e != hipSuccess) return e;
            Packed* a = static_cast<Packed*>(ws.ptr);
            Packed* b = reinterpret_cast<Packed*>(static_cast<char*>(ws.ptr) + vals);
            void* scratch = static_cast<char*>(ws.ptr) + 2 * vals;
            Packed* sorted = a;
            const PackedTagSource<Config> src{keys + off, v.numBuckets};
            e = radix::sortBits<Packed>(src, m, 0, packedBits(v.numBuckets), a, b, scratch, s, &sorted);
            if (e != hipSuccess) return e;
            if ((e = insertPacked(v, sorted, m, block, s)) != hipSuccess) return e;
        }
        return hipSuccess;
    }
};

}  // namespace cuckoo
