// LSD radix sort over a bit range of 32/64-bit values, written for wave64.
//
// Used by insertManySorted to order packed tags ((primary bucket << bits) |
// fp, CuckooFilter.cuh:1185-1207) by primary bucket.  Only the bucket-index
// bits are sorted (ceil(log2 numBuckets) bits, 8 per pass): ordering inside a
// bucket has no effect on what a concurrent insert may produce, so the fp
// bits that a full 64-bit sort (CuckooFilter.cuh:485-498) would also order
// are skipped -- 3 passes instead of 8 for a 2^24-bucket table.
//
// Each pass is count -> scan -> scatter.  The scatter is stable: a tile of
// kTile values is split into 64-value rows, each wave owns a contiguous run of
// rows, and a value's rank inside its row comes from eight ballots (one per
// digit bit) -- no LDS atomics, no per-lane sort.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace cuckoo::radix {

inline constexpr int kDigitBits = 8;
inline constexpr int kRadix = 1 << kDigitBits;
inline constexpr int kBlock = 512;
inline constexpr int kWaves = kBlock / 64;
inline constexpr int kItems = 16;
inline constexpr int kTile = kBlock * kItems;      // 8192 values per tile
inline constexpr int kRowsPerWave = kTile / 64 / kWaves;
inline constexpr int kScanBlock = 1024;
inline constexpr int kScanItems = 4;
inline constexpr int kScanChunk = kScanBlock * kScanItems;

static_assert(kTile % (64 * kWaves) == 0, "tile must split evenly into wave rows");

__device__ __forceinline__ uint32_t waveInclusiveScan(uint32_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

/// Exclusive block scan of one value per thread; returns the block total.
template <int BLOCK>
__device__ __forceinline__ uint32_t blockExclusiveScan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t waveSums[BLOCK / 64];
    const int lane = __lane_id();
    const int wave = threadIdx.x / 64;
    const uint32_t inc = waveInclusiveScan(v);
    if (lane == 63) waveSums[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        uint32_t s = lane < BLOCK / 64 ? waveSums[lane] : 0;
        s = waveInclusiveScan(s);
        if (lane < BLOCK / 64) waveSums[lane] = s;
    }
    __syncthreads();
    const uint32_t waveBase = wave == 0 ? 0 : waveSums[wave - 1];
    total = waveSums[BLOCK / 64 - 1];
    __syncthreads();
    return waveBase + inc - v;
}

template <typename U>
__global__ __launch_bounds__(kBlock) void countKernel(const U* __restrict__ in, size_t n, int shift,
                                                      uint32_t* __restrict__ counts, uint32_t numTiles) {
    __shared__ uint32_t hist[kRadix];
    for (int i = threadIdx.x; i < kRadix; i += kBlock) hist[i] = 0;
    __syncthreads();
    const size_t base = static_cast<size_t>(blockIdx.x) * kTile;
#pragma unroll 4
    for (int k = 0; k < kItems; ++k) {
        const size_t idx = base + static_cast<size_t>(k) * kBlock + threadIdx.x;
        if (idx < n) atomicAdd(&hist[(in[idx] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kRadix; i += kBlock) counts[static_cast<size_t>(i) * numTiles + blockIdx.x] = hist[i];
}

/// Sum of each kScanChunk-long chunk of `data`.  (The scan kernels are
/// templates only so that every translation unit may instantiate them.)
template <typename C = uint32_t>
__global__ __launch_bounds__(kScanBlock) void scanReduceKernel(const C* __restrict__ data, size_t m,
                                                               C* __restrict__ partial) {
    const size_t base = static_cast<size_t>(blockIdx.x) * kScanChunk;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const size_t idx = base + static_cast<size_t>(threadIdx.x) * kScanItems + k;
        if (idx < m) s += data[idx];
    }
    uint32_t total;
    blockExclusiveScan<kScanBlock>(s, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

/// In-place exclusive scan of the chunk partials (one block, any count).
template <typename C = uint32_t>
__global__ __launch_bounds__(kScanBlock) void scanPartialsKernel(C* partial, uint32_t count) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < count; base += kScanBlock) {
        const uint32_t idx = base + threadIdx.x;
        const uint32_t v = idx < count ? partial[idx] : 0;
        uint32_t total;
        const uint32_t ex = blockExclusiveScan<kScanBlock>(v, total);
        if (idx < count) partial[idx] = carry + ex;
        carry += total;
    }
}

/// In-place exclusive scan of each chunk, offset by its partial.
template <typename C = uint32_t>
__global__ __launch_bounds__(kScanBlock) void scanDownKernel(C* __restrict__ data, size_t m,
                                                             const C* __restrict__ partial) {
    const size_t base = static_cast<size_t>(blockIdx.x) * kScanChunk + static_cast<size_t>(threadIdx.x) * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < m) ? data[base + k] : 0;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = blockExclusiveScan<kScanBlock>(s, total) + partial[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < m) data[base + k] = run;
        run += v[k];
    }
}

/// Lanes of the wave holding the same digit as this lane (among `valid` lanes).
__device__ __forceinline__ uint64_t digitPeers(uint32_t d, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kDigitBits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

template <typename U>
__global__ __launch_bounds__(kBlock) void scatterKernel(const U* __restrict__ in, U* __restrict__ out, size_t n,
                                                        int shift, const uint32_t* __restrict__ offsets,
                                                        uint32_t numTiles) {
    __shared__ uint32_t cursor[kWaves][kRadix];
    const int lane = __lane_id();
    const int wave = threadIdx.x / 64;
    for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&cursor[0][0])[i] = 0;
    __syncthreads();

    const size_t tileBase = static_cast<size_t>(blockIdx.x) * kTile;
    const size_t waveBase = tileBase + static_cast<size_t>(wave) * kRowsPerWave * 64;
    const uint64_t ltMask = (1ull << lane) - 1;

    // per-wave digit counts over this wave's rows
    for (int r = 0; r < kRowsPerWave; ++r) {
        const size_t idx = waveBase + static_cast<size_t>(r) * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = valid ? static_cast<uint32_t>((in[idx] >> shift) & (kRadix - 1)) : 0;
        const uint64_t peers = digitPeers(d, valid);
        if (valid && (peers & ltMask) == 0) cursor[wave][d] += static_cast<uint32_t>(__popcll(peers));
    }
    __syncthreads();
    // turn counts into output cursors: global offset of (digit, tile) + earlier waves
    for (int d = threadIdx.x; d < kRadix; d += kBlock) {
        uint32_t run = offsets[static_cast<size_t>(d) * numTiles + blockIdx.x];
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = cursor[w][d];
            cursor[w][d] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int r = 0; r < kRowsPerWave; ++r) {
        const size_t idx = waveBase + static_cast<size_t>(r) * 64 + lane;
        const bool valid = idx < n;
        const U val = valid ? in[idx] : U(0);
        const uint32_t d = valid ? static_cast<uint32_t>((val >> shift) & (kRadix - 1)) : 0;
        const uint64_t peers = digitPeers(d, valid);
        const uint32_t pos = cursor[wave][d] + static_cast<uint32_t>(__popcll(peers & ltMask));
        if (valid) out[pos] = val;
        if (valid && (peers & ltMask) == 0) cursor[wave][d] += static_cast<uint32_t>(__popcll(peers));
    }
}

/// Scratch needed to sort n values: 2 value buffers are the caller's; this is
/// the count/scan space in bytes.
inline size_t scratchBytes(size_t n) {
    const size_t tiles = (n + kTile - 1) / kTile;
    const size_t m = tiles * kRadix;
    const size_t parts = (m + kScanChunk - 1) / kScanChunk;
    return (m + parts) * sizeof(uint32_t) + 256;
}

/// Sorts n values by bits [lo, hi).  Ping-pongs between a and b; returns the
/// buffer that holds the sorted result.  n must be below 2^32.
template <typename U>
inline hipError_t sortBits(U* a, U* b, size_t n, int lo, int hi, void* scratch, hipStream_t stream,
                           U** result) {
    *result = a;
    if (n <= 1 || hi <= lo) return hipSuccess;
    const uint32_t tiles = static_cast<uint32_t>((n + kTile - 1) / kTile);
    const size_t m = static_cast<size_t>(tiles) * kRadix;
    const uint32_t parts = static_cast<uint32_t>((m + kScanChunk - 1) / kScanChunk);
    uint32_t* counts = static_cast<uint32_t*>(scratch);
    uint32_t* partial = counts + m;
    U* src = a;
    U* dst = b;
    for (int shift = lo; shift < hi; shift += kDigitBits) {
        countKernel<U><<<tiles, kBlock, 0, stream>>>(src, n, shift, counts, tiles);
        scanReduceKernel<uint32_t><<<parts, kScanBlock, 0, stream>>>(counts, m, partial);
        scanPartialsKernel<uint32_t><<<1, kScanBlock, 0, stream>>>(partial, parts);
        scanDownKernel<uint32_t><<<parts, kScanBlock, 0, stream>>>(counts, m, partial);
        scatterKernel<U><<<tiles, kBlock, 0, stream>>>(src, dst, n, shift, counts, tiles);
        U* t = src;
        src = dst;
        dst = t;
    }
    *result = src;
    return hipGetLastError();
}

}  // namespace cuckoo::radix
