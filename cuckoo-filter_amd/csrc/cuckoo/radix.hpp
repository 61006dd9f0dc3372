// Radix partitions over a bit range of 32/64-bit values, written for wave64:
// LSD passes of one digit each (sortBits) and a two-level MSD partition
// (partitionTwoLevel).
//
// Used by the bucket-binned insert (binned.hpp) to group packed
// (primary bucket, fingerprint) entries by the table slice their primary
// bucket falls in.  Only the slice-index bits are sorted, 8 per pass: order
// inside a slice has no effect on what a concurrent insert may produce, so
// the reference's full 64-bit sort (CuckooFilter.cuh:485-498) is not needed.
//
// Each pass is count -> scan -> scatter over tiles of kTile values:
//  * count: per-tile digit histogram in LDS, written as one contiguous row
//    of a tile-major count matrix;
//  * scan: the digit-major exclusive scan of that matrix, done over chunks of
//    rows so that every access to it is a contiguous row;
//  * scatter: a tile is split into 64-value rows, each wave owns a contiguous
//    run of rows; values are first placed digit-ordered in an LDS staging
//    tile, then written out so that each digit's run is a contiguous,
//    coalesced store.  Passes after the first of an LSD partition must be
//    stable: there a value's rank inside its row comes from eight ballots
//    (one per digit bit).  A first pass need not be, and ranks each value
//    with one LDS atomic on its digit's counter.
// The first pass reads its values through a source functor, so the values
// (e.g. entries computed from keys) need not be materialised beforehand; its
// scatter runs as a resident grid that loads its next tile while it hashes,
// ranks and writes the current one.  Later passes read arrays and scatter one
// tile per workgroup.
//
// partitionTwoLevel is the MSD variant for two digits: after the first pass
// each partition is split by the second digit on its own, so the bits of
// the first digit can be dropped from the values (a scatter may write its
// values through an output transform), and the group starts come out of
// the second pass's scan.  Neither of its passes needs a stable order.  Its
// first pass works on wide tiles (kWidePartsFor<value> tiles of kTile
// values) for longer output runs per digit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace cuckoo::radix {

inline constexpr int kDigitBits = 8;
inline constexpr int kRadix = 1 << kDigitBits;
inline constexpr int kBlock = 512;
inline constexpr int kWaves = kBlock / 64;
inline constexpr int kItems = 8;
inline constexpr int kTile = kBlock * kItems;      // 4096 values per tile (staging <= 32 KiB of LDS)
inline constexpr int kRowsPerWave = kTile / 64 / kWaves;
inline constexpr int kScanBlock = 1024;
inline constexpr int kScanItems = 4;
inline constexpr int kScanChunk = kScanBlock * kScanItems;

static_assert(kTile % (64 * kWaves) == 0, "tile must split evenly into wave rows");
static_assert(kRowsPerWave == kItems, "each lane keeps one value per row in registers");

/// Reads element i of an array.  Sources split a read into load() (the
/// memory access, base()[i]) and cook() (any arithmetic on it), so that a
/// kernel can issue the loads of its next tile before it needs the values.
template <typename U>
struct ArraySource {
    using Raw = U;
    const U* data;
    __device__ __forceinline__ Raw load(size_t i) const { return data[i]; }
    /// load(tileBase + off) with a uniform tileBase and a 32-bit offset
    __device__ __forceinline__ Raw loadRel(size_t tileBase, uint32_t off) const { return (data + tileBase)[off]; }
    __device__ __forceinline__ U cook(Raw r) const { return r; }
    __device__ __forceinline__ U operator()(size_t i) const { return data[i]; }
};

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

template <typename U, typename Src>
__global__ __launch_bounds__(kBlock) void countKernel(Src src, size_t n, int shift, uint32_t* __restrict__ counts,
                                                      uint32_t numTiles) {
    __shared__ uint32_t hist[kRadix];
    for (int i = threadIdx.x; i < kRadix; i += kBlock) hist[i] = 0;
    __syncthreads();
    const size_t base = static_cast<size_t>(blockIdx.x) * kTile;
    // all loads first and unconditional (past-the-end lanes read element
    // n - 1 and are not counted), then the arithmetic and the LDS counts
    const size_t last = n - 1;
    typename Src::Raw raw[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const size_t idx = base + static_cast<size_t>(k) * kBlock + threadIdx.x;
        raw[k] = src.load(idx < last ? idx : last);
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const size_t idx = base + static_cast<size_t>(k) * kBlock + threadIdx.x;
        const uint32_t d = static_cast<uint32_t>(static_cast<U>(src.cook(raw[k])) >> shift) & (kRadix - 1);
        if (idx < n) atomicAdd(&hist[d], 1u);
    }
    __syncthreads();
    // tile-major: this tile's 256 counts are one contiguous row
    for (int i = threadIdx.x; i < kRadix; i += kBlock) counts[static_cast<size_t>(blockIdx.x) * kRadix + i] = hist[i];
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

/// The digit-ordered exclusive scan of a tile-major count matrix
/// (counts[tile][digit]) runs over chunks of kChunkTiles rows, so that every
/// access to the matrix is a contiguous row:
///  1. chunkSumKernel: per chunk, the per-digit sums, stored digit-major
///     (chunkSum[digit][chunk], a small array);
///  2. the linear scan above over chunkSum: its exclusive prefix in
///     digit-major order is each (digit, chunk)'s global start;
///  3. chunkDownKernel: per chunk, each digit's running offset down the rows,
///     written back in place.
inline constexpr uint32_t kChunkTiles = 256;

template <typename C = uint32_t>
__global__ __launch_bounds__(kRadix) void chunkSumKernel(const C* __restrict__ counts, uint32_t numTiles,
                                                         C* __restrict__ chunkSum, uint32_t numChunks) {
    const uint32_t c = blockIdx.x;
    const uint32_t t0 = c * kChunkTiles;
    const uint32_t t1 = numTiles - t0 < kChunkTiles ? numTiles : t0 + kChunkTiles;
    uint32_t sum = 0;
#pragma unroll 8
    for (uint32_t t = t0; t < t1; ++t) sum += counts[static_cast<size_t>(t) * kRadix + threadIdx.x];
    chunkSum[static_cast<size_t>(threadIdx.x) * numChunks + c] = sum;
}

template <typename C = uint32_t>
__global__ __launch_bounds__(kRadix) void chunkDownKernel(C* __restrict__ counts, uint32_t numTiles,
                                                          const C* __restrict__ chunkStart, uint32_t numChunks) {
    const uint32_t c = blockIdx.x;
    const uint32_t t0 = c * kChunkTiles;
    const uint32_t t1 = numTiles - t0 < kChunkTiles ? numTiles : t0 + kChunkTiles;
    uint32_t run = chunkStart[static_cast<size_t>(threadIdx.x) * numChunks + c];
#pragma unroll 8
    for (uint32_t t = t0; t < t1; ++t) {
        C* p = counts + static_cast<size_t>(t) * kRadix + threadIdx.x;
        const uint32_t v = *p;
        *p = run;
        run += v;
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

/// Output transform of a scatter: the value written for a staged value
/// (the digit is always taken from the staged value).  Identity by default.
template <typename U>
struct Identity {
    using Out = U;
    __device__ __forceinline__ Out operator()(U v) const { return v; }
};

/// Dynamic LDS bytes the scatter kernel needs for value type U.
template <typename U>
constexpr size_t scatterLdsBytes() {
    return static_cast<size_t>(kTile) * sizeof(U);
}

/// Dynamic LDS bytes of the unstable scatters for output values O: the
/// staged values, then one digit byte per value.
template <typename O>
constexpr size_t unstableLdsBytes() {
    return static_cast<size_t>(kTile) * (sizeof(O) + 1);
}

/// Steps of an unstable one-tile scatter, for passes whose order inside a
/// digit does not matter (the first pass of any partition, and the second
/// pass of the two-level one).  A value's rank among the tile's values of
/// its digit is the return of one LDS atomic on that digit's counter (kept
/// in a register), so a value is ranked once, with no ballots.  `vals` holds
/// this lane's output values (row r of wave w is tile positions (w *
/// kRowsPerWave + r) * 64 + lane), already transformed, and `digits` their
/// digits (8 bits per row); the first `tileCount` are valid.  `myBase` is,
/// for threads below kRadix, the global offset of digit threadIdx.x for this
/// tile.  `hist` (kRadix counters) must be zero on entry, with a barrier
/// passed.
template <typename O>
__device__ __forceinline__ void scatterTileUnstable(const O (&vals)[kRowsPerWave], uint64_t digits,
                                                    uint32_t tileCount, uint32_t myBase, O* __restrict__ out,
                                                    O* stage, uint8_t* dstage, uint32_t* hist, uint32_t* runShift) {
    static_assert(kRowsPerWave * kDigitBits <= 64, "digits of a lane must pack into 64 bits");
    const int lane = __lane_id();
    const int wave = threadIdx.x / 64;
    const uint32_t waveBase = static_cast<uint32_t>(wave) * kRowsPerWave * 64;
    // 1. ranks (two 16-bit ranks per register: a tile has 4096 values)
    uint32_t rank[kRowsPerWave / 2];
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
        const bool valid = waveBase + static_cast<uint32_t>(r) * 64 + lane < tileCount;
        const uint32_t d = static_cast<uint32_t>(digits >> (r * kDigitBits)) & (kRadix - 1);
        const uint32_t k = valid ? atomicAdd(&hist[d], 1u) : 0u;
        if (r % 2 == 0) {
            rank[r / 2] = k;
        } else {
            rank[r / 2] |= k << 16;
        }
    }
    __syncthreads();
    // 2. digit starts inside the tile, and each digit's shift to its global place
    const uint32_t c = threadIdx.x < kRadix ? hist[threadIdx.x] : 0u;
    uint32_t total;
    const uint32_t start = blockExclusiveScan<kBlock>(c, total);
    if (threadIdx.x < kRadix) {
        hist[threadIdx.x] = start;
        runShift[threadIdx.x] = myBase - start;
    }
    __syncthreads();
    // 3. stage values and their digits digit-ordered in LDS
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
        const bool valid = waveBase + static_cast<uint32_t>(r) * 64 + lane < tileCount;
        const uint32_t d = static_cast<uint32_t>(digits >> (r * kDigitBits)) & (kRadix - 1);
        const uint32_t k = (rank[r / 2] >> (16 * (r % 2))) & 0xFFFFu;
        if (valid) {
            const uint32_t pos = hist[d] + k;
            stage[pos] = vals[r];
            dstage[pos] = static_cast<uint8_t>(d);
        }
    }
    __syncthreads();
    // 4. contiguous runs per digit to their global place.  A fixed count of
    //    unconditional stores (past-the-end lanes repeat the tile's last value
    //    at its own place, an identical write): the compiler can then count
    //    them, so that a pipelined caller's wait for its prefetched loads does
    //    not also wait for these stores, and no loop preheader flushes the
    //    counter.
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint32_t j0 = static_cast<uint32_t>(it) * kBlock + threadIdx.x;
        const uint32_t j = j0 < tileCount ? j0 : tileCount - 1;
        out[static_cast<uint32_t>(runShift[dstage[j]] + j)] = stage[j];  // exact: the sum is below n < 2^32
        if (it % 2 == 1) __builtin_amdgcn_sched_barrier(0);  // bounds the registers held in flight
    }
}

/// Pipelined, for sources that compute their values (such as hashing keys):
/// a resident grid in which workgroup b handles tiles b, b + gridDim.x, ...
/// and issues the loads of its next tile before the current tile is ranked,
/// staged and written, so the loads overlap the arithmetic and LDS work.  It
/// is held to 80 VGPRs so that three workgroups fit per CU; at that bound the
/// compiler spills a few loop invariants, which measured faster (5.94 against
/// 6.03 ms per 255M-key insert) than 96 VGPRs without spills and two
/// workgroups per CU.
template <typename U, typename Src, typename P = Identity<U>>
__global__ __launch_bounds__(kBlock, 6) void scatterKernel(Src src, typename P::Out* __restrict__ out, size_t n,
                                                           int shift, const uint32_t* __restrict__ offsets,
                                                           uint32_t numTiles, P pack = P{}) {
    using Raw = typename Src::Raw;
    using O = typename P::Out;
    extern __shared__ __align__(16) unsigned char ldsRaw[];
    O* stage = reinterpret_cast<O*>(ldsRaw);
    uint8_t* dstage = ldsRaw + static_cast<size_t>(kTile) * sizeof(O);
    __shared__ uint32_t hist[kRadix];
    __shared__ uint32_t runShift[kRadix];  // global start of the digit's run minus its start in the tile

    const int lane = __lane_id();
    const int wave = threadIdx.x / 64;

    uint32_t t = blockIdx.x;
    if (t >= numTiles) return;
    // Loads are unconditional (past-the-end lanes read element n - 1 and are
    // masked out later), so the compiler can count them: a wait for one
    // value then does not wait for the loads issued after it.  Addresses are
    // a tile base (uniform, scalar registers) plus a 32-bit lane offset.
    const uint32_t laneBase = static_cast<uint32_t>(wave) * kRowsPerWave * 64 + lane;
    auto loadTile = [&](uint32_t tile, Raw (&dst)[kRowsPerWave]) {
        const size_t tb = static_cast<size_t>(tile) * kTile;
        const uint32_t lastLocal = static_cast<uint32_t>(n - 1 - tb < kTile - 1 ? n - 1 - tb : kTile - 1);
#pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r) {
            const uint32_t li = laneBase + static_cast<uint32_t>(r) * 64;
            dst[r] = src.loadRel(tb, (li < lastLocal ? li : lastLocal) & (kTile - 1));
        }
    };
    Raw raw[kRowsPerWave];
    loadTile(t, raw);
    for (; t < numTiles; t += gridDim.x) {
        const size_t tileBase = static_cast<size_t>(t) * kTile;
        const uint32_t tileCount =
            static_cast<uint32_t>(n - tileBase < static_cast<size_t>(kTile) ? n - tileBase : static_cast<size_t>(kTile));
        if (threadIdx.x < kRadix) hist[threadIdx.x] = 0;
        // this tile's output values and digits; then the next tile's loads go out
        O vals[kRowsPerWave];
        uint64_t digits = 0;
#pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r) {
            const U c = src.cook(raw[r]);
            digits |= static_cast<uint64_t>(static_cast<uint32_t>(c >> shift) & (kRadix - 1)) << (r * kDigitBits);
            vals[r] = pack(c);
            if (r % 2 == 1) __builtin_amdgcn_sched_barrier(0);  // bounds the hashes in flight
        }
        // this tile's digit bases, loaded before the prefetch so that waiting
        // for them does not wait for the next tile's loads
        const uint32_t myBase =
            offsets[static_cast<size_t>(t) * kRadix + (threadIdx.x < kRadix ? threadIdx.x : kRadix - 1)];
        // (past the last tile this reloads the current one; those values go unused)
        loadTile(t + gridDim.x < numTiles ? t + gridDim.x : t, raw);
        __syncthreads();  // hist cleared
        scatterTileUnstable<O>(vals, digits, tileCount, myBase, out, stage, dstage, hist, runShift);
        __syncthreads();  // stage, hist and the bases are reused by the next tile
    }
}

/// Values per tile of the wide first pass: PARTS parts of kTile, so that a
/// digit's run in the output is PARTS times as long (with 256 digits, 16 *
/// PARTS values: fewer partial-line writes) while a lane still holds one
/// part's values at a time.
template <int PARTS>
inline constexpr int kWideTileOf = PARTS * kTile;

/// Parts per wide tile in the two-level partition's first pass, by output
/// value size.  32-bit values: three (48-value runs, 60 KiB of LDS, two
/// workgroups per CU) measured faster than two (32-value runs, three
/// workgroups per CU): 5.53 against 5.63 ms per 255M-key insert; four would
/// leave one workgroup per CU.  64-bit values: two (72 KiB, two per CU).
template <typename O>
inline constexpr int kWidePartsFor = sizeof(O) <= 4 ? 3 : 2;

/// Digit histogram per wide tile (tile-major rows, as countKernel).
template <typename U, typename Src, int PARTS>
__global__ __launch_bounds__(kBlock) void countWideKernel(Src src, size_t n, int shift, uint32_t* __restrict__ counts) {
    constexpr int kWideTile = kWideTileOf<PARTS>;
    __shared__ uint32_t hist[kRadix];
    for (int i = threadIdx.x; i < kRadix; i += kBlock) hist[i] = 0;
    __syncthreads();
    const size_t base = static_cast<size_t>(blockIdx.x) * kWideTile;
    const size_t last = n - 1;
    typename Src::Raw raw[PARTS * kItems];
#pragma unroll
    for (int k = 0; k < PARTS * kItems; ++k) {
        const size_t idx = base + static_cast<size_t>(k) * kBlock + threadIdx.x;
        raw[k] = src.load(idx < last ? idx : last);
    }
#pragma unroll
    for (int k = 0; k < PARTS * kItems; ++k) {
        const size_t idx = base + static_cast<size_t>(k) * kBlock + threadIdx.x;
        const uint32_t d = static_cast<uint32_t>(static_cast<U>(src.cook(raw[k])) >> shift) & (kRadix - 1);
        if (idx < n) atomicAdd(&hist[d], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kRadix; i += kBlock) counts[static_cast<size_t>(blockIdx.x) * kRadix + i] = hist[i];
}

/// Dynamic LDS of scatterWideKernel for output values O.
template <typename O, int PARTS>
constexpr size_t wideLdsBytes() {
    return static_cast<size_t>(kWideTileOf<PARTS>) * (sizeof(O) + 1);
}

/// The pipelined first-pass scatter over wide tiles (unstable).  A tile's
/// digit counts are differences of the scanned offset matrix (the offset of
/// (digit, tile + 1), or of (digit + 1, tile 0) after the last tile, minus
/// that of (digit, tile)), so each digit's place in the tile's staging area
/// is known before any value is ranked: a value's staging position is the
/// return of one LDS atomic on its digit's cursor, and each part of the tile
/// is staged as soon as it is computed.  The loads of the next part (of this
/// tile, then of the next tile) go out before the current part is staged.
template <typename U, typename Src, typename P, int PARTS>
__global__ __launch_bounds__(kBlock, 6) void scatterWideKernel(Src src, typename P::Out* __restrict__ out, size_t n,
                                                               int shift, const uint32_t* __restrict__ offsets,
                                                               uint32_t numTiles, P pack) {
    constexpr int kWideTile = kWideTileOf<PARTS>;
    using Raw = typename Src::Raw;
    using O = typename P::Out;
    extern __shared__ __align__(16) unsigned char ldsRaw[];
    O* stage = reinterpret_cast<O*>(ldsRaw);
    uint8_t* dstage = ldsRaw + static_cast<size_t>(kWideTile) * sizeof(O);
    __shared__ uint32_t cursor[kRadix];
    __shared__ uint32_t runShift[kRadix];

    uint32_t t = blockIdx.x;
    if (t >= numTiles) return;
    const int lane = __lane_id();
    const int wave = threadIdx.x / 64;
    const uint32_t laneBase = static_cast<uint32_t>(wave) * kRowsPerWave * 64 + lane;
    // Loads are unconditional: past-the-end lanes (and a whole half past the
    // end) read element n - 1, and are masked out when staged.
    auto loadHalf = [&](uint32_t tile, int half, Raw (&dst)[kRowsPerWave]) {
        const size_t hb0 = static_cast<size_t>(tile) * kWideTile + static_cast<size_t>(half) * kTile;
        const size_t hb = hb0 < n ? hb0 : n - 1;
        const uint32_t lastLocal = static_cast<uint32_t>(n - 1 - hb < kTile - 1 ? n - 1 - hb : kTile - 1);
#pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r) {
            const uint32_t li = laneBase + static_cast<uint32_t>(r) * 64;
            dst[r] = src.loadRel(hb, (li < lastLocal ? li : lastLocal) & (kTile - 1));
        }
    };
    // a half's output values and digits (8 bits per row), from the loaded raw values
    O vals[kRowsPerWave];
    uint64_t digits = 0;
    auto cookHalf = [&](const Raw (&in)[kRowsPerWave]) {
        digits = 0;
#pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r) {
            const U c = src.cook(in[r]);
            digits |= static_cast<uint64_t>(static_cast<uint32_t>(c >> shift) & (kRadix - 1)) << (r * kDigitBits);
            vals[r] = pack(c);
        }
    };
    // stages a half: each valid value at its digit's cursor
    auto stageHalf = [&](size_t halfBase) {
#pragma unroll
        for (int r = 0; r < kRowsPerWave; ++r) {
            const uint32_t d = static_cast<uint32_t>(digits >> (r * kDigitBits)) & (kRadix - 1);
            if (halfBase + laneBase + static_cast<uint32_t>(r) * 64 < n) {
                const uint32_t pos = atomicAdd(&cursor[d], 1u);
                stage[pos] = vals[r];
                dstage[pos] = static_cast<uint8_t>(d);
            }
        }
    };
    Raw raw[kRowsPerWave];
    loadHalf(t, 0, raw);
    for (; t < numTiles; t += gridDim.x) {
        const size_t tileBase = static_cast<size_t>(t) * kWideTile;
        const uint32_t tileCount = static_cast<uint32_t>(
            n - tileBase < static_cast<size_t>(kWideTile) ? n - tileBase : static_cast<size_t>(kWideTile));
        // this tile's digit counts and bases
        uint32_t myBase = 0, cnt = 0;
        if (threadIdx.x < kRadix) {
            const uint32_t d = threadIdx.x;
            myBase = offsets[static_cast<size_t>(t) * kRadix + d];
            const uint32_t next = t + 1 < numTiles ? offsets[static_cast<size_t>(t + 1) * kRadix + d]
                                                   : (d + 1 < kRadix ? offsets[d + 1] : static_cast<uint32_t>(n));
            cnt = next - myBase;
        }
        uint32_t total;
        const uint32_t start = blockExclusiveScan<kBlock>(cnt, total);
        if (threadIdx.x < kRadix) {
            cursor[threadIdx.x] = start;
            runShift[threadIdx.x] = myBase - start;
        }
        __syncthreads();  // cursors set
        // each part (the first already loaded); the next part's loads go out before it is staged
#pragma unroll
        for (int h = 0; h < PARTS; ++h) {
            cookHalf(raw);
            if (h + 1 < PARTS) {
                loadHalf(t, h + 1, raw);
            } else {
                loadHalf(t + gridDim.x < numTiles ? t + gridDim.x : t, 0, raw);
            }
            stageHalf(tileBase + static_cast<size_t>(h) * kTile);
        }
        __syncthreads();
        // contiguous runs per digit to their global place: a fixed count of
        // unconditional stores (past-the-end lanes repeat the last value)
#pragma unroll
        for (int it = 0; it < PARTS * kItems; ++it) {
            const uint32_t j0 = static_cast<uint32_t>(it) * kBlock + threadIdx.x;
            const uint32_t j = j0 < tileCount ? j0 : tileCount - 1;
            out[static_cast<uint32_t>(runShift[dstage[j]] + j)] = stage[j];  // exact: below n < 2^32
            if (it % 2 == 1) __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();  // staging area and cursors are reused by the next tile
    }
}

/// Steps 1-4 of a one-tile scatter, shared by the array and segmented passes.
/// `vals` holds this lane's values of the tile (row r of wave w is tile
/// positions (w * kRowsPerWave + r) * 64 + lane), of which the first
/// `tileCount` are valid; `rowOffsets` is the tile's row of global digit
/// offsets.  `cursor` must be zero on entry (and a barrier passed since).
template <typename U, typename P>
__device__ __forceinline__ void scatterOneTile(const U (&vals)[kRowsPerWave], uint32_t tileCount, int shift,
                                               const uint32_t* __restrict__ rowOffsets, P pack,
                                               typename P::Out* __restrict__ out, U* stage,
                                               uint32_t (*cursor)[kRadix], uint32_t* runShift) {
    const int lane = __lane_id();
    const int wave = threadIdx.x / 64;
    const uint32_t waveBase = static_cast<uint32_t>(wave) * kRowsPerWave * 64;
    const uint64_t ltMask = (1ull << lane) - 1;
    static_assert(kRowsPerWave * kDigitBits <= 64, "digits of a lane must pack into 64 bits");
    // 1. count digits per wave (the ballot ranking loop stays rolled, with the
    //    lane's digits packed into one register)
    uint64_t digits = 0;
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
        digits |= static_cast<uint64_t>(static_cast<uint32_t>(vals[r] >> shift) & (kRadix - 1)) << (r * kDigitBits);
    }
#pragma unroll 1
    for (int r = 0; r < kRowsPerWave; ++r) {
        const bool valid = waveBase + static_cast<uint32_t>(r) * 64 + lane < tileCount;
        const uint32_t d = static_cast<uint32_t>(digits >> (r * kDigitBits)) & (kRadix - 1);
        const uint64_t peers = digitPeers(d, valid);
        if (valid && (peers & ltMask) == 0) cursor[wave][d] += static_cast<uint32_t>(__popcll(peers));
    }
    __syncthreads();
    // 2. digit start inside the tile, each wave's offset inside the digit, and the global base
    uint32_t tileTotal = 0;
    if (threadIdx.x < kRadix) {
        const int d = threadIdx.x;
        for (int w = 0; w < kWaves; ++w) tileTotal += cursor[w][d];
    }
    uint32_t total;
    const uint32_t start = blockExclusiveScan<kBlock>(threadIdx.x < kRadix ? tileTotal : 0u, total);
    if (threadIdx.x < kRadix) {
        const int d = threadIdx.x;
        runShift[d] = rowOffsets[d] - start;
        uint32_t run = start;
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = cursor[w][d];
            cursor[w][d] = run;
            run += c;
        }
    }
    __syncthreads();
    // 3. stage values digit-ordered (stable) in LDS
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
        const bool valid = waveBase + static_cast<uint32_t>(r) * 64 + lane < tileCount;
        const uint32_t d = static_cast<uint32_t>(digits >> (r * kDigitBits)) & (kRadix - 1);
        const uint64_t peers = digitPeers(d, valid);
        const uint32_t pos = cursor[wave][d] + static_cast<uint32_t>(__popcll(peers & ltMask));
        if (valid) stage[pos] = vals[r];
        if (valid && (peers & ltMask) == 0) cursor[wave][d] += static_cast<uint32_t>(__popcll(peers));
    }
    __syncthreads();
    // 4. contiguous runs per digit to their global place
    for (uint32_t j = threadIdx.x; j < tileCount; j += kBlock) {
        const U v = stage[j];
        const uint32_t d = static_cast<uint32_t>(v >> shift) & (kRadix - 1);
        out[static_cast<uint32_t>(runShift[d] + j)] = pack(v);  // exact: the sum is below n < 2^32
    }
}

/// One tile per workgroup (the grid is the tile count): used for passes
/// that read their values straight from an array.
template <typename U, typename Src>
__global__ __launch_bounds__(kBlock) void scatterTileKernel(Src src, U* __restrict__ out, size_t n, int shift,
                                                        const uint32_t* __restrict__ offsets, uint32_t numTiles) {
    extern __shared__ __align__(16) unsigned char ldsRaw[];
    U* stage = reinterpret_cast<U*>(ldsRaw);
    __shared__ uint32_t cursor[kWaves][kRadix];
    __shared__ uint32_t runShift[kRadix];  // global start of the digit's run minus its start in the tile

    const int lane = __lane_id();
    const int wave = threadIdx.x / 64;
    for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&cursor[0][0])[i] = 0;
    __syncthreads();

    const size_t tileBase = static_cast<size_t>(blockIdx.x) * kTile;
    const size_t waveBase = tileBase + static_cast<size_t>(wave) * kRowsPerWave * 64;
    const uint32_t tileCount =
        static_cast<uint32_t>(n - tileBase < static_cast<size_t>(kTile) ? n - tileBase : static_cast<size_t>(kTile));
    // all rows of a lane are loaded first (many loads in flight per wave)
    U vals[kRowsPerWave];
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
        const size_t idx = waveBase + static_cast<size_t>(r) * 64 + lane;
        vals[r] = idx < n ? static_cast<U>(src(idx)) : U(0);
    }
    scatterOneTile<U, Identity<U>>(vals, tileCount, shift, offsets + static_cast<size_t>(blockIdx.x) * kRadix,
                                   Identity<U>{}, out, stage, cursor, runShift);
}

// Segmented pass: after a first pass has grouped the values into up to
// kRadix partitions (the most significant digit), each partition is split
// by the next digit on its own.  Tiles never straddle a partition, so the
// next digit's offsets are per partition and the partition's own bits need
// not be kept in the values.  Tile t belongs to the partition p with
// tileBase[p] <= t < tileBase[p + 1] and covers partStart[p] + (t -
// tileBase[p]) * kTile onwards, up to the partition's end.

/// partStart[p] (p <= kRadix, partStart[kRadix] = n) and tileBase[p] from the
/// first pass's offset matrix (row 0 holds every digit's global start).
template <typename C = uint32_t>
__global__ __launch_bounds__(kRadix) void segmentKernel(const C* __restrict__ firstOffsets, uint32_t n,
                                                        C* __restrict__ partStart, C* __restrict__ tileBase) {
    const uint32_t p = threadIdx.x;
    const uint32_t s = firstOffsets[p];
    const uint32_t e = p + 1 < kRadix ? firstOffsets[p + 1] : n;
    const uint32_t tiles = (e - s + kTile - 1) / kTile;
    uint32_t total;
    const uint32_t before = blockExclusiveScan<kRadix>(tiles, total);
    partStart[p] = s;
    tileBase[p] = before;
    if (p == 0) {
        partStart[kRadix] = n;
        tileBase[kRadix] = total;
    }
}

/// Partition p and tile range [lo, hi) of segmented tile t; false if t is
/// past the last tile.
__device__ __forceinline__ bool segmentTile(const uint32_t* __restrict__ partStart,
                                            const uint32_t* __restrict__ tileBase, uint32_t t, uint32_t& p,
                                            uint32_t& lo, uint32_t& hi) {
    if (t >= tileBase[kRadix]) return false;
    uint32_t a = 0, b = kRadix;  // the last p with tileBase[p] <= t
    while (b - a > 1) {
        const uint32_t mid = (a + b) / 2;
        if (tileBase[mid] <= t) {
            a = mid;
        } else {
            b = mid;
        }
    }
    p = a;
    lo = partStart[p] + (t - tileBase[p]) * kTile;
    const uint32_t end = partStart[p + 1];
    hi = end - lo < static_cast<uint32_t>(kTile) ? end : lo + kTile;
    return true;
}

/// Digit histogram of each segmented tile, one row of the tile-major matrix
/// per tile (only the first `radix` columns are written).  The digit is the
/// `radix` (a power of two) values at `shift`: bits above it, such as the
/// first-level digit of whole entries, are masked off.
template <typename U>
__global__ __launch_bounds__(kBlock) void countSegKernel(const U* __restrict__ in, const uint32_t* __restrict__ partStart,
                                                         const uint32_t* __restrict__ tileBase, int shift, uint32_t radix,
                                                         uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[kRadix];
    uint32_t p, lo, hi;
    if (!segmentTile(partStart, tileBase, blockIdx.x, p, lo, hi)) return;
    for (int i = threadIdx.x; i < kRadix; i += kBlock) hist[i] = 0;
    __syncthreads();
    const uint32_t last = hi - 1;
    U raw[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint32_t idx = lo + static_cast<uint32_t>(k) * kBlock + threadIdx.x;
        raw[k] = in[idx < last ? idx : last];
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint32_t idx = lo + static_cast<uint32_t>(k) * kBlock + threadIdx.x;
        if (idx < hi) atomicAdd(&hist[static_cast<uint32_t>(raw[k] >> shift) & (radix - 1)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < radix; i += kBlock) counts[static_cast<size_t>(blockIdx.x) * kRadix + i] = hist[i];
}

/// Per partition (one workgroup each): the digit-major exclusive scan of its
/// tiles' rows, offset by the partition's start, written back in place; and
/// the start of every (partition, digit) run as groupStart[p * radix + d]
/// for the groups below numGroups (groupStart[numGroups] = n).
template <typename C = uint32_t>
__global__ __launch_bounds__(kRadix) void segScanKernel(C* __restrict__ counts, const C* __restrict__ partStart,
                                                        const C* __restrict__ tileBase, uint32_t radix,
                                                        C* __restrict__ groupStart, uint32_t numGroups) {
    const uint32_t p = blockIdx.x;
    const uint32_t d = threadIdx.x;
    const uint32_t t0 = tileBase[p];
    const uint32_t t1 = tileBase[p + 1];
    uint32_t sum = 0;
    if (d < radix) {
#pragma unroll 8
        for (uint32_t t = t0; t < t1; ++t) sum += counts[static_cast<size_t>(t) * kRadix + d];
    }
    uint32_t total;
    uint32_t run = partStart[p] + blockExclusiveScan<kRadix>(sum, total);
    if (d < radix) {
        const uint32_t g = p * radix + d;
        if (g < numGroups) groupStart[g] = run;
#pragma unroll 8
        for (uint32_t t = t0; t < t1; ++t) {
            C* q = counts + static_cast<size_t>(t) * kRadix + d;
            const uint32_t v = *q;
            *q = run;
            run += v;
        }
    }
    if (p == 0 && d == 0) groupStart[numGroups] = partStart[kRadix];
}

/// The segmented pass's scatter: one segmented tile per workgroup; digits as
/// in countSegKernel.
template <typename U, typename P>
__global__ __launch_bounds__(kBlock) void scatterSegKernel(const U* __restrict__ in, typename P::Out* __restrict__ out,
                                                           const uint32_t* __restrict__ partStart,
                                                           const uint32_t* __restrict__ tileBase, int shift,
                                                           uint32_t radix, const uint32_t* __restrict__ offsets,
                                                           P pack) {
    using O = typename P::Out;
    extern __shared__ __align__(16) unsigned char ldsRaw[];
    O* stage = reinterpret_cast<O*>(ldsRaw);
    uint8_t* dstage = ldsRaw + static_cast<size_t>(kTile) * sizeof(O);
    __shared__ uint32_t hist[kRadix];
    __shared__ uint32_t runShift[kRadix];
    uint32_t p, lo, hi;
    if (!segmentTile(partStart, tileBase, blockIdx.x, p, lo, hi)) return;
    if (threadIdx.x < kRadix) hist[threadIdx.x] = 0;
    const int lane = __lane_id();
    const int wave = threadIdx.x / 64;
    const uint32_t waveBase = lo + static_cast<uint32_t>(wave) * kRowsPerWave * 64;
    const uint32_t last = hi - 1;
    O vals[kRowsPerWave];
    uint64_t digits = 0;
#pragma unroll
    for (int r = 0; r < kRowsPerWave; ++r) {
        const uint32_t idx = waveBase + static_cast<uint32_t>(r) * 64 + lane;
        const U v = in[idx < last ? idx : last];  // past-the-end lanes are masked by tileCount
        digits |= static_cast<uint64_t>(static_cast<uint32_t>(v >> shift) & (radix - 1)) << (r * kDigitBits);
        vals[r] = pack(v);
    }
    const uint32_t myBase =
        offsets[static_cast<size_t>(blockIdx.x) * kRadix + (threadIdx.x < kRadix ? threadIdx.x : kRadix - 1)];
    __syncthreads();  // hist cleared
    scatterTileUnstable<O>(vals, digits, hi - lo, myBase, out, stage, dstage, hist, runShift);
}

/// Grid of the pipelined scatter kernel: its resident workgroups on the
/// current device (at most one per tile).
template <typename U, typename Src, typename P = Identity<U>>
inline uint32_t scatterGrid(uint32_t numTiles) {
    int dev = 0;
    int cus = 0;
    int perCu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, scatterKernel<U, Src, P>, kBlock,
                                                     unstableLdsBytes<typename P::Out>()) !=
            hipSuccess ||
        cus <= 0 || perCu <= 0) {
        return numTiles;
    }
    const uint32_t resident = static_cast<uint32_t>(cus) * static_cast<uint32_t>(perCu);
    return numTiles < resident ? numTiles : resident;
}

/// Grid of the wide first-pass kernel: its resident workgroups (at most one per tile).
template <typename U, typename Src, typename P, int PARTS>
inline uint32_t wideGrid(uint32_t numTiles) {
    int dev = 0;
    int cus = 0;
    int perCu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, scatterWideKernel<U, Src, P, PARTS>, kBlock,
                                                     wideLdsBytes<typename P::Out, PARTS>()) != hipSuccess ||
        cus <= 0 || perCu <= 0) {
        return numTiles;
    }
    const uint32_t resident = static_cast<uint32_t>(cus) * static_cast<uint32_t>(perCu);
    return numTiles < resident ? numTiles : resident;
}

/// Scratch needed to partition n values: the two value buffers are the
/// caller's; this is the count/scan space in bytes.
inline size_t scratchBytes(size_t n) {
    const size_t tiles = (n + kTile - 1) / kTile;
    const size_t m = tiles * kRadix;
    const size_t chunks = (tiles + kChunkTiles - 1) / kChunkTiles;
    const size_t cm = chunks * kRadix;
    const size_t parts = (cm + kScanChunk - 1) / kScanChunk;
    return (m + cm + parts) * sizeof(uint32_t) + 256;
}

/// Digit-major exclusive scan of a tile-major count matrix, in place.
inline void scanCounts(uint32_t* counts, uint32_t tiles, uint32_t* chunkSum, uint32_t* partial, hipStream_t stream) {
    const uint32_t chunks = (tiles + kChunkTiles - 1) / kChunkTiles;
    const size_t cm = static_cast<size_t>(chunks) * kRadix;
    const uint32_t parts = static_cast<uint32_t>((cm + kScanChunk - 1) / kScanChunk);
    chunkSumKernel<uint32_t><<<chunks, kRadix, 0, stream>>>(counts, tiles, chunkSum, chunks);
    scanReduceKernel<uint32_t><<<parts, kScanBlock, 0, stream>>>(chunkSum, cm, partial);
    scanPartialsKernel<uint32_t><<<1, kScanBlock, 0, stream>>>(partial, parts);
    scanDownKernel<uint32_t><<<parts, kScanBlock, 0, stream>>>(chunkSum, cm, partial);
    chunkDownKernel<uint32_t><<<chunks, kRadix, 0, stream>>>(counts, tiles, chunkSum, chunks);
}

/// Scratch of partitionTwoLevel for n values, in bytes.
inline size_t twoLevelScratchBytes(size_t n) {
    // the second pass's count rows (up to kRadix more tiles), then the partition starts and tile bases
    return scratchBytes(n) + static_cast<size_t>(kRadix) * kRadix * sizeof(uint32_t) +
           2 * (kRadix + 1) * sizeof(uint32_t);
}

/// Two-level (MSD) partition of n values into groups p * radix2 + d:
///  1. the values computed by `first` are split by the kDigitBits-wide digit
///     at `shift1` into partitions p and written to `a` through `pack1`;
///  2. each partition is split by the digit at `shift2` of the packed value
///     (radix2 <= kRadix values) and written to `b` through `pack2`.
/// The packed values may be 32 or 64 bits wide.  groupStart[g] receives the
/// start of group g for g < numGroups, and groupStart[numGroups] = n.  n must
/// be below 2^31.
template <typename U, typename Src, typename P1, typename P2>
inline hipError_t partitionTwoLevel(Src first, size_t n, int shift1, P1 pack1, int shift2, uint32_t radix2, P2 pack2,
                                    typename P1::Out* a, typename P2::Out* b, void* scratch, uint32_t* groupStart,
                                    uint32_t numGroups, hipStream_t stream) {
    using V = typename P1::Out;
    if (n == 0) return hipSuccess;
    const uint32_t tiles = static_cast<uint32_t>((n + kTile - 1) / kTile);
    const uint32_t tiles2 = tiles + kRadix;  // each partition adds at most one partial tile
    const uint32_t chunks = (tiles + kChunkTiles - 1) / kChunkTiles;
    uint32_t* counts = static_cast<uint32_t*>(scratch);                      // [tile][digit], both passes
    uint32_t* chunkSum = counts + static_cast<size_t>(tiles2) * kRadix;      // [digit][chunk]
    uint32_t* partial = chunkSum + static_cast<size_t>(chunks) * kRadix;
    uint32_t* partStart = partial + (static_cast<size_t>(chunks) * kRadix + kScanChunk - 1) / kScanChunk;
    uint32_t* tileBase = partStart + kRadix + 1;
    constexpr int PARTS = kWidePartsFor<V>;
    const uint32_t wideTiles = static_cast<uint32_t>((n + kWideTileOf<PARTS> - 1) / kWideTileOf<PARTS>);
    countWideKernel<U, Src, PARTS><<<wideTiles, kBlock, 0, stream>>>(first, n, shift1, counts);
    scanCounts(counts, wideTiles, chunkSum, partial, stream);
    scatterWideKernel<U, Src, P1, PARTS>
        <<<wideGrid<U, Src, P1, PARTS>(wideTiles), kBlock, wideLdsBytes<V, PARTS>(), stream>>>(
            first, a, n, shift1, counts, wideTiles, pack1);
    segmentKernel<uint32_t><<<1, kRadix, 0, stream>>>(counts, static_cast<uint32_t>(n), partStart, tileBase);
    countSegKernel<V><<<tiles2, kBlock, 0, stream>>>(a, partStart, tileBase, shift2, radix2, counts);
    segScanKernel<uint32_t><<<kRadix, kRadix, 0, stream>>>(counts, partStart, tileBase, radix2, groupStart, numGroups);
    scatterSegKernel<V, P2><<<tiles2, kBlock, unstableLdsBytes<typename P2::Out>(), stream>>>(
        a, b, partStart, tileBase, shift2, radix2, counts, pack2);
    return hipGetLastError();
}

/// Sorts n values by bits [lo, hi), 8 bits per pass.  The first pass reads
/// through `first` (which may compute the values); later passes ping-pong
/// between a and b.  *result receives the buffer holding the sorted values.
/// n must be below 2^32 and hi > lo.
template <typename U, typename Src>
inline hipError_t sortBits(Src first, size_t n, int lo, int hi, U* a, U* b, void* scratch, hipStream_t stream,
                           U** result) {
    *result = a;
    if (n == 0 || hi <= lo) return hipSuccess;
    const uint32_t tiles = static_cast<uint32_t>((n + kTile - 1) / kTile);
    const size_t m = static_cast<size_t>(tiles) * kRadix;
    const uint32_t chunks = (tiles + kChunkTiles - 1) / kChunkTiles;
    uint32_t* counts = static_cast<uint32_t*>(scratch);                  // [tile][digit]
    uint32_t* chunkSum = counts + m;                                      // [digit][chunk]
    uint32_t* partial = chunkSum + static_cast<size_t>(chunks) * kRadix;  // chunk-sum partials
    U* dst = a;
    U* other = b;
    bool firstPass = true;
    for (int shift = lo; shift < hi; shift += kDigitBits) {
        if (firstPass) {
            countKernel<U, Src><<<tiles, kBlock, 0, stream>>>(first, n, shift, counts, tiles);
        } else {
            countKernel<U, ArraySource<U>><<<tiles, kBlock, 0, stream>>>(ArraySource<U>{other}, n, shift, counts,
                                                                          tiles);
        }
        scanCounts(counts, tiles, chunkSum, partial, stream);
        if (firstPass) {
            scatterKernel<U, Src><<<scatterGrid<U, Src>(tiles), kBlock, unstableLdsBytes<U>(), stream>>>(
                first, dst, n, shift, counts, tiles);
        } else {
            scatterTileKernel<U, ArraySource<U>><<<tiles, kBlock, scatterLdsBytes<U>(), stream>>>(
                ArraySource<U>{other}, dst, n, shift, counts, tiles);
        }
        firstPass = false;
        *result = dst;
        U* t = dst;
        dst = other;
        other = t;
    }
    return hipGetLastError();
}

}  // namespace cuckoo::radix
