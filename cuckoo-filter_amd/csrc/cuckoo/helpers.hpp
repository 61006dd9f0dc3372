// Small shared helpers: power-of-two arithmetic, SWAR lane tests over packed
// tag words, the HIP error macro and wave64 utilities.
//
// SWAR semantics follow include/helpers.cuh:78-151 of the reference: for a
// word of equal-width lanes, getZeroMask sets the high bit of every zero lane;
// its LOWEST set bit always marks the lowest zero lane exactly (bits above it
// may be borrow artefacts), which is the only bit the filter ever uses.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

// helpers.cuh:14-16
__host__ __device__ constexpr bool powerOfTwo(size_t n) {
    return n != 0 && (n & (n - 1)) == 0;
}

// helpers.cuh:31-45
__host__ __device__ constexpr size_t nextPowerOfTwo(size_t n) {
    if (powerOfTwo(n)) return n;
    n--;
    n |= n >> 1;
    n |= n >> 2;
    n |= n >> 4;
    n |= n >> 8;
    n |= n >> 16;
    n |= n >> 32;
    n++;
    return n;
}

// helpers.cuh:196
#define SDIV(x, y) (((x) + (y) - 1) / (y))

// helpers.cuh:202-210 with the HIP runtime: print file:line and the error
// string, then exit with the error code (the reference's error convention).
#define HIP_CALL(err)                                                                 \
    do {                                                                              \
        hipError_t err_ = (err);                                                      \
        if (err_ == hipSuccess) [[likely]] {                                          \
            break;                                                                    \
        }                                                                             \
        std::printf("%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(err_));       \
        std::exit(static_cast<int>(err_));                                            \
    } while (0)

// helpers.cuh:22-24 (one-dimensional grids)
__device__ __forceinline__ uint32_t globalThreadId() {
    return blockIdx.x * blockDim.x + threadIdx.x;
}

namespace cuckoo::detail {

template <typename TagType, typename WordType>
__host__ __device__ constexpr WordType laneOnes() {
    WordType r = 0;
    for (size_t s = 0; s < sizeof(WordType) * 8; s += sizeof(TagType) * 8) r |= WordType(1) << s;
    return r;
}

}  // namespace cuckoo::detail

// High bit of every zero lane (lowest set bit exact).  helpers.cuh:78-103.
template <typename TagType, typename WordType>
__host__ __device__ __forceinline__ constexpr WordType getZeroMask(WordType v) {
    static_assert(sizeof(WordType) == 4 || sizeof(WordType) == 8, "WordType must be 32 or 64 bits");
    static_assert(sizeof(TagType) <= sizeof(WordType), "TagType must fit within WordType");
    constexpr WordType ones = cuckoo::detail::laneOnes<TagType, WordType>();
    constexpr WordType highs = ones << (sizeof(TagType) * 8 - 1);
    return (v - ones) & ~v & highs;
}

// helpers.cuh:113-116
template <typename TagType, typename WordType>
__host__ __device__ __forceinline__ constexpr bool hasZero(WordType v) {
    return getZeroMask<TagType, WordType>(v) != 0;
}

// Tag broadcast into every lane.  helpers.cuh:126-151.
template <typename TagType, typename WordType>
__host__ __device__ __forceinline__ constexpr WordType replicateTag(TagType tag) {
    return static_cast<WordType>(tag) * cuckoo::detail::laneOnes<TagType, WordType>();
}

// Index of the lowest set bit (v != 0).
template <typename WordType>
__host__ __device__ __forceinline__ constexpr uint32_t lowestBit(WordType v) {
    if constexpr (sizeof(WordType) == 8) {
        return static_cast<uint32_t>(__builtin_ctzll(static_cast<unsigned long long>(v)));
    } else {
        return static_cast<uint32_t>(__builtin_ctz(static_cast<unsigned int>(v)));
    }
}
