// XXH64 for the cuckoo filter: fingerprints and bucket indices are derived
// from the 64-bit XXH64 (seed 0) of the raw little-endian bytes of a key or a
// tag, exactly as the reference does (include/hashutil.cuh:44-161).  The same
// constexpr code runs on the host and on gfx950.
//
// Device note: CDNA has no 64-bit integer multiply instruction; each 64-bit
// product below lowers to v_mul_lo_u32 / v_mul_hi_u32 / v_mad_u64_u32.  For
// an 8-byte key the whole hash is five such products, for a 2-byte tag six.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace xxhash {

inline constexpr uint64_t PRIME64_1 = 11400714785074694791ULL;
inline constexpr uint64_t PRIME64_2 = 14029467366897019727ULL;
inline constexpr uint64_t PRIME64_3 = 1609587929392839161ULL;
inline constexpr uint64_t PRIME64_4 = 9650029242287828579ULL;
inline constexpr uint64_t PRIME64_5 = 2870177450012600261ULL;

__host__ __device__ __forceinline__ constexpr uint64_t rotl64(uint64_t x, int r) {
    return (x << r) | (x >> (64 - r));
}

__host__ __device__ __forceinline__ constexpr uint64_t round64(uint64_t acc, uint64_t lane) {
    return rotl64(acc + lane * PRIME64_2, 31) * PRIME64_1;
}

__host__ __device__ __forceinline__ constexpr uint64_t merge64(uint64_t h, uint64_t v) {
    return (h ^ round64(0, v)) * PRIME64_1 + PRIME64_4;
}

__host__ __device__ __forceinline__ constexpr uint64_t avalanche(uint64_t h) {
    h ^= h >> 33;
    h *= PRIME64_2;
    h ^= h >> 29;
    h *= PRIME64_3;
    h ^= h >> 32;
    return h;
}

// Tail steps of XXH64 (8-, 4- and 1-byte inputs).
__host__ __device__ __forceinline__ constexpr uint64_t step8(uint64_t h, uint64_t v) {
    return rotl64(h ^ round64(0, v), 27) * PRIME64_1 + PRIME64_4;
}
__host__ __device__ __forceinline__ constexpr uint64_t step4(uint64_t h, uint32_t v) {
    return rotl64(h ^ (uint64_t(v) * PRIME64_1), 23) * PRIME64_2 + PRIME64_3;
}
__host__ __device__ __forceinline__ constexpr uint64_t step1(uint64_t h, uint8_t v) {
    return rotl64(h ^ (uint64_t(v) * PRIME64_5), 11) * PRIME64_1;
}

// XXH64 of a byte range (any length).
__host__ __device__ inline uint64_t xxhash64_bytes(const uint8_t* p, size_t len, uint64_t seed = 0) {
    const uint8_t* const end = p + len;
    auto ld64 = [](const uint8_t* q) {
        uint64_t v = 0;
        for (int i = 7; i >= 0; --i) v = (v << 8) | q[i];
        return v;
    };
    auto ld32 = [](const uint8_t* q) {
        return uint32_t(q[0]) | (uint32_t(q[1]) << 8) | (uint32_t(q[2]) << 16) |
               (uint32_t(q[3]) << 24);
    };
    uint64_t h;
    if (len >= 32) {
        uint64_t v1 = seed + PRIME64_1 + PRIME64_2, v2 = seed + PRIME64_2, v3 = seed,
                 v4 = seed - PRIME64_1;
        do {
            v1 = round64(v1, ld64(p));
            v2 = round64(v2, ld64(p + 8));
            v3 = round64(v3, ld64(p + 16));
            v4 = round64(v4, ld64(p + 24));
            p += 32;
        } while (p + 32 <= end);
        h = rotl64(v1, 1) + rotl64(v2, 7) + rotl64(v3, 12) + rotl64(v4, 18);
        h = merge64(merge64(merge64(merge64(h, v1), v2), v3), v4);
    } else {
        h = seed + PRIME64_5;
    }
    h += uint64_t(len);
    for (; p + 8 <= end; p += 8) h = step8(h, ld64(p));
    if (p + 4 <= end) {
        h = step4(h, ld32(p));
        p += 4;
    }
    for (; p < end; ++p) h = step1(h, *p);
    return avalanche(h);
}

// XXH64 (seed 0 by default) of the object representation of a trivially
// copyable value.  1-, 2-, 4- and 8-byte integers take closed forms that the
// compiler reduces to a handful of multiplies; anything else goes through the
// byte routine.  Matches hashutil.cuh:70-161 for every size.
template <typename T>
__host__ __device__ __forceinline__ uint64_t xxhash64(const T& key, uint64_t seed = 0) {
    static_assert(std::is_trivially_copyable_v<T>, "key must be trivially copyable");
    if constexpr (std::is_integral_v<T> && sizeof(T) == 8) {
        return avalanche(step8(seed + PRIME64_5 + 8, uint64_t(key)));
    } else if constexpr (std::is_integral_v<T> && sizeof(T) == 4) {
        return avalanche(step4(seed + PRIME64_5 + 4, uint32_t(key)));
    } else if constexpr (std::is_integral_v<T> && sizeof(T) == 2) {
        const uint32_t v = uint16_t(key);
        return avalanche(step1(step1(seed + PRIME64_5 + 2, uint8_t(v)), uint8_t(v >> 8)));
    } else if constexpr (std::is_integral_v<T> && sizeof(T) == 1) {
        return avalanche(step1(seed + PRIME64_5 + 1, uint8_t(key)));
    } else {
        uint8_t bytes[sizeof(T)];
        __builtin_memcpy(bytes, &key, sizeof(T));
        return xxhash64_bytes(bytes, sizeof(T), seed);
    }
}

}  // namespace xxhash
