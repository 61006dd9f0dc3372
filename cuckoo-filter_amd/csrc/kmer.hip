// k-mer keys for the filter: 2-bit packing of DNA k-mers on the device.
//
// Encoding as the reference's converter writes k-mer keys
// (src/kmc_to_binary.cpp:17-46): A=0, C=1, G=2, T=3, first base in the
// highest used bits, k <= 32.  A k-mer containing any other character gets
// no key (the converter skips it); here it is marked with the all-ones key,
// which no valid k-mer of k <= 31 can produce.  With `canonical` set the key
// is the smaller of the k-mer and its reverse complement (the form k-mer
// counters such as KMC store by default).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

__device__ __forceinline__ uint32_t baseCode(uint8_t c) {
    switch (c | 0x20) {  // lower-case
        case 'a':
            return 0;
        case 'c':
            return 1;
        case 'g':
            return 2;
        case 't':
            return 3;
        default:
            return 4;
    }
}

__global__ __launch_bounds__(256) void encodeKmersKernel(const uint8_t* __restrict__ seq, size_t count, uint32_t k,
                                                         int canonical, uint64_t* __restrict__ out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const uint64_t mask = k == 32 ? ~0ull : ((1ull << (2 * k)) - 1);
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        uint64_t fwd = 0;
        uint64_t rev = 0;
        bool valid = true;
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t c = baseCode(seq[i + j]);
            valid &= c < 4;
            fwd = (fwd << 2) | (c & 3);
            rev |= static_cast<uint64_t>(3 - (c & 3)) << (2 * j);  // complement, reversed
        }
        uint64_t key = fwd & mask;
        if (canonical && (rev & mask) < key) key = rev & mask;
        out[i] = valid ? key : ~0ull;
    }
}

}  // namespace

extern "C" int cfk_encode_kmers_device(const void* d_seq, size_t len, uint32_t k, int canonical, uint64_t* d_out,
                                       void* stream) {
    if (k == 0 || k > 32 || len < k) return -1;
    const size_t count = len - k + 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t blocks = (count + 255) / 256;
    const uint32_t grid = static_cast<uint32_t>(blocks < 65536 ? blocks : 65536);
    encodeKmersKernel<<<grid, 256, 0, s>>>(static_cast<const uint8_t*>(d_seq), count, k, canonical, d_out);
    return static_cast<int>(hipGetLastError());
}
