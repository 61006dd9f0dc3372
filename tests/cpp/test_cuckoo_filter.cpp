// The reference's behavioural tests (tests/test_cuckoo_filter.cu:32-344),
// rewritten against this build's C++ template drop-in
// (cuckoo-filter_amd/csrc/cuckoo/CuckooFilter.hpp) and its rocThrust
// device_vector overloads.  Same fixture: CuckooConfig<uint32_t, 16, 500,
// 256, 16> (DFS default, uint64 words), keys from std::mt19937 seeded 42 with
// uniform_int_distribution<uint32_t>(1, UINT32_MAX).
//
// Build: hipcc --offload-arch=gfx950 -std=c++20 -O2 -I cuckoo-filter_amd/csrc
// Run:   ./test_cuckoo_filter   (exit status 0 iff every check holds)
#include <hipcub/hipcub.hpp>
#include <thrust/copy.h>
#include <thrust/device_vector.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <limits>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "cuckoo/CuckooFilter.hpp"

namespace {

using Config = CuckooConfig<uint32_t, 16, 500, 256, 16>;
using Filter = CuckooFilter<Config>;
static_assert(Config::evictionPolicy == EvictionPolicy::DFS, "DFS is the default eviction policy");

int g_failures = 0;

void check(bool ok, const std::string& what) {
    if (!ok) {
        ++g_failures;
        std::printf("    FAILED: %s\n", what.c_str());
    }
}

std::vector<uint32_t> randomKeys(std::mt19937& rng, size_t n) {
    std::uniform_int_distribution<uint32_t> dist(1, std::numeric_limits<uint32_t>::max());
    std::vector<uint32_t> keys(n);
    for (auto& k : keys) k = dist(rng);
    return keys;
}

std::vector<uint8_t> lookup(Filter& f, const std::vector<uint32_t>& keys) {
    thrust::device_vector<uint32_t> d(keys.begin(), keys.end());
    thrust::device_vector<uint8_t> out(keys.size());
    f.containsMany(d, out);
    std::vector<uint8_t> h(keys.size());
    thrust::copy(out.begin(), out.end(), h.begin());
    return h;
}

size_t insert(Filter& f, const std::vector<uint32_t>& keys) {
    thrust::device_vector<uint32_t> d(keys.begin(), keys.end());
    return f.insertMany(d);
}

size_t removeKeys(Filter& f, const std::vector<uint32_t>& keys, std::vector<uint8_t>* ok) {
    thrust::device_vector<uint32_t> d(keys.begin(), keys.end());
    thrust::device_vector<uint8_t> out(keys.size());
    const size_t remaining = f.deleteMany(d, out);
    if (ok) {
        ok->resize(keys.size());
        thrust::copy(out.begin(), out.end(), ok->begin());
    }
    return remaining;
}

size_t countTrue(const std::vector<uint8_t>& v) {
    return static_cast<size_t>(std::count_if(v.begin(), v.end(), [](uint8_t x) { return x != 0; }));
}

void basicInsertAndQuery(std::mt19937&) {
    Filter f(10000);
    const std::vector<uint32_t> keys = {1, 2, 3, 4, 5, 100, 200, 300};
    check(insert(f, keys) == keys.size(), "insertMany returns the number of keys");
    check(countTrue(lookup(f, keys)) == keys.size(), "every inserted key is found");
}

void emptyFilter(std::mt19937&) {
    Filter f(10000);
    check(countTrue(lookup(f, {1, 2, 3, 4, 5})) == 0, "an empty filter finds nothing");
}

void clearOperation(std::mt19937&) {
    Filter f(10000);
    const std::vector<uint32_t> keys = {1, 2, 3, 4, 5};
    insert(f, keys);
    check(countTrue(lookup(f, keys)) == keys.size(), "found before clear()");
    f.clear();
    check(countTrue(lookup(f, keys)) == 0, "nothing found after clear()");
    check(f.loadFactor() == 0.0f, "load factor 0 after clear()");
}

void loadFactorTest(std::mt19937& rng) {
    Filter f(10000);
    check(f.loadFactor() == 0.0f, "empty load factor");
    const auto keys = randomKeys(rng, 5000);
    const size_t inserted = insert(f, keys);
    const float lf = f.loadFactor();
    check(lf > 0.0f && lf <= 1.0f, "load factor in (0, 1]");
    check(lf == static_cast<float>(inserted) / static_cast<float>(f.capacity()), "load factor = inserted/capacity");
    check(inserted > 0, "some keys inserted");
}

void nearCapacityInsertion(std::mt19937& rng) {
    const size_t n = static_cast<size_t>((1u << 20) * 0.95);
    Filter f(n);
    const auto keys = randomKeys(rng, n);
    const size_t inserted = insert(f, keys);
    check(inserted > n * 0.99, "at least 99% of the keys inserted");
    check(countTrue(lookup(f, keys)) > inserted * 0.99, "at least 99% of the inserted keys found");
}

void duplicateInsertions(std::mt19937&) {
    Filter f(10000);
    const std::vector<uint32_t> keys = {1, 1, 2, 2, 3, 3, 4, 4, 5, 5};
    check(insert(f, keys) == keys.size(), "duplicates occupy separate slots");
    check(countTrue(lookup(f, {1, 2, 3, 4, 5})) == 5, "every distinct key found");
}

void basicDeletion(std::mt19937&) {
    Filter f(10000);
    const std::vector<uint32_t> keys = {1, 2, 3, 4, 5, 100, 200, 300};
    check(insert(f, keys) == keys.size(), "all keys inserted");
    std::vector<uint8_t> ok;
    const size_t remaining = removeKeys(f, keys, &ok);
    check(countTrue(ok) == keys.size(), "every key deleted");
    check(remaining == 0, "no slot occupied after deleting everything");
    check(countTrue(lookup(f, keys)) == 0, "deleted keys not found");
}

void deleteNonExistentKeys(std::mt19937&) {
    Filter f(10000);
    std::vector<uint8_t> ok;
    removeKeys(f, {1, 2, 3, 4, 5}, &ok);
    check(countTrue(ok) == 0, "deleting absent keys succeeds for none");
}

void partialDeletion(std::mt19937&) {
    Filter f(10000);
    check(insert(f, {1, 2, 3, 4, 5, 6, 7, 8, 9, 10}) == 10, "10 keys inserted");
    std::vector<uint8_t> ok;
    const std::vector<uint32_t> evens = {2, 4, 6, 8, 10};
    const size_t remaining = removeKeys(f, evens, &ok);
    check(countTrue(ok) == 5, "every even key deleted");
    check(remaining == 5, "5 slots remain occupied");
    check(countTrue(lookup(f, evens)) == 0, "deleted keys not found");
    check(countTrue(lookup(f, {1, 3, 5, 7, 9})) == 5, "remaining keys still found");
}

void deleteAndReinsert(std::mt19937&) {
    Filter f(10000);
    const std::vector<uint32_t> keys = {1, 2, 3, 4, 5};
    insert(f, keys);
    removeKeys(f, keys, nullptr);
    check(countTrue(lookup(f, keys)) == 0, "gone after delete");
    check(insert(f, keys) == keys.size(), "reinsertion fills the slots again");
    check(countTrue(lookup(f, keys)) == keys.size(), "found after reinsertion");
}

void loadFactorAfterDeletion(std::mt19937& rng) {
    Filter f(10000);
    const auto keys = randomKeys(rng, 5000);
    insert(f, keys);
    const float before = f.loadFactor();
    check(before > 0.0f, "positive load factor after insert");
    const std::vector<uint32_t> half(keys.begin(), keys.begin() + 2500);
    const size_t remaining = removeKeys(f, half, nullptr);
    const float after = f.loadFactor();
    check(after < before, "load factor falls after deleting");
    check(after == static_cast<float>(remaining) / static_cast<float>(f.capacity()), "load = remaining/capacity");
}

void deleteDuplicates(std::mt19937&) {
    Filter f(10000);
    insert(f, {1, 1, 1, 2, 2, 3});
    std::vector<uint8_t> ok;
    removeKeys(f, {1, 1, 1}, &ok);
    check(countTrue(ok) >= 1, "at least one copy deleted");
}

void sortedInsertAndCounters(std::mt19937& rng) {
    // insertManySorted, countOccupiedSlots, capacity/numBuckets/sizeInBytes
    const size_t n = static_cast<size_t>((1u << 20) * 0.95);
    Filter f(n);
    const auto keys = randomKeys(rng, n);
    thrust::device_vector<uint32_t> d(keys.begin(), keys.end());
    const size_t inserted = f.insertManySorted(d);
    check(inserted > n * 0.99, "sorted insert places at least 99% of the keys");
    check(f.countOccupiedSlots() == inserted, "counter equals non-empty slots");
    check(f.capacity() == f.getNumBuckets() * Filter::bucketSize, "capacity = buckets * bucketSize");
    check(f.sizeInBytes() == f.getNumBuckets() * sizeof(Filter::Bucket), "sizeInBytes = buckets * sizeof(Bucket)");
    check(countTrue(lookup(f, keys)) > inserted * 0.99, "sorted-inserted keys found");
}

void hostHashMatchesDevice(std::mt19937& rng) {
    // getCandidateBuckets is __host__ __device__: host results equal device packed tags
    Filter f(1 << 16);
    const auto keys = randomKeys(rng, 4096);
    thrust::device_vector<uint32_t> d(keys.begin(), keys.end());
    thrust::device_vector<Filter::PackedTagType> packed(keys.size());
    computePackedTagsKernel<Config><<<16, 256>>>(thrust::raw_pointer_cast(d.data()),
                                                 thrust::raw_pointer_cast(packed.data()), keys.size(),
                                                 f.getNumBuckets());
    HIP_CALL(hipDeviceSynchronize());
    std::vector<Filter::PackedTagType> h(keys.size());
    thrust::copy(packed.begin(), packed.end(), h.begin());
    size_t same = 0;
    for (size_t i = 0; i < keys.size(); ++i) {
        const auto [i1, i2, fp] = Filter::getCandidateBuckets(keys[i], f.getNumBuckets());
        (void)i2;
        same += Filter::PackedTag(fp, i1).value == h[i];
    }
    check(same == keys.size(), "host getCandidateBuckets == device computePackedTagsKernel");
}

void presortedInsert(std::mt19937& rng) {
    // The reference's InsertPresorted benchmark flow
    // (benchmark/cuckoo_sorted_vs_unsorted.cu:48-86), written as it is there:
    // computePackedTagsKernel, a device radix sort of the packed tags, then
    // insertKernelSorted launched with the filter's address.
    using PackedTagType = Filter::PackedTagType;
    const size_t n = static_cast<size_t>((1u << 20) * 0.95);
    auto filter = std::make_unique<Filter>(n);
    const auto keys = randomKeys(rng, n);
    thrust::device_vector<uint32_t> d_keys(keys.begin(), keys.end());
    PackedTagType* d_packedTags = nullptr;
    HIP_CALL(hipMalloc(&d_packedTags, n * sizeof(PackedTagType)));
    const size_t numBlocks = SDIV(n, Config::blockSize);
    computePackedTagsKernel<Config><<<numBlocks, Config::blockSize>>>(thrust::raw_pointer_cast(d_keys.data()),
                                                                      d_packedTags, n, filter->numBuckets);
    void* d_tempStorage = nullptr;
    size_t tempStorageBytes = 0;
    HIP_CALL(hipcub::DeviceRadixSort::SortKeys(d_tempStorage, tempStorageBytes, d_packedTags, d_packedTags, n));
    HIP_CALL(hipMalloc(&d_tempStorage, tempStorageBytes));
    HIP_CALL(hipcub::DeviceRadixSort::SortKeys(d_tempStorage, tempStorageBytes, d_packedTags, d_packedTags, n));
    HIP_CALL(hipFree(d_tempStorage));
    HIP_CALL(hipDeviceSynchronize());
    std::vector<PackedTagType> h(n);
    HIP_CALL(hipMemcpy(h.data(), d_packedTags, n * sizeof(PackedTagType), hipMemcpyDeviceToHost));
    check(std::is_sorted(h.begin(), h.end()), "packed tags sorted");
    filter->clear();
    insertKernelSorted<Config><<<numBlocks, Config::blockSize>>>(thrust::raw_pointer_cast(d_keys.data()),
                                                                 d_packedTags, n, filter.get());
    HIP_CALL(hipDeviceSynchronize());
    const size_t inserted = filter->occupiedSlots();
    check(inserted > n * 0.99, "presorted insert places at least 99% of the keys");
    check(filter->countOccupiedSlots() == inserted, "counter equals non-empty slots");
    check(countTrue(lookup(*filter, keys)) >= inserted, "presorted-inserted keys found");
    // the variant for workgroups of up to 256 threads, launched the same way
    filter->clear();
    insertKernelSortedSmall<Config><<<numBlocks, Config::blockSize>>>(d_packedTags, n, filter.get());
    HIP_CALL(hipDeviceSynchronize());
    const size_t insertedSmall = filter->occupiedSlots();
    check(insertedSmall > n * 0.99, "small-block presorted insert places at least 99% of the keys");
    check(filter->countOccupiedSlots() == insertedSmall, "counter equals non-empty slots (small-block kernel)");
    check(countTrue(lookup(*filter, keys)) >= insertedSmall, "small-block presorted-inserted keys found");
    HIP_CALL(hipFree(d_packedTags));
}

}  // namespace

int main() {
    const std::vector<std::pair<const char*, std::function<void(std::mt19937&)>>> tests = {
        {"BasicInsertAndQuery", basicInsertAndQuery},
        {"EmptyFilter", emptyFilter},
        {"ClearOperation", clearOperation},
        {"LoadFactor", loadFactorTest},
        {"NearCapacityInsertion", nearCapacityInsertion},
        {"DuplicateInsertions", duplicateInsertions},
        {"BasicDeletion", basicDeletion},
        {"DeleteNonExistentKeys", deleteNonExistentKeys},
        {"PartialDeletion", partialDeletion},
        {"DeleteAndReinsert", deleteAndReinsert},
        {"LoadFactorAfterDeletion", loadFactorAfterDeletion},
        {"DeleteDuplicates", deleteDuplicates},
        {"SortedInsertAndCounters", sortedInsertAndCounters},
        {"HostHashMatchesDevice", hostHashMatchesDevice},
        {"PresortedInsert", presortedInsert},
    };
    for (const auto& [name, fn] : tests) {
        std::mt19937 rng(42);
        const int before = g_failures;
        fn(rng);
        std::printf("[%s] %s\n", g_failures == before ? "  OK  " : "FAILED", name);
        std::fflush(stdout);
    }
    std::printf("%d check(s) failed\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
