// C ABI of the cuckoo filter (include/cuckoo_filter.h) over the per-config
// kernel tables built by capi_inst.hip.
#include "../../include/cuckoo_filter.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "capi_internal.hpp"

extern "C" int cfk_encode_kmers_device(const void* d_seq, size_t len, uint32_t k, int canonical, uint64_t* d_out,
                                       void* stream);

using namespace cfabi::internal;

namespace {

uint32_t tagBytes(uint32_t bits) { return bits <= 8 ? 1u : (bits <= 16 ? 2u : 4u); }

bool isPow2(uint64_t v) { return v != 0 && (v & (v - 1)) == 0; }

// Mirrors the reference's static_asserts (CuckooFilter.cuh:67-71, 147-152, 226-236).
int validate(const cf_config* c) {
    if (c == nullptr) return fail(CF_ERR_INVALID, "config is NULL");
    if (c->key_bytes != 4 && c->key_bytes != 8) return fail(CF_ERR_INVALID, "key_bytes must be 4 or 8");
    if (c->bits_per_tag != 8 && c->bits_per_tag != 16 && c->bits_per_tag != 32)
        return fail(CF_ERR_INVALID, "The tag must be 8, 16 or 32 bits");
    if (c->word_bytes != 4 && c->word_bytes != 8) return fail(CF_ERR_INVALID, "WordType must be uint32_t or uint64_t");
    if (tagBytes(c->bits_per_tag) > c->word_bytes) return fail(CF_ERR_INVALID, "TagType must fit within WordType");
    if (!isPow2(c->bucket_size)) return fail(CF_ERR_INVALID, "Bucket size must be a power of 2");
    if (c->bucket_size % (c->word_bytes / tagBytes(c->bits_per_tag)) != 0)
        return fail(CF_ERR_INVALID, "bucketSize must be divisible by tagsPerWord");
    if (c->policy > CF_POLICY_OFFSET) return fail(CF_ERR_INVALID, "unknown alternate-bucket policy %u", c->policy);
    if (c->eviction > CF_EVICT_DFS) return fail(CF_ERR_INVALID, "unknown eviction policy %u", c->eviction);
    if (c->block_size < 1 || c->block_size > 1024) return fail(CF_ERR_INVALID, "block_size must be in [1, 1024]");
    return CF_OK;
}

const cfabi::Ops* findOps(const cf_config* c) {
    static const std::vector<const cfabi::Ops*> all = [] {
        std::vector<const cfabi::Ops*> v;
        auto add = [&](const cfabi::Ops* t, size_t n) {
            for (size_t i = 0; i < n; ++i) v.push_back(&t[i]);
        };
        size_t n = 0;
        const cfabi::Ops* t;
        t = cfabi_table_0_4(&n); add(t, n);
        t = cfabi_table_0_8(&n); add(t, n);
        t = cfabi_table_1_4(&n); add(t, n);
        t = cfabi_table_1_8(&n); add(t, n);
        t = cfabi_table_2_4(&n); add(t, n);
        t = cfabi_table_2_8(&n); add(t, n);
        return v;
    }();
    for (const cfabi::Ops* o : all) {
        if (o->policy == c->policy && o->keyBytes == c->key_bytes && o->bitsPerTag == c->bits_per_tag &&
            o->bucketSize == c->bucket_size && o->wordBytes == c->word_bytes) {
            return o;
        }
    }
    return nullptr;
}

/// Workspace of the binned insert, shared by every filter on a device: calls
/// are synchronous, so a call owns it from launch to return.  Memory held is
/// the largest call's need, not a sum over filters.
struct SharedWorkspace {
    std::mutex mu;
    std::map<int, cuckoo::Workspace> perDevice;
};

SharedWorkspace& sharedWorkspace() {
    static SharedWorkspace* ws = new SharedWorkspace();  // lives until process exit
    return *ws;
}

}  // namespace

// ---- scratch of cf_insert_many_noalloc ------------------------------------
// A file-scope device array: every device the library runs on holds its own
// copy from the module's load until the process ends.  Host code never uses
// the array's name as an address (on the host it only stands in for the
// symbol); hipGetSymbolAddress gives each device's copy, looked up once per
// device on that device's first no-allocation call, and only that address
// reaches the kernels.
static __device__ __align__(256) unsigned char scratch[CF_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

struct NoAllocScratch {
    std::mutex mu;  // held by a call from launch to completion: calls on one process serialize
    void* ptr[kMaxDevices] = {};
};

NoAllocScratch& noAllocScratch() {
    static NoAllocScratch s;
    return s;
}

int readOccupied(cf_filter* f, hipStream_t s, size_t* out) {
    if (out == nullptr) return CF_OK;
    unsigned long long v = 0;
    hipError_t e = hipMemcpyAsync(&v, f->counters, sizeof(v), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hipFail(e, "reading the occupancy counter");
    *out = static_cast<size_t>(v);
    return CF_OK;
}

int checkFilter(const cf_filter* f) {
    if (f == nullptr || f->ops == nullptr) return fail(CF_ERR_INVALID, "filter is NULL");
    return CF_OK;
}

int checkKeys(const void* keys, size_t n) {
    if (n != 0 && keys == nullptr) return fail(CF_ERR_INVALID, "key pointer is NULL");
    return CF_OK;
}

}  // namespace

extern "C" {

const char* cf_last_error(void) { return g_lastError.c_str(); }

int cf_config_init(cf_config* cfg, uint32_t key_bytes, uint32_t bits_per_tag) {
    if (cfg == nullptr) return fail(CF_ERR_INVALID, "config is NULL");
    cfg->key_bytes = key_bytes;
    cfg->bits_per_tag = bits_per_tag;
    cfg->max_evictions = 500;
    cfg->block_size = 256;
    cfg->bucket_size = 16;
    cfg->policy = CF_POLICY_XOR;
    cfg->eviction = CF_EVICT_DFS;
    cfg->word_bytes = 8;
    return validate(cfg);
}

int cf_config_supported(const cf_config* cfg) {
    const int r = validate(cfg);
    if (r != CF_OK) return r;
    return findOps(cfg) != nullptr ? 1 : 0;
}

size_t cf_calculate_num_buckets(const cf_config* cfg, size_t capacity) {
    if (validate(cfg) != CF_OK) return 0;
    const cfabi::Ops* o = findOps(cfg);
    return o ? o->numBuckets(capacity) : 0;
}

int cf_candidate_buckets_host(const cf_config* cfg, size_t num_buckets, const void* h_keys, size_t n,
                              uint64_t* h_i1, uint64_t* h_i2, uint32_t* h_fp) {
    int r = validate(cfg);
    if (r != CF_OK) return r;
    const cfabi::Ops* o = findOps(cfg);
    if (o == nullptr) return fail(CF_ERR_UNSUPPORTED, "configuration not built into this library");
    if (num_buckets == 0) return fail(CF_ERR_INVALID, "num_buckets must be positive");
    if (n != 0 && (h_keys == nullptr || h_i1 == nullptr || h_i2 == nullptr || h_fp == nullptr))
        return fail(CF_ERR_INVALID, "NULL buffer");
    o->hostCandidates(num_buckets, h_keys, n, h_i1, h_i2, h_fp);
    return CF_OK;
}

int cf_create(const cf_config* cfg, size_t capacity, cf_filter** out) {
    if (out == nullptr) return fail(CF_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int r = validate(cfg);
    if (r != CF_OK) return r;
    const cfabi::Ops* ops = findOps(cfg);
    if (ops == nullptr) return fail(CF_ERR_UNSUPPORTED, "configuration not built into this library");
    const size_t nb = ops->numBuckets(capacity);
    if (nb == 0) return fail(CF_ERR_INVALID, "capacity %zu gives no buckets", capacity);

    cf_filter* f = new (std::nothrow) cf_filter();
    if (f == nullptr) return fail(CF_ERR_ALLOC, "host allocation failed");
    f->cfg = *cfg;
    f->ops = ops;
    f->numBuckets = nb;
    if (hipGetDevice(&f->device) != hipSuccess) {
        delete f;
        return fail(CF_ERR_HIP, "no HIP device");
    }
    hipError_t e = hipMalloc(&f->buckets, f->tableBytes());
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&f->counters), 2 * sizeof(unsigned long long));
    if (e != hipSuccess) {
        if (f->buckets) (void)hipFree(f->buckets);
        delete f;
        return fail(CF_ERR_ALLOC, "hipMalloc of %zu table bytes: %s", nb * ops->bucketBytes, hipGetErrorString(e));
    }
    r = cf_clear(f);
    if (r != CF_OK) {
        cf_destroy(f);
        return r;
    }
    *out = f;
    return CF_OK;
}

int cf_destroy(cf_filter* f) {
    if (f == nullptr) return CF_OK;
    DeviceGuard g(f->device);
    hipError_t e = hipSuccess;
    if (f->buckets) e = hipFree(f->buckets);
    if (f->counters) (void)hipFree(f->counters);
    if (f->hist) (void)hipFree(f->hist);
    delete f;
    return e == hipSuccess ? CF_OK : hipFail(e, "hipFree");
}

int cf_insert_many(cf_filter* f, const void* d_keys, size_t n, void* stream, size_t* occupied) {
    int r = checkFilter(f);
    if (r == CF_OK) r = checkKeys(d_keys, n);
    if (r != CF_OK) return r;
    DeviceGuard g(f->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e;
    {
        SharedWorkspace& sw = sharedWorkspace();
        std::lock_guard<std::mutex> lock(sw.mu);
        e = f->ops->insert(f->view(), d_keys, n, f->cfg.block_size, s, sw.perDevice[f->device]);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) return hipFail(e, "insertMany");
    return readOccupied(f, s, occupied);
}

int cf_insert_many_noalloc(cf_filter* f, const void* d_keys, size_t n, void* stream, size_t* occupied) {
    int r = checkFilter(f);
    if (r == CF_OK) r = checkKeys(d_keys, n);
    if (r != CF_OK) return r;
    if (n > CF_NOALLOC_MAX_KEYS)
        return fail(CF_ERR_INVALID, "cf_insert_many_noalloc takes at most %zu keys per call, got %zu",
                    static_cast<size_t>(CF_NOALLOC_MAX_KEYS), n);
    if (f->device < 0 || f->device >= kMaxDevices) return fail(CF_ERR_INVALID, "device %d out of range", f->device);
    const size_t need = f->ops->insertWorkspaceBytes(n, f->numBuckets);
    if (need > CF_NOALLOC_SCRATCH_BYTES)
        return fail(CF_ERR_INVALID, "this insert needs %zu bytes of scratch; cf_insert_many_noalloc has %zu", need,
                    static_cast<size_t>(CF_NOALLOC_SCRATCH_BYTES));
    DeviceGuard g(f->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    NoAllocScratch& ns = noAllocScratch();
    hipError_t e;
    {
        std::lock_guard<std::mutex> lock(ns.mu);
        if (ns.ptr[f->device] == nullptr) {
            void* p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return hipFail(e, "hipGetSymbolAddress(scratch)");
            ns.ptr[f->device] = p;
        }
        cuckoo::Workspace ws;
        ws.ptr = ns.ptr[f->device];
        ws.bytes = CF_NOALLOC_SCRATCH_BYTES;
        ws.borrowed = true;
        e = f->ops->insert(f->view(), d_keys, n, f->cfg.block_size, s, ws);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) return hipFail(e, "insertMany (no allocation)");
    return readOccupied(f, s, occupied);
}

int cf_insert_many_sorted(cf_filter* f, const void* d_keys, size_t n, void* stream, size_t* occupied) {
    int r = checkFilter(f);
    if (r == CF_OK) r = checkKeys(d_keys, n);
    if (r != CF_OK) return r;
    DeviceGuard g(f->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e;
    {
        SharedWorkspace& sw = sharedWorkspace();
        std::lock_guard<std::mutex> lock(sw.mu);
        e = f->ops->insertSorted(f->view(), d_keys, n, f->cfg.block_size, s, sw.perDevice[f->device]);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) return hipFail(e, "insertManySorted");
    return readOccupied(f, s, occupied);
}

int cf_contains_many(cf_filter* f, const void* d_keys, size_t n, bool* d_output, void* stream) {
    int r = checkFilter(f);
    if (r == CF_OK) r = checkKeys(d_keys, n);
    if (r != CF_OK) return r;
    if (n != 0 && d_output == nullptr) return fail(CF_ERR_INVALID, "output pointer is NULL");
    DeviceGuard g(f->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = f->ops->contains(f->view(), d_keys, n, d_output, f->cfg.block_size, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? CF_OK : hipFail(e, "containsMany");
}

int cf_delete_many(cf_filter* f, const void* d_keys, size_t n, bool* d_output, void* stream, size_t* occupied) {
    int r = checkFilter(f);
    if (r == CF_OK) r = checkKeys(d_keys, n);
    if (r != CF_OK) return r;
    DeviceGuard g(f->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e;
    {
        SharedWorkspace& sw = sharedWorkspace();
        std::lock_guard<std::mutex> lock(sw.mu);
        e = f->ops->remove(f->view(), d_keys, n, d_output, f->cfg.block_size, s, sw.perDevice[f->device]);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) return hipFail(e, "deleteMany");
    return readOccupied(f, s, occupied);
}

int cf_clear(cf_filter* f) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    DeviceGuard g(f->device);
    hipError_t e = hipMemset(f->buckets, 0, f->tableBytes());
    if (e == hipSuccess) e = hipMemset(f->counters, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? CF_OK : hipFail(e, "clear");
}

int cf_occupied_slots(cf_filter* f, size_t* occupied) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    if (occupied == nullptr) return fail(CF_ERR_INVALID, "out is NULL");
    DeviceGuard g(f->device);
    unsigned long long v = 0;
    hipError_t e = hipMemcpy(&v, f->counters, sizeof(v), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipFail(e, "occupiedSlots");
    *occupied = static_cast<size_t>(v);
    return CF_OK;
}

int cf_load_factor(cf_filter* f, float* load_factor) {
    if (load_factor == nullptr) return fail(CF_ERR_INVALID, "out is NULL");
    size_t occ = 0;
    int r = cf_occupied_slots(f, &occ);
    if (r != CF_OK) return r;
    *load_factor = static_cast<float>(occ) / static_cast<float>(f->numBuckets * f->cfg.bucket_size);
    return CF_OK;
}

size_t cf_capacity(const cf_filter* f) { return f ? f->numBuckets * f->cfg.bucket_size : 0; }
size_t cf_num_buckets(const cf_filter* f) { return f ? f->numBuckets : 0; }
size_t cf_size_in_bytes(const cf_filter* f) { return f ? f->tableBytes() : 0; }

int cf_count_occupied_slots(cf_filter* f, size_t* count) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    if (count == nullptr) return fail(CF_ERR_INVALID, "out is NULL");
    std::vector<unsigned char> img;
    try {
        img.resize(f->tableBytes());
    } catch (...) {
        return fail(CF_ERR_ALLOC, "host allocation of %zu bytes failed", f->tableBytes());
    }
    r = cf_export_table(f, img.data(), img.size());
    if (r != CF_OK) return r;
    const uint32_t tb = tagBytes(f->cfg.bits_per_tag);
    size_t c = 0;
    // slots are equal-width little-endian lanes, so the image is a flat tag array
    for (size_t off = 0; off < img.size(); off += tb) {
        uint32_t t = 0;
        std::memcpy(&t, img.data() + off, tb);
        c += t != 0;
    }
    *count = c;
    return CF_OK;
}

int cf_eviction_count(cf_filter* f, size_t* count) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    if (count == nullptr) return fail(CF_ERR_INVALID, "out is NULL");
    DeviceGuard g(f->device);
    unsigned long long v = 0;
    hipError_t e = hipMemcpy(&v, f->counters + 1, sizeof(v), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipFail(e, "evictionCount");
    *count = static_cast<size_t>(v);
    return CF_OK;
}

int cf_reset_eviction_count(cf_filter* f) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    DeviceGuard g(f->device);
    hipError_t e = hipMemset(f->counters + 1, 0, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? CF_OK : hipFail(e, "resetEvictionCount");
}

int cf_set_eviction_histogram(cf_filter* f, int enable) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    DeviceGuard g(f->device);
    if (!enable) {
        if (f->hist) {
            hipError_t e = hipFree(f->hist);
            f->hist = nullptr;
            if (e != hipSuccess) return hipFail(e, "hipFree");
        }
        return CF_OK;
    }
    const size_t bins = static_cast<size_t>(f->cfg.max_evictions) + 2;
    hipError_t e = hipSuccess;
    if (f->hist == nullptr) e = hipMalloc(reinterpret_cast<void**>(&f->hist), bins * sizeof(unsigned int));
    if (e == hipSuccess) e = hipMemset(f->hist, 0, bins * sizeof(unsigned int));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? CF_OK : hipFail(e, "eviction histogram");
}

int cf_eviction_histogram(cf_filter* f, uint64_t* h_bins, size_t nbins) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    const size_t bins = static_cast<size_t>(f->cfg.max_evictions) + 2;
    if (f->hist == nullptr) return fail(CF_ERR_INVALID, "eviction histogram is not enabled");
    if (h_bins == nullptr || nbins < bins) return fail(CF_ERR_INVALID, "need %zu bins", bins);
    DeviceGuard g(f->device);
    std::vector<unsigned int> tmp(bins);
    hipError_t e = hipMemcpy(tmp.data(), f->hist, bins * sizeof(unsigned int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipFail(e, "eviction histogram");
    for (size_t i = 0; i < nbins; ++i) h_bins[i] = i < bins ? tmp[i] : 0;
    return CF_OK;
}

int cf_export_table(cf_filter* f, void* h_dst, size_t bytes) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    if (h_dst == nullptr || bytes != f->tableBytes())
        return fail(CF_ERR_INVALID, "table image is %zu bytes", f->tableBytes());
    DeviceGuard g(f->device);
    hipError_t e = hipMemcpy(h_dst, f->buckets, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? CF_OK : hipFail(e, "export table");
}

int cf_import_table(cf_filter* f, const void* h_src, size_t bytes, size_t occupied) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    if (h_src == nullptr || bytes != f->tableBytes())
        return fail(CF_ERR_INVALID, "table image is %zu bytes", f->tableBytes());
    DeviceGuard g(f->device);
    const unsigned long long occ = occupied;
    hipError_t e = hipMemcpy(f->buckets, h_src, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(f->counters, &occ, sizeof(occ), hipMemcpyHostToDevice);
    return e == hipSuccess ? CF_OK : hipFail(e, "import table");
}

int cf_copy_filter(cf_filter* dst, cf_filter* src, void* stream) {
    int r = checkFilter(dst);
    if (r == CF_OK) r = checkFilter(src);
    if (r != CF_OK) return r;
    if (dst->ops != src->ops || dst->numBuckets != src->numBuckets || dst->device != src->device)
        return fail(CF_ERR_INVALID, "filters differ in configuration, size or device");
    DeviceGuard g(dst->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemcpyAsync(dst->buckets, src->buckets, dst->tableBytes(), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dst->counters, src->counters, 2 * sizeof(unsigned long long), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? CF_OK : hipFail(e, "copy filter");
}

int cf_compute_candidates(cf_filter* f, const void* d_keys, size_t n, uint64_t* d_i1, uint64_t* d_i2,
                          uint32_t* d_fp, void* stream) {
    int r = checkFilter(f);
    if (r == CF_OK) r = checkKeys(d_keys, n);
    if (r != CF_OK) return r;
    if (n != 0 && (d_i1 == nullptr || d_i2 == nullptr || d_fp == nullptr)) return fail(CF_ERR_INVALID, "NULL output");
    DeviceGuard g(f->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = f->ops->candidates(d_keys, n, f->numBuckets, d_i1, d_i2, d_fp, f->cfg.block_size, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? CF_OK : hipFail(e, "compute candidates");
}

int cf_compute_packed_tags(cf_filter* f, const void* d_keys, size_t n, void* d_packed, void* stream) {
    int r = checkFilter(f);
    if (r == CF_OK) r = checkKeys(d_keys, n);
    if (r != CF_OK) return r;
    if (n != 0 && d_packed == nullptr) return fail(CF_ERR_INVALID, "NULL output");
    DeviceGuard g(f->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = f->ops->packedTags(d_keys, n, f->numBuckets, d_packed, f->cfg.block_size, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? CF_OK : hipFail(e, "compute packed tags");
}

int cf_sort_packed_tags(cf_filter* f, void* d_packed, size_t n, void* stream) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    if (n != 0 && d_packed == nullptr) return fail(CF_ERR_INVALID, "packed-tag pointer is NULL");
    if (n >= (size_t(1) << 32)) return fail(CF_ERR_INVALID, "at most 2^32 - 1 packed tags per sort");
    DeviceGuard g(f->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e;
    {
        SharedWorkspace& sw = sharedWorkspace();
        std::lock_guard<std::mutex> lock(sw.mu);
        e = f->ops->sortPacked(d_packed, n, f->numBuckets, s, sw.perDevice[f->device]);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    return e == hipSuccess ? CF_OK : hipFail(e, "sort packed tags");
}

int cf_insert_packed_sorted(cf_filter* f, const void* d_packed, size_t n, void* stream, size_t* occupied) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    if (n != 0 && d_packed == nullptr) return fail(CF_ERR_INVALID, "packed-tag pointer is NULL");
    DeviceGuard g(f->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = f->ops->insertPacked(f->view(), d_packed, n, f->cfg.block_size, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hipFail(e, "insertKernelSorted");
    return readOccupied(f, s, occupied);
}

}  // extern "C"

struct cf_batch {
    std::vector<std::pair<const cfabi::Ops*, void*>> executors;  // one per configuration used
    std::vector<cuckoo::BatchOpDesc> descs;
};

extern "C" {

int cf_batch_create(cf_batch** out) {
    if (out == nullptr) return fail(CF_ERR_INVALID, "out is NULL");
    *out = new (std::nothrow) cf_batch();
    return *out != nullptr ? CF_OK : fail(CF_ERR_ALLOC, "host allocation failed");
}

int cf_batch_destroy(cf_batch* b) {
    if (b == nullptr) return CF_OK;
    for (auto& [ops, state] : b->executors) {
        if (state != nullptr) ops->batchRelease(state);
    }
    delete b;
    return CF_OK;
}

int cf_batch_execute(cf_batch* b, const cf_batch_op* ops, size_t count, void* stream, uint64_t* h_occupied) {
    if (b == nullptr) return fail(CF_ERR_INVALID, "batch is NULL");
    if (count == 0) return CF_OK;
    if (ops == nullptr) return fail(CF_ERR_INVALID, "ops is NULL");
    const cfabi::Ops* kind = nullptr;
    int device = -1;
    b->descs.resize(count);
    for (size_t i = 0; i < count; ++i) {
        const cf_batch_op& o = ops[i];
        int r = checkFilter(o.filter);
        if (r != CF_OK) return r;
        if (kind == nullptr) {
            kind = o.filter->ops;
            device = o.filter->device;
        } else if (o.filter->ops != kind) {
            return fail(CF_ERR_INVALID, "op %zu: every filter of a batch must share one configuration", i);
        } else if (o.filter->device != device) {
            return fail(CF_ERR_INVALID, "op %zu: every filter of a batch must live on one device", i);
        }
        if (o.op > CF_OP_DELETE) return fail(CF_ERR_INVALID, "op %zu: unknown operation %u", i, o.op);
        if (o.n != 0 && o.d_keys == nullptr) return fail(CF_ERR_INVALID, "op %zu: key pointer is NULL", i);
        if (o.op == CF_OP_LOOKUP && o.n != 0 && o.d_output == nullptr)
            return fail(CF_ERR_INVALID, "op %zu: lookup needs an output", i);
        b->descs[i] = describe(o.filter, o.op, o.d_keys, o.n, o.d_output);
    }
    void** state = nullptr;
    for (auto& [k, st] : b->executors) {
        if (k == kind) state = &st;
    }
    if (state == nullptr) {
        b->executors.emplace_back(kind, nullptr);
        state = &b->executors.back().second;
    }
    DeviceGuard g(device);
    const hipError_t e = kind->batchExecute(state, b->descs.data(), count, static_cast<hipStream_t>(stream), h_occupied);
    return e == hipSuccess ? CF_OK : hipFail(e, "batch execute");
}

int cf_run_plain_calls(const cf_batch_op* ops, size_t count, void* stream, uint64_t* h_occupied) {
    if (count != 0 && ops == nullptr) return fail(CF_ERR_INVALID, "ops is NULL");
    for (size_t i = 0; i < count; ++i) {
        const cf_batch_op& o = ops[i];
        size_t occ = 0;
        int r;
        switch (o.op) {
            case CF_OP_INSERT:
                r = cf_insert_many(o.filter, o.d_keys, o.n, stream, &occ);
                break;
            case CF_OP_LOOKUP:
                r = cf_contains_many(o.filter, o.d_keys, o.n, o.d_output, stream);
                if (r == CF_OK) r = cf_occupied_slots(o.filter, &occ);
                break;
            case CF_OP_DELETE:
                r = cf_delete_many(o.filter, o.d_keys, o.n, o.d_output, stream, &occ);
                break;
            default:
                return fail(CF_ERR_INVALID, "op %zu: unknown operation %u", i, o.op);
        }
        if (r != CF_OK) return r;
        if (h_occupied != nullptr) h_occupied[i] = occ;
    }
    return CF_OK;
}

static_assert(sizeof(cf_table_header) == 80, "table file header layout");

int cf_save_table(cf_filter* f, const char* path) {
    int r = checkFilter(f);
    if (r != CF_OK) return r;
    if (path == nullptr) return fail(CF_ERR_INVALID, "path is NULL");
    cf_table_header h{};
    std::memcpy(h.magic, "CFAMDTB1", 8);
    h.header_bytes = sizeof(cf_table_header);
    h.config = f->cfg;
    h.num_buckets = f->numBuckets;
    DeviceGuard g(f->device);
    unsigned long long counters[2] = {0, 0};
    hipError_t e = hipMemcpy(counters, f->counters, sizeof(counters), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hipFail(e, "reading counters");
    h.occupied = counters[0];
    h.evictions = counters[1];
    h.image_bytes = f->tableBytes();
    std::vector<unsigned char> img;
    try {
        img.resize(f->tableBytes());
    } catch (...) {
        return fail(CF_ERR_ALLOC, "host allocation of %zu bytes failed", f->tableBytes());
    }
    r = cf_export_table(f, img.data(), img.size());
    if (r != CF_OK) return r;
    std::unique_ptr<FILE, int (*)(FILE*)> fp(std::fopen(path, "wb"), &std::fclose);
    if (!fp) return fail(CF_ERR_INVALID, "cannot open %s for writing", path);
    if (std::fwrite(&h, sizeof(h), 1, fp.get()) != 1 || std::fwrite(img.data(), 1, img.size(), fp.get()) != img.size())
        return fail(CF_ERR_INVALID, "short write to %s", path);
    return CF_OK;
}

int cf_load_table(const char* path, cf_filter** out) {
    if (out == nullptr || path == nullptr) return fail(CF_ERR_INVALID, "NULL argument");
    *out = nullptr;
    std::unique_ptr<FILE, int (*)(FILE*)> fp(std::fopen(path, "rb"), &std::fclose);
    if (!fp) return fail(CF_ERR_INVALID, "cannot open %s", path);
    cf_table_header h{};
    if (std::fread(&h, sizeof(h), 1, fp.get()) != 1 || std::memcmp(h.magic, "CFAMDTB1", 8) != 0 ||
        h.header_bytes != sizeof(cf_table_header))
        return fail(CF_ERR_INVALID, "%s is not a cuckoo filter table file", path);
    int r = validate(&h.config);
    if (r != CF_OK) return r;
    const cfabi::Ops* ops = findOps(&h.config);
    if (ops == nullptr) return fail(CF_ERR_UNSUPPORTED, "configuration not built into this library");
    // capacity that reproduces num_buckets under the configuration's policy
    const size_t capacity = static_cast<size_t>(h.num_buckets) * h.config.bucket_size;
    if (ops->numBuckets(capacity) != h.num_buckets)
        return fail(CF_ERR_INVALID, "bucket count %llu is not one this policy produces",
                    static_cast<unsigned long long>(h.num_buckets));
    if (h.image_bytes != h.num_buckets * ops->bucketBytes)
        return fail(CF_ERR_INVALID, "%s: image size does not match its header", path);
    cf_filter* f = nullptr;
    r = cf_create(&h.config, capacity, &f);
    if (r != CF_OK) return r;
    std::vector<unsigned char> img;
    try {
        img.resize(f->tableBytes());
    } catch (...) {
        cf_destroy(f);
        return fail(CF_ERR_ALLOC, "host allocation failed");
    }
    if (std::fread(img.data(), 1, img.size(), fp.get()) != img.size()) {
        cf_destroy(f);
        return fail(CF_ERR_INVALID, "%s: table image is truncated", path);
    }
    r = cf_import_table(f, img.data(), img.size(), static_cast<size_t>(h.occupied));
    if (r == CF_OK) {
        DeviceGuard g(f->device);
        const unsigned long long ev = h.evictions;
        const hipError_t e = hipMemcpy(f->counters + 1, &ev, sizeof(ev), hipMemcpyHostToDevice);
        if (e != hipSuccess) r = hipFail(e, "restoring the eviction counter");
    }
    if (r != CF_OK) {
        cf_destroy(f);
        return r;
    }
    *out = f;
    return CF_OK;
}

int cf_kmer_file_write(const char* path, const uint64_t* h_keys, size_t n) {
    if (path == nullptr || (n != 0 && h_keys == nullptr)) return fail(CF_ERR_INVALID, "NULL argument");
    std::unique_ptr<FILE, int (*)(FILE*)> fp(std::fopen(path, "wb"), &std::fclose);
    if (!fp) return fail(CF_ERR_INVALID, "cannot open %s for writing", path);
    const uint64_t count = n;
    if (std::fwrite(&count, sizeof(count), 1, fp.get()) != 1 ||
        (n != 0 && std::fwrite(h_keys, sizeof(uint64_t), n, fp.get()) != n))
        return fail(CF_ERR_INVALID, "short write to %s", path);
    return CF_OK;
}

int cf_kmer_file_count(const char* path, size_t* n) {
    if (path == nullptr || n == nullptr) return fail(CF_ERR_INVALID, "NULL argument");
    std::unique_ptr<FILE, int (*)(FILE*)> fp(std::fopen(path, "rb"), &std::fclose);
    if (!fp) return fail(CF_ERR_INVALID, "cannot open %s", path);
    uint64_t count = 0;
    if (std::fread(&count, sizeof(count), 1, fp.get()) != 1) return fail(CF_ERR_INVALID, "%s: no k-mer count", path);
    *n = static_cast<size_t>(count);
    return CF_OK;
}

int cf_kmer_file_read(const char* path, uint64_t* h_keys, size_t n) {
    if (path == nullptr || (n != 0 && h_keys == nullptr)) return fail(CF_ERR_INVALID, "NULL argument");
    std::unique_ptr<FILE, int (*)(FILE*)> fp(std::fopen(path, "rb"), &std::fclose);
    if (!fp) return fail(CF_ERR_INVALID, "cannot open %s", path);
    uint64_t count = 0;
    if (std::fread(&count, sizeof(count), 1, fp.get()) != 1) return fail(CF_ERR_INVALID, "%s: no k-mer count", path);
    if (count != n) return fail(CF_ERR_INVALID, "%s holds %llu k-mers, not %zu", path,
                                static_cast<unsigned long long>(count), n);
    if (n != 0 && std::fread(h_keys, sizeof(uint64_t), n, fp.get()) != n)
        return fail(CF_ERR_INVALID, "%s: k-mer list is truncated", path);
    return CF_OK;
}

int cf_encode_kmers(const void* d_seq, size_t len, uint32_t k, int canonical, uint64_t* d_out, void* stream) {
    if (d_seq == nullptr || d_out == nullptr) return fail(CF_ERR_INVALID, "NULL argument");
    if (k == 0 || k > 32) return fail(CF_ERR_INVALID, "k must be in [1, 32]");
    if (len < k) return fail(CF_ERR_INVALID, "sequence shorter than k");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int rc = cfk_encode_kmers_device(d_seq, len, k, canonical, d_out, stream);
    if (rc != 0) return hipFail(static_cast<hipError_t>(rc), "encode k-mers");
    const hipError_t e = hipStreamSynchronize(s);
    return e == hipSuccess ? CF_OK : hipFail(e, "encode k-mers");
}

}  // extern "C"
