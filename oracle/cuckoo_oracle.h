/*
 * cuckoo_oracle.h -- CPU oracle for the cuckoo-filter hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (tiltedtoast/cuckoo-filter, include/CuckooFilter.cuh,
 * include/bucket_policies.cuh, include/hashutil.cuh, include/helpers.cuh).
 * It is linked by the tests, by __graft_entry__.smoke() and by the CPU
 * baseline leg of bench.py, and only ever as the checker; the product library
 * never calls it.
 *
 * Every configuration axis of the reference's compile-time CuckooConfig is a
 * run-time field here, so one build covers all of them.  Sequential calls
 * (threads == 1) are deterministic: for the same keys in the same order they
 * give one table image, one failure count and one eviction histogram.
 *
 * Parity pinning (what checks each part; see tests/test_oracle_*.py):
 *   cfo_xxh64              -- python-xxhash (independent implementation) and
 *                             the published XXH64 test vectors.
 *   cfo_candidates / alt   -- golden triples generated from python-xxhash by
 *                             tests/golden/gen_golden.py; involution and range
 *                             invariants.
 *   bucket scan / insert / -- brute force on tiny tables (all slot orders
 *   remove / BFS / DFS        enumerated by hand-built cases), occupancy and
 *                             "every stored tag belongs to a key" invariants,
 *                             closed-form FPR 2*b*alpha/2^f.
 *   sorted insert          -- same invariants; order = ascending packed tag.
 */
#ifndef CUCKOO_ORACLE_H
#define CUCKOO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { CFO_POLICY_XOR = 0, CFO_POLICY_ADDSUB = 1, CFO_POLICY_OFFSET = 2 };
enum { CFO_EVICT_BFS = 0, CFO_EVICT_DFS = 1 };

/* Same field order as cf_config in include/cuckoo_filter.h. */
typedef struct {
    uint32_t key_bytes;     /* sizeof(KeyType): 4 or 8 */
    uint32_t bits_per_tag;  /* 8, 16 or 32 */
    uint32_t max_evictions; /* DFS chain bound */
    uint32_t block_size;    /* unused by the oracle */
    uint32_t bucket_size;   /* slots per bucket, power of two */
    uint32_t policy;        /* CFO_POLICY_* */
    uint32_t eviction;      /* CFO_EVICT_* */
    uint32_t word_bytes;    /* 4 or 8 */
} cfo_config;

typedef struct cfo_filter cfo_filter;

/* XXH64 of len bytes (hashutil.cuh:70-161). */
uint64_t cfo_xxh64(const void* data, size_t len, uint64_t seed);

/* helpers.cuh:31-45 */
uint64_t cfo_next_pow2(uint64_t n);

/* 0 on a valid config, else a negative code (mirrors the reference's static_asserts). */
int cfo_config_check(const cfo_config* cfg);

/* bucket_policies.cuh:80-83, 172-181, 347-350 */
uint64_t cfo_num_buckets(const cfo_config* cfg, uint64_t capacity);

/* Per-key (i1, i2, fp) -- bucket_policies.cuh:45-61, 126-143, 263-285. */
void cfo_candidates(const cfo_config* cfg, uint64_t nb, const void* keys, size_t n,
                    uint64_t* i1, uint64_t* i2, uint32_t* fp);

/* getAlternateBucketWithNewFp -- CuckooFilter.cuh:388-400. */
uint64_t cfo_alt_bucket(const cfo_config* cfg, uint64_t nb, uint64_t bucket, uint32_t fp,
                        uint32_t* new_fp);

/* computePackedTagsKernel -- CuckooFilter.cuh:1185-1207 (PackedTagType truncation kept). */
void cfo_packed_tags(const cfo_config* cfg, uint64_t nb, const void* keys, size_t n,
                     uint64_t* packed);

/* Lifetime. capacity as in CuckooFilter(size_t capacity), CuckooFilter.cuh:420-428. */
cfo_filter* cfo_create(const cfo_config* cfg, uint64_t capacity);
void cfo_destroy(cfo_filter* f);
void cfo_clear(cfo_filter* f);
uint64_t cfo_num_buckets_of(const cfo_filter* f);
uint64_t cfo_size_bytes(const cfo_filter* f);
uint64_t cfo_occupied(const cfo_filter* f);
uint64_t cfo_evictions(const cfo_filter* f);
uint64_t cfo_count_occupied_slots(const cfo_filter* f);
uint8_t* cfo_table(cfo_filter* f);
void cfo_set_occupied(cfo_filter* f, uint64_t occupied);

/*
 * Batch operations.  threads == 1 is the sequential, deterministic oracle;
 * threads > 1 runs the same lock-free CAS algorithm on that many host
 * threads (the CPU timing baseline), keys split into contiguous ranges.
 *
 * ok[i]   (nullable) -- per-key success (insert / delete) or membership.
 * chain[i](nullable) -- eviction swaps performed for key i (insert only).
 * Return: occupied slots after the call (insert / delete), as
 *         CuckooFilter.cuh:463,547.
 */
uint64_t cfo_insert_many(cfo_filter* f, const void* keys, size_t n, uint8_t* ok,
                         uint32_t* chain, int threads);
uint64_t cfo_insert_many_sorted(cfo_filter* f, const void* keys, size_t n, uint32_t* chain);
/* Sequential run of the insert scheduled by step: all primary attempts, then
 * all secondary attempts, then (BFS) all BFS candidate loops, then DFS. */
uint64_t cfo_insert_many_phased(cfo_filter* f, const void* keys, size_t n, uint8_t* ok, uint32_t* chain);
void cfo_contains_many(cfo_filter* f, const void* keys, size_t n, uint8_t* out, int threads);
uint64_t cfo_delete_many(cfo_filter* f, const void* keys, size_t n, uint8_t* ok, int threads);

/* Algorithmic traffic of the calls since the last reset: bucket reads
 * (tryInsert/tryRemove scans; BFS/DFS steps read one bucket word more, counted
 * with the write) and words written by successful CAS. */
void cfo_stats(const cfo_filter* f, uint64_t* bucket_reads, uint64_t* word_writes);
void cfo_reset_stats(cfo_filter* f);

/* Membership of fp in bucket b of an arbitrary table image (bucket scan semantics). */
int cfo_bucket_contains(const cfo_filter* f, uint64_t b, uint32_t tag);

/* Number of bucket loads containsMany needs per key on this image (1 or 2). */
void cfo_probe_counts(cfo_filter* f, const void* keys, size_t n, uint8_t* probes);

/* Deterministic key stream: key_j = mix64(seed*C1 + j*C2), distinct for distinct j. */
void cfo_gen_keys(uint64_t seed, uint64_t start, size_t n, uint32_t key_bytes, void* out);

/* SWAR helpers restated literally from helpers.cuh:78-151 (for the exhaustive SWAR test). */
uint64_t cfo_swar_zero_mask(uint64_t v, uint32_t tag_bits, uint32_t word_bits);
uint64_t cfo_swar_replicate(uint32_t tag, uint32_t tag_bits, uint32_t word_bits);

#ifdef __cplusplus
}
#endif

#endif
