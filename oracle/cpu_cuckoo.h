/*
 * cpu_cuckoo.h -- single-threaded CPU cuckoo filter after the published
 * design that the reference benchmarks its GPU filter against.
 *
 * TEST INFRASTRUCTURE ONLY (a reported CPU baseline, never a parity target
 * and never part of the product library).  The reference's CPU comparison
 * (benchmark/cuckoo_cpu_vs_gpu.cu:52-155) times
 * cuckoofilter::CuckooFilter<uint64_t, 16> from efficient/cuckoofilter, a
 * network-only dependency (subprojects/cpu-cuckoo.wrap) whose source is not
 * available here.  This is a clean-room restatement of the design published
 * by Fan, Andersen, Kaminsky and Mitzenmacher ("Cuckoo Filter: Practically
 * Better Than Bloom", CoNEXT 2014) in the shape that library packages it:
 *   - 4 slots of 16-bit tags per bucket, a power-of-two bucket count sized
 *     for the requested key count, doubled when that would exceed 96% load;
 *   - a 64-bit multiply-shift hash of the key (128-bit product, fixed odd
 *     multiplier and addend); bucket = high 32 bits, tag = low 16 bits, tag 0
 *     becomes 1;
 *   - partial-key cuckoo hashing: alt = bucket ^ (tag * 0x5bd1e995), masked;
 *   - insert: first free slot of the first bucket, then of the alternate;
 *     otherwise kick a pseudo-random slot and move its tag to that tag's
 *     alternate, at most 500 times; a tag left over is kept in a one-entry
 *     victim cache, and the filter refuses further inserts while it is full;
 *   - lookup: either bucket holds the tag, or the victim matches;
 *   - delete: clear one matching slot of either bucket (or the victim), then
 *     re-insert a cached victim.
 * Its hash and kick choices are this restatement's own, so table images are
 * not comparable with the library's; the tests pin its behaviour instead (no
 * false negatives, FPR against the closed form, sizing, delete round trips).
 */
#ifndef CPU_CUCKOO_H
#define CPU_CUCKOO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { CCF_OK = 0, CCF_NOT_FOUND = 1, CCF_NOT_ENOUGH_SPACE = 2 };

typedef struct ccf_filter ccf_filter;

/* Filter sized for max_keys keys; NULL when allocation fails. */
ccf_filter* ccf_create(uint64_t max_keys);
void ccf_destroy(ccf_filter* f);

uint64_t ccf_num_buckets(const ccf_filter* f);
uint64_t ccf_size_bytes(const ccf_filter* f);
uint64_t ccf_items(const ccf_filter* f);  /* tags stored in buckets */
int ccf_victim_used(const ccf_filter* f);
uint64_t ccf_alt_index(const ccf_filter* f, uint64_t index, uint32_t tag);
/* Bucket and tag of one key. */
void ccf_index_tag(const ccf_filter* f, uint64_t key, uint64_t* index, uint32_t* tag);
/* FNV-1a over the bucket array, for determinism checks. */
uint64_t ccf_checksum(const ccf_filter* f);

int ccf_add(ccf_filter* f, uint64_t key);
int ccf_contain(const ccf_filter* f, uint64_t key);
int ccf_delete(ccf_filter* f, uint64_t key);

/* Loops over keys in order; each returns how many calls gave CCF_OK.
   out (optional) receives 1 for CCF_OK, else 0. */
uint64_t ccf_add_many(ccf_filter* f, const uint64_t* keys, size_t n, uint8_t* out);
uint64_t ccf_contain_many(const ccf_filter* f, const uint64_t* keys, size_t n, uint8_t* out);
uint64_t ccf_delete_many(ccf_filter* f, const uint64_t* keys, size_t n, uint8_t* out);

#ifdef __cplusplus
}
#endif

#endif /* CPU_CUCKOO_H */
