/*
 * cuckoo_oracle.c -- CPU oracle for the cuckoo-filter hot path.
 *
 * TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py's CPU
 * baseline).  A sequential plain-C restatement of the reference algorithm;
 * each function cites the reference file:line it follows.  Slot searches are
 * written as plain lane loops (the semantics the reference's SWAR code
 * computes), so the oracle does not share the product's SWAR arithmetic.
 * The literal SWAR formulas are restated separately (cfo_swar_*) only so the
 * tests can prove them equal to the lane loops.
 *
 * All word accesses go through relaxed atomics so that the very same code,
 * run on several host threads, is the lock-free multi-threaded CPU baseline.
 */
#include "cuckoo_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* XXH64 (hashutil.cuh:44-161; standard XXH64, seed 0 on the path)           */
/* ------------------------------------------------------------------------- */

static const uint64_t P1 = 11400714785074694791ULL;
static const uint64_t P2 = 14029467366897019727ULL;
static const uint64_t P3 = 1609587929392839161ULL;
static const uint64_t P4 = 9650029242287828579ULL;
static const uint64_t P5 = 2870177450012600261ULL;

static inline uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

static inline uint64_t rd64(const uint8_t* p) {
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}

static inline uint32_t rd32(const uint8_t* p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

static inline uint64_t round64(uint64_t acc, uint64_t lane) {
    acc += lane * P2;
    acc = rotl(acc, 31);
    return acc * P1;
}

static inline uint64_t merge64(uint64_t h, uint64_t v) {
    h ^= round64(0, v);
    return h * P1 + P4;
}

uint64_t cfo_xxh64(const void* data, size_t len, uint64_t seed) {
    const uint8_t* p = (const uint8_t*)data;
    const uint8_t* end = p + len;
    uint64_t h;
    if (len >= 32) { /* hashutil.cuh:78-126 */
        uint64_t v1 = seed + P1 + P2, v2 = seed + P2, v3 = seed, v4 = seed - P1;
        const uint8_t* limit = end - 32;
        do {
            v1 = round64(v1, rd64(p));
            v2 = round64(v2, rd64(p + 8));
            v3 = round64(v3, rd64(p + 16));
            v4 = round64(v4, rd64(p + 24));
            p += 32;
        } while (p <= limit);
        h = rotl(v1, 1) + rotl(v2, 7) + rotl(v3, 12) + rotl(v4, 18);
        h = merge64(h, v1);
        h = merge64(h, v2);
        h = merge64(h, v3);
        h = merge64(h, v4);
    } else {
        h = seed + P5; /* hashutil.cuh:127-129 */
    }
    h += (uint64_t)len; /* :131 */
    while (p + 8 <= end) { /* :133-141 */
        h ^= round64(0, rd64(p));
        h = rotl(h, 27) * P1 + P4;
        p += 8;
    }
    if (p + 4 <= end) { /* :143-149 */
        h ^= (uint64_t)rd32(p) * P1;
        h = rotl(h, 23) * P2 + P3;
        p += 4;
    }
    while (p < end) { /* :151-158 */
        h ^= (uint64_t)(*p) * P5;
        h = rotl(h, 11) * P1;
        ++p;
    }
    /* finalize, hashutil.cuh:61-68 */
    h ^= h >> 33;
    h *= P2;
    h ^= h >> 29;
    h *= P3;
    h ^= h >> 32;
    return h;
}

/* helpers.cuh:14-16, 31-45 */
uint64_t cfo_next_pow2(uint64_t n) {
    if (n != 0 && (n & (n - 1)) == 0) return n;
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

/* ------------------------------------------------------------------------- */
/* configuration                                                              */
/* ------------------------------------------------------------------------- */

static inline int is_pow2(uint64_t n) { return n != 0 && (n & (n - 1)) == 0; }

static inline uint32_t tag_bytes(const cfo_config* c) {
    /* CuckooFilter.cuh:61-64: TagType = u8 / u16 / u32 by bitsPerTag. */
    return c->bits_per_tag <= 8 ? 1u : (c->bits_per_tag <= 16 ? 2u : 4u);
}

int cfo_config_check(const cfo_config* c) {
    if (c->key_bytes != 4 && c->key_bytes != 8) return -1;
    /* CuckooFilter.cuh:147-150 */
    if (c->bits_per_tag != 8 && c->bits_per_tag != 16 && c->bits_per_tag != 32) return -2;
    /* CuckooFilter.cuh:67-71 */
    if (c->word_bytes != 4 && c->word_bytes != 8) return -3;
    if (tag_bytes(c) > c->word_bytes) return -3;
    /* CuckooFilter.cuh:152, 226-236 */
    if (!is_pow2(c->bucket_size)) return -4;
    uint32_t tpw = c->word_bytes / tag_bytes(c);
    if (c->bucket_size % tpw != 0) return -5;
    if (c->policy > CFO_POLICY_OFFSET) return -6;
    if (c->eviction > CFO_EVICT_DFS) return -7;
    return 0;
}

/* bucket_policies.cuh:80-83 (XOR), 172-181 (AddSub), 347-350 (Offset) */
uint64_t cfo_num_buckets(const cfo_config* c, uint64_t capacity) {
    double required = ceil((double)capacity / (double)c->bucket_size);
    if (c->policy == CFO_POLICY_XOR) return cfo_next_pow2((uint64_t)required);
    if (c->policy == CFO_POLICY_ADDSUB) {
        if (((uint64_t)required) % 2 != 0) required += 1;
        return (uint64_t)required;
    }
    return (uint64_t)required;
}

/* ------------------------------------------------------------------------- */
/* alternate-bucket policies                                                  */
/* ------------------------------------------------------------------------- */

static inline uint64_t hash_key(const cfo_config* c, const uint8_t* key) {
    return cfo_xxh64(key, c->key_bytes, 0);
}

/* hash64 of a TagType value: sizeof(TagType) little-endian bytes
 * (bucket_policies.cuh:73, 159, 249 -> hashutil.cuh:70-73). */
static inline uint64_t hash_tag(const cfo_config* c, uint32_t tag) {
    uint8_t b[4];
    memcpy(b, &tag, 4);
    return cfo_xxh64(b, tag_bytes(c), 0);
}

static inline uint64_t fp_mask_of(const cfo_config* c) {
    return (c->bits_per_tag == 32) ? 0xFFFFFFFFull : ((1ull << c->bits_per_tag) - 1);
}

/* OffsetAltBucketPolicy helpers, bucket_policies.cuh:203-252 */
static inline uint32_t off_pure_mask(const cfo_config* c) {
    return (uint32_t)((1ull << (c->bits_per_tag - 1)) - 1);
}

static inline uint64_t off_offset(const cfo_config* c, uint32_t pure, uint64_t nb) {
    uint64_t o = hash_tag(c, pure) % nb; /* :248-252 */
    return o == 0 ? 1 : o;
}

uint64_t cfo_alt_bucket(const cfo_config* c, uint64_t nb, uint64_t b, uint32_t fp,
                        uint32_t* new_fp) {
    if (c->policy == CFO_POLICY_XOR) { /* bucket_policies.cuh:71-74 */
        if (new_fp) *new_fp = fp;
        return b ^ (hash_tag(c, fp) & (nb - 1));
    }
    if (c->policy == CFO_POLICY_ADDSUB) { /* :156-166 */
        uint64_t bpb = nb / 2;
        uint64_t fh = hash_tag(c, fp) % bpb;
        if (new_fp) *new_fp = fp;
        if (b < bpb) return ((b + fh) % bpb) + bpb;
        return (b - fh) % bpb;
    }
    /* Offset, :327-342 (getAlternateBucketWithNewFp) */
    uint32_t pos = c->bits_per_tag - 1;
    uint32_t pure = fp & off_pure_mask(c);
    uint32_t choice = (fp >> pos) & 1u;
    uint64_t off = off_offset(c, pure, nb);
    uint64_t alt = (choice == 0) ? (b + off) % nb : (b + nb - off) % nb;
    if (new_fp) *new_fp = pure | ((1u - choice) << pos);
    return alt;
}

static inline void candidates_one(const cfo_config* c, uint64_t nb, const uint8_t* key,
                                  uint64_t* i1, uint64_t* i2, uint32_t* fp) {
    uint64_t h = hash_key(c, key);
    uint32_t h_fp = (uint32_t)(h >> 32);
    uint32_t h_bucket = (uint32_t)(h & 0xFFFFFFFFull);
    if (c->policy == CFO_POLICY_OFFSET) { /* bucket_policies.cuh:263-285 */
        uint32_t masked = h_fp & off_pure_mask(c);
        uint32_t pure = masked == 0 ? 1u : masked;
        *fp = pure; /* choice bit 0 */
        *i1 = (uint64_t)h_bucket % nb;
        *i2 = (*i1 + off_offset(c, pure, nb)) % nb;
        return;
    }
    uint32_t masked = (uint32_t)(h_fp & fp_mask_of(c)); /* :49-53, :130-133 */
    *fp = masked == 0 ? 1u : masked;
    if (c->policy == CFO_POLICY_XOR) {
        *i1 = (uint64_t)h_bucket & (nb - 1); /* :56-57 */
    } else {
        *i1 = (uint64_t)h_bucket % (nb / 2); /* :136-139 */
    }
    *i2 = cfo_alt_bucket(c, nb, *i1, *fp, NULL);
}

void cfo_candidates(const cfo_config* c, uint64_t nb, const void* keys, size_t n,
                    uint64_t* i1, uint64_t* i2, uint32_t* fp) {
    const uint8_t* k = (const uint8_t*)keys;
    for (size_t i = 0; i < n; ++i) candidates_one(c, nb, k + i * c->key_bytes, &i1[i], &i2[i], &fp[i]);
}

/* computePackedTagsKernel, CuckooFilter.cuh:159, 1185-1207: PackedTagType is
 * u32 for bitsPerTag <= 8, else u64; (PackedTagType)i1 << bitsPerTag | fp. */
static inline uint64_t pack_one(const cfo_config* c, uint64_t i1, uint32_t fp) {
    if (c->bits_per_tag <= 8) {
        uint32_t v = ((uint32_t)i1 << c->bits_per_tag) | (uint32_t)fp;
        return v;
    }
    return (i1 << c->bits_per_tag) | (uint64_t)fp;
}

void cfo_packed_tags(const cfo_config* c, uint64_t nb, const void* keys, size_t n,
                     uint64_t* packed) {
    const uint8_t* k = (const uint8_t*)keys;
    for (size_t i = 0; i < n; ++i) {
        uint64_t i1, i2;
        uint32_t fp;
        candidates_one(c, nb, k + i * c->key_bytes, &i1, &i2, &fp);
        packed[i] = pack_one(c, i1, fp);
    }
}

/* ------------------------------------------------------------------------- */
/* filter state and bucket layout (CuckooFilter.cuh:209-249)                  */
/* ------------------------------------------------------------------------- */

struct cfo_filter {
    cfo_config cfg;
    uint64_t nb;
    uint32_t tpw;      /* tags per word */
    uint32_t wc;       /* words per bucket */
    uint32_t bits;
    uint64_t fp_mask;
    uint64_t word_mask;
    uint8_t* table;    /* nb * wc words, little-endian */
    uint64_t occupied; /* d_numOccupied */
    uint64_t evictions;
    /* algorithmic traffic of the sequential run: buckets read, words written */
    uint64_t stat_bucket_reads;
    uint64_t stat_word_writes;
};

cfo_filter* cfo_create(const cfo_config* cfg, uint64_t capacity) {
    if (cfo_config_check(cfg) != 0) return NULL;
    uint64_t nb = cfo_num_buckets(cfg, capacity);
    if (nb == 0) return NULL;
    cfo_filter* f = (cfo_filter*)calloc(1, sizeof(cfo_filter));
    if (!f) return NULL;
    f->cfg = *cfg;
    f->nb = nb;
    f->tpw = cfg->word_bytes / tag_bytes(cfg);
    f->wc = cfg->bucket_size / f->tpw;
    f->bits = cfg->bits_per_tag;
    f->fp_mask = fp_mask_of(cfg);
    f->word_mask = cfg->word_bytes == 8 ? ~0ull : 0xFFFFFFFFull;
    size_t bytes = (size_t)nb * f->wc * cfg->word_bytes;
    size_t alloc = (bytes + 63) & ~(size_t)63;
    f->table = (uint8_t*)aligned_alloc(64, alloc ? alloc : 64);
    if (!f->table) {
        free(f);
        return NULL;
    }
    cfo_clear(f);
    return f;
}

void cfo_destroy(cfo_filter* f) {
    if (!f) return;
    free(f->table);
    free(f);
}

void cfo_clear(cfo_filter* f) { /* CuckooFilter.cuh:675-683 */
    memset(f->table, 0, (size_t)f->nb * f->wc * f->cfg.word_bytes);
    f->occupied = 0;
    f->evictions = 0;
    f->stat_bucket_reads = 0;
    f->stat_word_writes = 0;
}

void cfo_stats(const cfo_filter* f, uint64_t* bucket_reads, uint64_t* word_writes) {
    *bucket_reads = f->stat_bucket_reads;
    *word_writes = f->stat_word_writes;
}

void cfo_reset_stats(cfo_filter* f) {
    f->stat_bucket_reads = 0;
    f->stat_word_writes = 0;
}

uint64_t cfo_num_buckets_of(const cfo_filter* f) { return f->nb; }
uint64_t cfo_size_bytes(const cfo_filter* f) { return f->nb * f->wc * f->cfg.word_bytes; }
uint64_t cfo_occupied(const cfo_filter* f) { return f->occupied; }
uint64_t cfo_evictions(const cfo_filter* f) { return f->evictions; }
uint8_t* cfo_table(cfo_filter* f) { return f->table; }
void cfo_set_occupied(cfo_filter* f, uint64_t occupied) { f->occupied = occupied; }

static inline uint64_t ld_word(const cfo_filter* f, uint64_t b, uint32_t w) {
    size_t idx = (size_t)b * f->wc + w;
    if (f->cfg.word_bytes == 8) return __atomic_load_n(((uint64_t*)f->table) + idx, __ATOMIC_RELAXED);
    return __atomic_load_n(((uint32_t*)f->table) + idx, __ATOMIC_RELAXED);
}

/* compare_exchange (relaxed): on failure *expected receives the current word. */
static inline int cas_word(cfo_filter* f, uint64_t b, uint32_t w, uint64_t* expected,
                           uint64_t desired) {
    size_t idx = (size_t)b * f->wc + w;
    if (f->cfg.word_bytes == 8) {
        return __atomic_compare_exchange_n(((uint64_t*)f->table) + idx, expected, desired, 0,
                                           __ATOMIC_RELAXED, __ATOMIC_RELAXED);
    }
    uint32_t e = (uint32_t)*expected;
    int ok = __atomic_compare_exchange_n(((uint32_t*)f->table) + idx, &e, (uint32_t)desired, 0,
                                         __ATOMIC_RELAXED, __ATOMIC_RELAXED);
    *expected = e;
    return ok;
}

/* Bucket::extractTag / replaceTag, CuckooFilter.cuh:239-249 */
static inline uint32_t lane_get(const cfo_filter* f, uint64_t word, uint32_t lane) {
    return (uint32_t)((word >> (lane * f->bits)) & f->fp_mask);
}

static inline uint64_t lane_set(const cfo_filter* f, uint64_t word, uint32_t lane, uint32_t tag) {
    uint32_t shift = lane * f->bits;
    uint64_t cleared = word & ~(f->fp_mask << shift);
    return (cleared | ((uint64_t)tag << shift)) & f->word_mask;
}

/* lowest lane of word equal to v, or -1 (what the SWAR ffs picks). */
static inline int lowest_lane_eq(const cfo_filter* f, uint64_t word, uint32_t v) {
    for (uint32_t l = 0; l < f->tpw; ++l)
        if (lane_get(f, word, l) == v) return (int)l;
    return -1;
}

uint64_t cfo_count_occupied_slots(const cfo_filter* f) { /* CuckooFilter.cuh:759-785 */
    uint64_t cnt = 0;
    for (uint64_t b = 0; b < f->nb; ++b)
        for (uint32_t w = 0; w < f->wc; ++w) {
            uint64_t word = ld_word(f, b, w);
            for (uint32_t l = 0; l < f->tpw; ++l) cnt += lane_get(f, word, l) != 0;
        }
    return cnt;
}

/* Bucket::contains, CuckooFilter.cuh:290-346: true iff some slot holds tag.
 * The reference's scan order (start chunk, wrap) only changes where it stops. */
int cfo_bucket_contains(const cfo_filter* f, uint64_t b, uint32_t tag) {
    for (uint32_t w = 0; w < f->wc; ++w)
        if (lowest_lane_eq(f, ld_word(f, b, w), tag) >= 0) return 1;
    return 0;
}

/* tryRemoveAtBucket, CuckooFilter.cuh:805-848 */
static int try_remove(cfo_filter* f, uint64_t b, uint32_t tag) {
    uint32_t start_word = (tag & (f->cfg.bucket_size - 1)) / f->tpw;
    __atomic_fetch_add(&f->stat_bucket_reads, 1, __ATOMIC_RELAXED);
    for (uint32_t i = 0; i < f->wc; ++i) {
        uint32_t w = (start_word + i) & (f->wc - 1);
        for (;;) {
            uint64_t expected = ld_word(f, b, w);
            int lane = lowest_lane_eq(f, expected, tag);
            if (lane < 0) break;
            uint64_t desired = lane_set(f, expected, (uint32_t)lane, 0);
            if (cas_word(f, b, w, &expected, desired)) {
                __atomic_fetch_add(&f->stat_word_writes, 1, __ATOMIC_RELAXED);
                return 1;
            }
        }
    }
    return 0;
}

/* tryInsertAtBucket, CuckooFilter.cuh:859-904 */
static int try_insert(cfo_filter* f, uint64_t b, uint32_t tag) {
    uint32_t start_word = (tag & (f->cfg.bucket_size - 1)) / f->tpw;
    __atomic_fetch_add(&f->stat_bucket_reads, 1, __ATOMIC_RELAXED);
    for (uint32_t i = 0; i < f->wc; ++i) {
        uint32_t w = (start_word + i) & (f->wc - 1);
        uint64_t expected = ld_word(f, b, w);
        for (;;) {
            int lane = lowest_lane_eq(f, expected, 0);
            if (lane < 0) break;
            uint64_t desired = lane_set(f, expected, (uint32_t)lane, tag);
            if (cas_word(f, b, w, &expected, desired)) {
                __atomic_fetch_add(&f->stat_word_writes, 1, __ATOMIC_RELAXED);
                return 1;
            }
        }
    }
    return 0;
}

/* insertWithEvictionDFS, CuckooFilter.cuh:917-954 */
static int insert_dfs(cfo_filter* f, uint32_t fp, uint64_t start, uint32_t* chain) {
    uint32_t cur_fp = fp;
    uint64_t cur_b = start;
    for (uint64_t e = 0; e < f->cfg.max_evictions; ++e) {
        uint32_t slot = (uint32_t)(((uint64_t)cur_fp + e * 7) & (f->cfg.bucket_size - 1));
        uint32_t w = slot / f->tpw;
        uint32_t lane = slot & (f->tpw - 1);
        uint64_t expected = ld_word(f, cur_b, w);
        uint32_t evicted;
        for (;;) {
            evicted = lane_get(f, expected, lane);
            uint64_t desired = lane_set(f, expected, lane, cur_fp);
            if (cas_word(f, cur_b, w, &expected, desired)) break;
        }
        __atomic_fetch_add(&f->evictions, 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&f->stat_word_writes, 1, __ATOMIC_RELAXED);
        if (chain) *chain = (uint32_t)(e + 1);
        uint32_t new_fp;
        cur_b = cfo_alt_bucket(&f->cfg, f->nb, cur_b, evicted, &new_fp);
        cur_fp = new_fp;
        if (try_insert(f, cur_b, cur_fp)) return 1;
    }
    return 0;
}

/* The candidate loop of insertWithEvictionBFS, CuckooFilter.cuh:967-1015
 * (everything before its DFS fallback). */
static int bfs_candidates(cfo_filter* f, uint32_t fp, uint64_t start, uint32_t* chain) {
    uint32_t half = f->cfg.bucket_size / 2;
    uint32_t ncand = half < 8 ? half : 8;
    for (uint32_t i = 0; i < ncand; ++i) {
        uint32_t slot = (uint32_t)(((uint64_t)fp + (uint64_t)i * 7) & (f->cfg.bucket_size - 1));
        uint32_t w = slot / f->tpw;
        uint32_t lane = slot & (f->tpw - 1);
        uint32_t cand = lane_get(f, ld_word(f, start, w), lane);
        if (cand == 0) {
            if (try_insert(f, start, fp)) return 1;
            continue;
        }
        uint32_t alt_fp;
        uint64_t alt = cfo_alt_bucket(&f->cfg, f->nb, start, cand, &alt_fp);
        if (try_insert(f, alt, alt_fp)) {
            uint64_t expected = ld_word(f, start, w);
            if (lane_get(f, expected, lane) == cand) {
                uint64_t desired = lane_set(f, expected, lane, fp);
                if (cas_word(f, start, w, &expected, desired)) {
                    __atomic_fetch_add(&f->evictions, 1, __ATOMIC_RELAXED);
                    __atomic_fetch_add(&f->stat_word_writes, 1, __ATOMIC_RELAXED);
                    if (chain) *chain = 1;
                    return 1;
                }
            }
            try_remove(f, alt, cand); /* :1011-1012 (removes cand, as written) */
        }
    }
    return 0;
}

/* insertWithEvictionBFS, CuckooFilter.cuh:967-1018 */
static int insert_bfs(cfo_filter* f, uint32_t fp, uint64_t start, uint32_t* chain) {
    return bfs_candidates(f, fp, start, chain) || insert_dfs(f, fp, start, chain);
}

static int evict(cfo_filter* f, uint32_t fp, uint64_t start, uint32_t* chain) {
    if (f->cfg.eviction == CFO_EVICT_BFS) return insert_bfs(f, fp, start, chain);
    return insert_dfs(f, fp, start, chain);
}

/* CuckooFilter::insert, CuckooFilter.cuh:1029-1080, entered with (i1, fp). */
static int insert_from(cfo_filter* f, uint64_t i1, uint32_t fp, uint32_t* chain) {
    if (chain) *chain = 0;
    uint32_t fp2;
    uint64_t i2 = cfo_alt_bucket(&f->cfg, f->nb, i1, fp, &fp2);
    if (try_insert(f, i1, fp)) return 1;
    if (try_insert(f, i2, fp2)) return 1;
    /* :1065-1066 / :1243-1244: start bucket chosen by the low fp bit */
    int first = (fp & 1u) == 0;
    return evict(f, first ? fp : fp2, first ? i1 : i2, chain);
}

/* CuckooFilter::contains, CuckooFilter.cuh:1088-1098 */
static int contains_one(const cfo_filter* f, const uint8_t* key, int* probes) {
    uint64_t i1, i2;
    uint32_t fp, fp2;
    candidates_one(&f->cfg, f->nb, key, &i1, &i2, &fp);
    cfo_alt_bucket(&f->cfg, f->nb, i1, fp, &fp2);
    if (probes) *probes = 1;
    if (cfo_bucket_contains(f, i1, fp)) return 1;
    if (probes) *probes = (i1 == i2 && fp == fp2) ? 1 : 2;
    return cfo_bucket_contains(f, i2, fp2);
}

/* CuckooFilter::remove, CuckooFilter.cuh:1106-1116 */
static int remove_one(cfo_filter* f, const uint8_t* key) {
    uint64_t i1, i2;
    uint32_t fp, fp2;
    candidates_one(&f->cfg, f->nb, key, &i1, &i2, &fp);
    cfo_alt_bucket(&f->cfg, f->nb, i1, fp, &fp2);
    return try_remove(f, i1, fp) || try_remove(f, i2, fp2);
}

/* ------------------------------------------------------------------------- */
/* batch drivers (insertKernel / containsKernel / deleteKernel semantics)     */
/* ------------------------------------------------------------------------- */

typedef struct {
    cfo_filter* f;
    const uint8_t* keys;
    size_t lo, hi;
    uint8_t* ok;
    uint32_t* chain;
    int op; /* 0 insert, 1 contains, 2 delete */
    uint64_t successes;
} work_t;

static void* run_range(void* arg) {
    work_t* w = (work_t*)arg;
    cfo_filter* f = w->f;
    uint32_t kb = f->cfg.key_bytes;
    uint64_t s = 0;
    for (size_t i = w->lo; i < w->hi; ++i) {
        const uint8_t* key = w->keys + i * kb;
        int r;
        if (w->op == 0) {
            uint64_t i1, i2;
            uint32_t fp;
            candidates_one(&f->cfg, f->nb, key, &i1, &i2, &fp);
            r = insert_from(f, i1, fp, w->chain ? &w->chain[i] : NULL);
        } else if (w->op == 1) {
            r = contains_one(f, key, NULL);
        } else {
            r = remove_one(f, key);
        }
        if (w->ok) w->ok[i] = (uint8_t)r;
        s += (uint64_t)r;
    }
    w->successes = s;
    return NULL;
}

static uint64_t run_batch(cfo_filter* f, const void* keys, size_t n, uint8_t* ok, uint32_t* chain,
                          int op, int threads) {
    if (threads < 1) threads = 1;
    if ((size_t)threads > n) threads = n ? (int)n : 1;
    work_t* ws = (work_t*)calloc((size_t)threads, sizeof(work_t));
    pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    size_t per = (n + (size_t)threads - 1) / (size_t)threads;
    for (int t = 0; t < threads; ++t) {
        ws[t].f = f;
        ws[t].keys = (const uint8_t*)keys;
        ws[t].lo = (size_t)t * per < n ? (size_t)t * per : n;
        ws[t].hi = ws[t].lo + per < n ? ws[t].lo + per : n;
        ws[t].ok = ok;
        ws[t].chain = chain;
        ws[t].op = op;
    }
    if (threads == 1) {
        run_range(&ws[0]);
    } else {
        for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, run_range, &ws[t]);
        for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    }
    uint64_t s = 0;
    for (int t = 0; t < threads; ++t) s += ws[t].successes;
    free(ws);
    free(th);
    return s;
}

uint64_t cfo_insert_many(cfo_filter* f, const void* keys, size_t n, uint8_t* ok, uint32_t* chain,
                         int threads) {
    f->occupied += run_batch(f, keys, n, ok, chain, 0, threads); /* :1133-1140 */
    return f->occupied;
}

void cfo_contains_many(cfo_filter* f, const void* keys, size_t n, uint8_t* out, int threads) {
    run_batch(f, keys, n, out, NULL, 1, threads);
}

uint64_t cfo_delete_many(cfo_filter* f, const void* keys, size_t n, uint8_t* ok, int threads) {
    f->occupied -= run_batch(f, keys, n, ok, NULL, 2, threads); /* :1176-1182 */
    return f->occupied;
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* insertManySorted, CuckooFilter.cuh:475-508 + insertKernelSorted :1209-1266:
 * packed tags sorted ascending, then inserted from (packed >> bits, packed & mask). */
uint64_t cfo_insert_many_sorted(cfo_filter* f, const void* keys, size_t n, uint32_t* chain) {
    uint64_t* packed = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
    cfo_packed_tags(&f->cfg, f->nb, keys, n, packed);
    qsort(packed, n, sizeof(uint64_t), cmp_u64);
    uint64_t s = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t primary = packed[i] >> f->bits;
        uint32_t fp = (uint32_t)(packed[i] & f->fp_mask);
        s += (uint64_t)insert_from(f, primary, fp, chain ? &chain[i] : NULL);
    }
    free(packed);
    f->occupied += s;
    return f->occupied;
}

/* The reference's insert (CuckooFilter.cuh:1029-1080), every key's steps in
 * the same order as one key alone, but scheduled by step over the batch:
 * every key's primary-bucket attempt first, then the secondary attempts of
 * the keys still unplaced, then (BFS policy) their BFS candidate loops, then
 * DFS chains for the rest.  Each phase runs in key order.  This is the
 * schedule of the bucket-binned GPU insert (a legal interleaving of the
 * reference's concurrent insertKernel), run sequentially. */
uint64_t cfo_insert_many_phased(cfo_filter* f, const void* keys, size_t n, uint8_t* ok, uint32_t* chain) {
    const uint8_t* k = (const uint8_t*)keys;
    uint64_t* start = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
    uint32_t* sfp = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint8_t* done = (uint8_t*)calloc(n ? n : 1, 1);
    uint64_t s = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t i1, i2;
        uint32_t fp;
        if (chain) chain[i] = 0;
        candidates_one(&f->cfg, f->nb, k + i * f->cfg.key_bytes, &i1, &i2, &fp);
        start[i] = i1;
        sfp[i] = fp;
        done[i] = (uint8_t)try_insert(f, i1, fp);
    }
    for (size_t i = 0; i < n; ++i) {
        if (done[i]) continue;
        uint32_t fp = sfp[i], fp2;
        uint64_t i2 = cfo_alt_bucket(&f->cfg, f->nb, start[i], fp, &fp2);
        done[i] = (uint8_t)try_insert(f, i2, fp2);
        if (!done[i] && (fp & 1u) != 0) { /* :1065-1066: eviction starts at i2 for odd fp */
            start[i] = i2;
            sfp[i] = fp2;
        }
    }
    if (f->cfg.eviction == CFO_EVICT_BFS) {
        for (size_t i = 0; i < n; ++i)
            if (!done[i]) done[i] = (uint8_t)bfs_candidates(f, sfp[i], start[i], chain ? &chain[i] : NULL);
    }
    for (size_t i = 0; i < n; ++i)
        if (!done[i]) done[i] = (uint8_t)insert_dfs(f, sfp[i], start[i], chain ? &chain[i] : NULL);
    for (size_t i = 0; i < n; ++i) {
        s += done[i];
        if (ok) ok[i] = done[i];
    }
    free(start);
    free(sfp);
    free(done);
    f->occupied += s;
    return f->occupied;
}

void cfo_probe_counts(cfo_filter* f, const void* keys, size_t n, uint8_t* probes) {
    const uint8_t* k = (const uint8_t*)keys;
    for (size_t i = 0; i < n; ++i) {
        int p;
        contains_one(f, k + i * f->cfg.key_bytes, &p);
        probes[i] = (uint8_t)p;
    }
}

/* ------------------------------------------------------------------------- */
/* key generator and literal SWAR restatement                                 */
/* ------------------------------------------------------------------------- */

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void cfo_gen_keys(uint64_t seed, uint64_t start, size_t n, uint32_t key_bytes, void* out) {
    for (size_t i = 0; i < n; ++i) {
        uint64_t x = seed * 0xD1B54A32D192ED03ull + (start + i) * 0x9E3779B97F4A7C15ull;
        uint64_t k = mix64(x);
        if (key_bytes == 8) {
            ((uint64_t*)out)[i] = k;
        } else {
            ((uint32_t*)out)[i] = (uint32_t)k;
        }
    }
}

/* helpers.cuh:78-103 */
uint64_t cfo_swar_zero_mask(uint64_t v, uint32_t tag_bits, uint32_t word_bits) {
    uint64_t wmask = word_bits == 64 ? ~0ull : 0xFFFFFFFFull;
    uint64_t ones = cfo_swar_replicate(1, tag_bits, word_bits);
    uint64_t high = ones << (tag_bits - 1);
    v &= wmask;
    return ((v - ones) & ~v & high) & wmask;
}

/* helpers.cuh:126-151 */
uint64_t cfo_swar_replicate(uint32_t tag, uint32_t tag_bits, uint32_t word_bits) {
    uint64_t r = 0;
    for (uint32_t s = 0; s < word_bits; s += tag_bits) r |= (uint64_t)tag << s;
    return word_bits == 64 ? r : (r & 0xFFFFFFFFull);
}
