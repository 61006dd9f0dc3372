/*
 * cpu_cuckoo.c -- single-threaded CPU cuckoo filter (TEST INFRASTRUCTURE
 * ONLY: the reported CPU baseline that mirrors the reference's CPU
 * comparison, benchmark/cuckoo_cpu_vs_gpu.cu:52-155).  What it restates and
 * how it is checked: cpu_cuckoo.h.
 */
#include "cpu_cuckoo.h"

#include <stdlib.h>

enum { kSlots = 4, kMaxKicks = 500 };

struct ccf_filter {
    uint16_t* tags;  /* numBuckets * kSlots, 0 = empty */
    uint64_t numBuckets;
    uint64_t items;
    unsigned __int128 mul, add;
    uint64_t rng; /* xorshift64 state for the kicked slot */
    struct {
        uint64_t index;
        uint16_t tag;
        int used;
    } victim;
};

static uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t nextPow2(uint64_t n) {
    uint64_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

ccf_filter* ccf_create(uint64_t max_keys) {
    ccf_filter* f = (ccf_filter*)calloc(1, sizeof(ccf_filter));
    if (f == NULL) return NULL;
    uint64_t nb = nextPow2(max_keys / kSlots > 0 ? max_keys / kSlots : 1);
    if ((double)max_keys / (double)nb / kSlots > 0.96) nb <<= 1;
    f->numBuckets = nb;
    f->tags = (uint16_t*)calloc(nb * kSlots, sizeof(uint16_t));
    if (f->tags == NULL) {
        free(f);
        return NULL;
    }
    uint64_t s = 42;
    f->mul = ((unsigned __int128)splitmix64(&s) << 64) | (splitmix64(&s) | 1u);
    f->add = ((unsigned __int128)splitmix64(&s) << 64) | splitmix64(&s);
    f->rng = splitmix64(&s) | 1u;
    return f;
}

void ccf_destroy(ccf_filter* f) {
    if (f == NULL) return;
    free(f->tags);
    free(f);
}

uint64_t ccf_num_buckets(const ccf_filter* f) { return f->numBuckets; }
uint64_t ccf_size_bytes(const ccf_filter* f) { return f->numBuckets * kSlots * sizeof(uint16_t); }
uint64_t ccf_items(const ccf_filter* f) { return f->items; }
int ccf_victim_used(const ccf_filter* f) { return f->victim.used; }

uint64_t ccf_alt_index(const ccf_filter* f, uint64_t index, uint32_t tag) {
    return (index ^ ((uint64_t)tag * 0x5bd1e995ull)) & (f->numBuckets - 1);
}

void ccf_index_tag(const ccf_filter* f, uint64_t key, uint64_t* index, uint32_t* tag) {
    const uint64_t h = (uint64_t)((f->add + f->mul * (unsigned __int128)key) >> 64);
    *index = (h >> 32) & (f->numBuckets - 1);
    uint32_t t = (uint32_t)(h & 0xFFFFu);
    *tag = t == 0 ? 1u : t;
}

uint64_t ccf_checksum(const ccf_filter* f) {
    const unsigned char* p = (const unsigned char*)f->tags;
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint64_t i = 0; i < ccf_size_bytes(f); ++i) h = (h ^ p[i]) * 0x100000001b3ull;
    return h;
}

static int bucketHas(const ccf_filter* f, uint64_t i, uint32_t tag) {
    const uint16_t* b = f->tags + i * kSlots;
    return b[0] == tag || b[1] == tag || b[2] == tag || b[3] == tag;
}

/* Free slot of bucket i takes the tag (returns 1); else, with kick, a
   pseudo-random slot's tag goes to *old and the tag takes its place. */
static int putTag(ccf_filter* f, uint64_t i, uint16_t tag, int kick, uint16_t* old) {
    uint16_t* b = f->tags + i * kSlots;
    for (int j = 0; j < kSlots; ++j) {
        if (b[j] == 0) {
            b[j] = tag;
            return 1;
        }
    }
    if (kick) {
        f->rng ^= f->rng << 13;
        f->rng ^= f->rng >> 7;
        f->rng ^= f->rng << 17;
        const int r = (int)(f->rng % kSlots);
        *old = b[r];
        b[r] = tag;
    }
    return 0;
}

static void place(ccf_filter* f, uint64_t i, uint16_t tag) {
    uint64_t cur = i;
    uint16_t t = tag;
    for (int count = 0; count < kMaxKicks; ++count) {
        const int kick = count > 0;
        uint16_t old = 0;
        if (putTag(f, cur, t, kick, &old)) {
            f->items++;
            return;
        }
        if (kick) t = old;
        cur = ccf_alt_index(f, cur, t);
    }
    f->victim.index = cur;
    f->victim.tag = t;
    f->victim.used = 1;
}

int ccf_add(ccf_filter* f, uint64_t key) {
    if (f->victim.used) return CCF_NOT_ENOUGH_SPACE;
    uint64_t i;
    uint32_t tag;
    ccf_index_tag(f, key, &i, &tag);
    place(f, i, (uint16_t)tag);
    return CCF_OK;
}

static int victimMatches(const ccf_filter* f, uint64_t i1, uint64_t i2, uint32_t tag) {
    return f->victim.used && f->victim.tag == tag && (f->victim.index == i1 || f->victim.index == i2);
}

int ccf_contain(const ccf_filter* f, uint64_t key) {
    uint64_t i1;
    uint32_t tag;
    ccf_index_tag(f, key, &i1, &tag);
    const uint64_t i2 = ccf_alt_index(f, i1, tag);
    if (victimMatches(f, i1, i2, tag) || bucketHas(f, i1, tag) || bucketHas(f, i2, tag)) return CCF_OK;
    return CCF_NOT_FOUND;
}

static int clearTag(ccf_filter* f, uint64_t i, uint32_t tag) {
    uint16_t* b = f->tags + i * kSlots;
    for (int j = 0; j < kSlots; ++j) {
        if (b[j] == tag) {
            b[j] = 0;
            return 1;
        }
    }
    return 0;
}

int ccf_delete(ccf_filter* f, uint64_t key) {
    uint64_t i1;
    uint32_t tag;
    ccf_index_tag(f, key, &i1, &tag);
    const uint64_t i2 = ccf_alt_index(f, i1, tag);
    if (clearTag(f, i1, tag) || clearTag(f, i2, tag)) {
        f->items--;
        if (f->victim.used) {  /* room again: the cached tag goes back in */
            f->victim.used = 0;
            place(f, f->victim.index, f->victim.tag);
        }
        return CCF_OK;
    }
    if (victimMatches(f, i1, i2, tag)) {
        f->victim.used = 0;
        return CCF_OK;
    }
    return CCF_NOT_FOUND;
}

uint64_t ccf_add_many(ccf_filter* f, const uint64_t* keys, size_t n, uint8_t* out) {
    uint64_t ok = 0;
    for (size_t i = 0; i < n; ++i) {
        const int r = ccf_add(f, keys[i]) == CCF_OK;
        ok += (uint64_t)r;
        if (out != NULL) out[i] = (uint8_t)r;
    }
    return ok;
}

uint64_t ccf_contain_many(const ccf_filter* f, const uint64_t* keys, size_t n, uint8_t* out) {
    uint64_t ok = 0;
    for (size_t i = 0; i < n; ++i) {
        const int r = ccf_contain(f, keys[i]) == CCF_OK;
        ok += (uint64_t)r;
        if (out != NULL) out[i] = (uint8_t)r;
    }
    return ok;
}

uint64_t ccf_delete_many(ccf_filter* f, const uint64_t* keys, size_t n, uint8_t* out) {
    uint64_t ok = 0;
    for (size_t i = 0; i < n; ++i) {
        const int r = ccf_delete(f, keys[i]) == CCF_OK;
        ok += (uint64_t)r;
        if (out != NULL) out[i] = (uint8_t)r;
    }
    return ok;
}
