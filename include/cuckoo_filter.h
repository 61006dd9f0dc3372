/*
 * cuckoo_filter.h -- C ABI of the MI355X-native cuckoo filter (libcuckoo_amd.so).
 *
 * The reference (tiltedtoast/cuckoo-filter) is a header-only C++ template
 * library with no C ABI; its boundary is CuckooFilter<Config>
 * (include/CuckooFilter.cuh:139-1117).  This header exposes that same
 * boundary for fixed instantiations selected at run time by cf_config, so
 * that non-C++ hosts (ctypes, cgo, JNI, N-API) can bind it.  C++ callers can
 * instead include cuckoo-filter_amd/csrc/cuckoo/CuckooFilter.hpp, the
 * template drop-in.
 *
 * Conventions
 *   - All key / output pointers marked d_ are device pointers on the device
 *     that was current when the filter was created; h_ pointers are host.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Every operation
 *     is synchronous on that stream on return, like the reference's methods.
 *   - Functions return CF_OK (0) or a negative CF_ERR_* code; the message of
 *     the last failure on the calling thread is cf_last_error().  (The C++
 *     template API keeps the reference's print-and-exit convention instead.)
 *   - Insert / delete report the TOTAL number of occupied slots after the
 *     call through *occupied (CuckooFilter.cuh:463,547), not a per-call count.
 */
#ifndef CUCKOO_FILTER_AMD_H
#define CUCKOO_FILTER_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CF_OK 0
#define CF_ERR_INVALID (-1)     /* bad argument or configuration */
#define CF_ERR_UNSUPPORTED (-2) /* configuration valid but not built into this library */
#define CF_ERR_HIP (-3)         /* HIP runtime / launch error */
#define CF_ERR_ALLOC (-4)       /* device or host allocation failed */

enum cf_policy { CF_POLICY_XOR = 0, CF_POLICY_ADDSUB = 1, CF_POLICY_OFFSET = 2 };
enum cf_eviction { CF_EVICT_BFS = 0, CF_EVICT_DFS = 1 };

/* Run-time form of CuckooConfig<T, bitsPerTag, maxEvictions, blockSize,
 * bucketSize, AltBucketPolicy, evictionPolicy, WordType>
 * (CuckooFilter.cuh:44-74). */
typedef struct cf_config {
    uint32_t key_bytes;     /* sizeof(T): 4 (uint32_t keys) or 8 (uint64_t keys) */
    uint32_t bits_per_tag;  /* 8, 16 or 32 */
    uint32_t max_evictions; /* maxEvictions (default 500) */
    uint32_t block_size;    /* blockSize: threads per workgroup, 1..1024 (default 256) */
    uint32_t bucket_size;   /* bucketSize: slots per bucket, power of two (default 16) */
    uint32_t policy;        /* enum cf_policy (default XOR) */
    uint32_t eviction;      /* enum cf_eviction (default DFS, CuckooFilter.cuh:51) */
    uint32_t word_bytes;    /* sizeof(WordType): 4 or 8 (default 8) */
} cf_config;

typedef struct cf_filter cf_filter;

/* Message of the last failed call on this thread ("" if none). */
const char* cf_last_error(void);

/* Fills cfg with CuckooConfig's defaults for the given key width and tag bits
 * (CuckooFilter.cuh:44-52). */
int cf_config_init(cf_config* cfg, uint32_t key_bytes, uint32_t bits_per_tag);

/* 1 if this library holds kernels for cfg, 0 if not, CF_ERR_INVALID if cfg
 * violates the reference's static_asserts (CuckooFilter.cuh:67-71,147-152,
 * 226-236). */
int cf_config_supported(const cf_config* cfg);

/* calculateNumBuckets (CuckooFilter.cuh:406-408 -> bucket_policies.cuh:80-83,
 * 172-181, 347-350). */
size_t cf_calculate_num_buckets(const cf_config* cfg, size_t capacity);

/* Host-side getCandidateBuckets for n host keys (bucket_policies.cuh:45-61,
 * 126-143, 263-285; a __host__ __device__ static in the reference). */
int cf_candidate_buckets_host(const cf_config* cfg, size_t num_buckets, const void* h_keys, size_t n,
                              uint64_t* h_i1, uint64_t* h_i2, uint32_t* h_fp);

/* explicit CuckooFilter(size_t capacity)  (CuckooFilter.cuh:420-428) */
int cf_create(const cf_config* cfg, size_t capacity, cf_filter** out);

/* ~CuckooFilter()  (CuckooFilter.cuh:435-447) */
int cf_destroy(cf_filter* f);

/* size_t insertMany(const T* d_keys, size_t n, cudaStream_t)  (CuckooFilter.cuh:457-464) */
int cf_insert_many(cf_filter* f, const void* d_keys, size_t n, void* stream, size_t* occupied);

/* No counterpart in the reference.  cf_insert_many for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same insert and result, but every
 * buffer beyond the filter, the keys and the counter read-back comes from a
 * fixed device array of CF_NOALLOC_SCRATCH_BYTES that the library holds on
 * each device it runs on until the process ends; the call allocates nothing.
 * At most CF_NOALLOC_MAX_KEYS keys per call (CF_ERR_INVALID beyond that, or
 * when a table is so large that the insert's slice index would not fit).
 * Calls in one process serialize on a lock for the array; the array keeps the
 * last call's contents, and calls that overlap on one GPU from different
 * processes do not share it (each process has its own). */
#define CF_NOALLOC_MAX_KEYS (1u << 21)
#define CF_NOALLOC_SCRATCH_BYTES (64u << 20)
int cf_insert_many_noalloc(cf_filter* f, const void* d_keys, size_t n, void* stream, size_t* occupied);

/* size_t insertManySorted(const T* d_keys, size_t n, cudaStream_t)  (CuckooFilter.cuh:475-508) */
int cf_insert_many_sorted(cf_filter* f, const void* d_keys, size_t n, void* stream, size_t* occupied);

/* void containsMany(const T* d_keys, size_t n, bool* d_output, cudaStream_t)  (CuckooFilter.cuh:518-523) */
int cf_contains_many(cf_filter* f, const void* d_keys, size_t n, bool* d_output, void* stream);

/* size_t deleteMany(const T* d_keys, size_t n, bool* d_output = nullptr, cudaStream_t)
 * (CuckooFilter.cuh:536-548); d_output may be NULL. */
int cf_delete_many(cf_filter* f, const void* d_keys, size_t n, bool* d_output, void* stream,
                   size_t* occupied);

/* void clear()  (CuckooFilter.cuh:675-683) */
int cf_clear(cf_filter* f);

/* size_t occupiedSlots()  (CuckooFilter.cuh:699-704) */
int cf_occupied_slots(cf_filter* f, size_t* occupied);

/* float loadFactor()  (CuckooFilter.cuh:688-690) */
int cf_load_factor(cf_filter* f, float* load_factor);

/* size_t capacity() / getNumBuckets() / sizeInBytes()  (CuckooFilter.cuh:732-750) */
size_t cf_capacity(const cf_filter* f);
size_t cf_num_buckets(const cf_filter* f);
size_t cf_size_in_bytes(const cf_filter* f);

/* size_t countOccupiedSlots()  (CuckooFilter.cuh:759-785) */
int cf_count_occupied_slots(cf_filter* f, size_t* count);

/* evictionCount() / resetEvictionCount()  (CuckooFilter.cuh:706-726; always
 * counted here, as with CUCKOO_FILTER_COUNT_EVICTIONS). */
int cf_eviction_count(cf_filter* f, size_t* count);
int cf_reset_eviction_count(cf_filter* f);

/* Eviction-chain histogram of subsequent inserts (no reference counterpart;
 * extends the CUCKOO_FILTER_COUNT_EVICTIONS counter, eviction_benchmark.cu).
 * bins = max_evictions + 2: bin c (1 <= c <= max_evictions) counts successful
 * inserts that made c swaps, bin max_evictions+1 counts failed inserts; bin 0
 * is left 0 (callers derive it as keys - sum).  Enabling resets the bins. */
int cf_set_eviction_histogram(cf_filter* f, int enable);
int cf_eviction_histogram(cf_filter* f, uint64_t* h_bins, size_t nbins);

/* Raw table image (numBuckets * sizeof(Bucket) bytes, little-endian words;
 * CuckooFilter.cuh:237,350,748-750) to / from host memory.  Import also sets
 * the occupancy counter. */
int cf_export_table(cf_filter* f, void* h_dst, size_t bytes);
int cf_import_table(cf_filter* f, const void* h_src, size_t bytes, size_t occupied);

/* Copies src's table image and counters into dst on the device (same
 * configuration and bucket count; no reference counterpart). */
int cf_copy_filter(cf_filter* dst, cf_filter* src, void* stream);

/* getCandidateBuckets on the device for n device keys (the device half of the
 * reference's __host__ __device__ static, CuckooFilter.cuh:366-369). */
int cf_compute_candidates(cf_filter* f, const void* d_keys, size_t n, uint64_t* d_i1, uint64_t* d_i2,
                          uint32_t* d_fp, void* stream);

/* computePackedTagsKernel (CuckooFilter.cuh:1185-1207): d_packed receives n
 * PackedTagType values (uint32_t when bits_per_tag <= 8, else uint64_t). */
int cf_compute_packed_tags(cf_filter* f, const void* d_keys, size_t n, void* d_packed, void* stream);

/* The cub::DeviceRadixSort::SortKeys step of insertManySorted
 * (CuckooFilter.cuh:485-498; the reference's presorted benchmark calls it
 * directly, benchmark/cuckoo_sorted_vs_unsorted.cu:57-69): sorts n packed
 * tags of this filter's table in place, ascending.  Every bit such a value
 * can hold (bits_per_tag + ceil(log2(num_buckets)), capped at the
 * PackedTagType width) is ordered, so for values cf_compute_packed_tags
 * produced the result is the full ascending order.  n < 2^32. */
int cf_sort_packed_tags(cf_filter* f, void* d_packed, size_t n, void* stream);

/* insertKernelSorted (CuckooFilter.cuh:1209-1266) over n PackedTagType
 * values, as the reference's InsertPresorted benchmark launches it
 * (benchmark/cuckoo_sorted_vs_unsorted.cu:74-76); *occupied receives the
 * total occupancy afterwards.  The values need not be sorted; sorted ones
 * give neighbouring threads neighbouring buckets. */
int cf_insert_packed_sorted(cf_filter* f, const void* d_packed, size_t n, void* stream, size_t* occupied);

/* ------------------------------------------------------------------------ */
/* Batched many-filter execution (no reference counterpart: replaces a loop   */
/* of the calls above, each of which launches, synchronises and reads the    */
/* counter back -- CuckooFilter.cuh:457-548).                                */
/* ------------------------------------------------------------------------ */

enum cf_op { CF_OP_INSERT = 0, CF_OP_LOOKUP = 1, CF_OP_DELETE = 2 };

typedef struct cf_batch_op {
    cf_filter* filter;    /* target filter */
    uint32_t op;          /* enum cf_op */
    uint32_t reserved;    /* 0 */
    const void* d_keys;   /* device keys of the filter's key width */
    size_t n;             /* number of keys */
    bool* d_output;       /* lookup: required; delete: optional; insert: ignored */
} cf_batch_op;

typedef struct cf_batch cf_batch;

int cf_batch_create(cf_batch** out);
int cf_batch_destroy(cf_batch* b);

/* Executes ops[0..count) with the results of running them one by one in
 * array order through cf_insert_many / cf_contains_many / cf_delete_many:
 * ops on the same filter keep their order, ops on different filters run
 * together.  Every filter must live on the current device and share one
 * configuration apart from max_evictions and eviction (which stay per
 * filter).  Synchronous on `stream`.  h_occupied (optional, count entries)
 * receives each op's post-op occupancy -- what the plain call returns. */
int cf_batch_execute(cf_batch* b, const cf_batch_op* ops, size_t count, void* stream, uint64_t* h_occupied);

/* The same ops run one at a time through cf_insert_many / cf_contains_many /
 * cf_delete_many, in array order, from a C loop: the per-call pattern of the
 * reference's API (launch, synchronise, read the counter back;
 * CuckooFilter.cuh:457-548) without a scripting layer's overhead -- the
 * baseline cf_batch_execute replaces.  h_occupied (optional) receives each
 * op's post-op occupancy (for lookups, the occupancy at that point). */
int cf_run_plain_calls(const cf_batch_op* ops, size_t count, void* stream, uint64_t* h_occupied);

/* ------------------------------------------------------------------------ */
/* Serving loop (no counterpart in the reference; its nearest analogue is    */
/* the synchronous one-worker IPC server, include/CuckooFilterIPC.cuh:200-307) */
/* ------------------------------------------------------------------------ */

/* A server runs requests against resident filters with continuous batching.
 * Requests queue in arrival order; a launcher thread moves everything queued
 * (within the window limits) into the next free window slot and runs it as
 * one batched window (the cf_batch_execute semantics) on the server's
 * stream; a completer thread retires windows in order and publishes the
 * results.  Two window slots keep one window forming while the other runs.
 * Results equal running the requests one by one, in arrival order, through
 * the plain calls (or, under CF_ORDER_PHASED, in the phase order given
 * below).  Requests with CF_REQUEST_HOST_MEMORY pass host buffers:
 * their keys and results travel through a device staging arena, and such a
 * request is admitted to a window only when its bytes fit in the arena
 * (admission waits, in arrival order, for running windows to free space).
 * A filter used by a server must not be used by other calls meanwhile. */
typedef struct cf_server cf_server;

#define CF_REQUEST_HOST_MEMORY 1u /* keys / output are host memory */

typedef struct cf_server_options {
    int32_t device;               /* default: the current device */
    uint32_t max_window_requests; /* default 4096 */
    uint64_t max_window_keys;     /* default 1 << 22 (a larger single request runs alone) */
    uint64_t staging_bytes;       /* device staging arena for host-memory requests;
                                     0 = min(512 MiB, 1/16 of free HBM at creation) */
    uint32_t window_slots;        /* windows formed ahead: 2..8, default 2; 3 overlaps the
                                     staging of host-memory windows with their uploads */
    uint32_t max_window_levels;   /* 0 = unlimited (default); else a window stops taking requests
                                     before one that would need more levels than this (a level
                                     per insert / delete on one filter, and one per run of
                                     lookups between them; phased: at most two), which bounds
                                     a window's duration */
    uint32_t ordering;            /* CF_ORDER_ARRIVAL (default) or CF_ORDER_PHASED */
    uint32_t form_late;           /* 0 (default): the next window is formed as soon as a slot is
                                     free; 1: only once the window launched before it has reached
                                     its last level, so that requests arriving until then join it
                                     (lower latency below saturation, a short gap between windows) */
} cf_server_options;

/* Orderings of the requests on one filter.
 * CF_ORDER_ARRIVAL: results equal running the requests one by one, in
 *   arrival order, through the plain calls.
 * CF_ORDER_PHASED: the requests of one window are concurrent; windows run in
 *   arrival order.  On each filter the window's inserts run first, together
 *   as one concurrent insert, then its lookups and deletes together.  A
 *   lookup therefore sees every insert that arrived before it (and may see
 *   inserts of its window that arrived after it); for a key that a delete of
 *   its own window removes it may report either answer, and every other key
 *   gets the answer it gets after the window's inserts.  Deletes never move
 *   a tag, which is why they may run beside lookups; inserts move tags and
 *   run alone.  Every request's occupied count is the filter's after its
 *   phase.  A window takes at most two levels. */
#define CF_ORDER_ARRIVAL 0u
#define CF_ORDER_PHASED 1u

typedef struct cf_request {
    cf_filter* filter;
    uint32_t op;      /* cf_op */
    uint32_t flags;   /* CF_REQUEST_* */
    const void* keys; /* n keys of the filter's key width */
    uint64_t n;
    bool* output;     /* lookup: required; delete: optional; insert: unused */
} cf_request;

typedef struct cf_request_result {
    int32_t status;         /* CF_OK or a CF_ERR_* code */
    uint32_t window_levels; /* levels of the window it ran in */
    uint64_t window;        /* id of that window (0, 1, ...) */
    uint64_t occupied;      /* filter occupancy after the request (what insertMany /
                               deleteMany return) */
    uint64_t arrival_ns;    /* steady clock (cf_server_now_ns): submission, or the
                               scheduled arrival under cf_server_replay */
    uint64_t start_ns;      /* its window was launched */
    uint64_t done_ns;       /* its window had completed and its results were out */
} cf_request_result;

typedef struct cf_server_stats {
    uint64_t windows;
    uint64_t requests;
    uint64_t keys;
    uint64_t levels;        /* summed over windows */
    uint64_t staging_bytes; /* arena size in use */
    uint64_t max_queue;     /* deepest request queue seen */
} cf_server_stats;

int cf_server_options_init(cf_server_options* o);
int cf_server_create(const cf_server_options* o, cf_server** out);
/* Finishes every submitted request, then stops the threads and frees.  No
 * other call on this server may run concurrently with it or follow it. */
int cf_server_destroy(cf_server* s);
/* Queues one request; *ticket names it for cf_server_wait / cf_server_poll.
 * The request's buffers must stay valid until it has completed. */
int cf_server_submit(cf_server* s, const cf_request* r, uint64_t* ticket);
/* Blocks until the request has completed, fills *out and forgets the ticket. */
int cf_server_wait(cf_server* s, uint64_t ticket, cf_request_result* out);
/* 1 (and *out filled, ticket forgotten) if complete, 0 if still pending. */
int cf_server_poll(cf_server* s, uint64_t ticket, cf_request_result* out);
/* Blocks until every request submitted so far has completed. */
int cf_server_drain(cf_server* s);
int cf_server_stats_get(cf_server* s, cf_server_stats* out);
int cf_server_stats_reset(cf_server* s);
uint64_t cf_server_now_ns(void);
/* Open-loop replay: submits reqs[i] at arrival_s[i] seconds (non-decreasing)
 * after the call starts, from this thread, then waits for all of them;
 * results[i].arrival_ns is the scheduled arrival, so queueing behind a late
 * submission still counts in the latency. */
int cf_server_replay(cf_server* s, const cf_request* reqs, const double* arrival_s, size_t count,
                     cf_request_result* results);

/* ------------------------------------------------------------------------ */
/* File formats (no reference counterpart for the table file; the k-mer key  */
/* file is the reference's src/kmc_to_binary.cpp:1-7 format).                */
/* ------------------------------------------------------------------------ */

/* Table file: cf_table_header (80 bytes, little-endian) followed by the raw
 * table image of num_buckets * sizeof(Bucket) bytes. */
typedef struct cf_table_header {
    char magic[8];          /* "CFAMDTB1" */
    uint32_t header_bytes;  /* sizeof(cf_table_header) = 80 */
    uint32_t reserved;      /* 0 */
    cf_config config;       /* configuration the image was written with */
    uint64_t num_buckets;
    uint64_t occupied;      /* occupancy counter */
    uint64_t evictions;     /* eviction counter */
    uint64_t image_bytes;   /* num_buckets * sizeof(Bucket) */
} cf_table_header;

/* Writes the filter's header and table image to `path`. */
int cf_save_table(cf_filter* f, const char* path);

/* Creates a filter on the current device from a file cf_save_table wrote
 * (same configuration, bucket count, image and counters). */
int cf_load_table(const char* path, cf_filter** out);

/* k-mer key file: uint64 N, then N uint64 keys (little-endian), as written by
 * the reference's src/kmc_to_binary.cpp:92-123 and read by
 * benchmark/kmer_benchmark.cu:26-47. */
int cf_kmer_file_write(const char* path, const uint64_t* h_keys, size_t n);
int cf_kmer_file_count(const char* path, size_t* n);
int cf_kmer_file_read(const char* path, uint64_t* h_keys, size_t n);

/* 2-bit encodes every k-mer (1 <= k <= 32) of an ASCII DNA sequence on the
 * device: d_out[i] = key of seq[i .. i+k-1], len - k + 1 keys.  A=0 C=1 G=2
 * T=3, first base in the highest bits (src/kmc_to_binary.cpp:17-46); with
 * `canonical` the smaller of the k-mer and its reverse complement; a k-mer
 * holding another character gets the key UINT64_MAX.  That marker is also
 * the key of the valid k-mer of 32 T's when k = 32 and `canonical` is 0 (no
 * other k and no canonical key reaches it): callers that must tell the two
 * apart check the sequence for that run of T's. */
int cf_encode_kmers(const void* d_seq, size_t len, uint32_t k, int canonical, uint64_t* d_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CUCKOO_FILTER_AMD_H */
