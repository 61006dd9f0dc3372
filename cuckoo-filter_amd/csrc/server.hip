// Serving loop over resident filters: continuous batching of queued requests
// into windows run by the batched many-filter executor (cuckoo/batch.hpp).
//
// Threads: callers submit into a FIFO queue; the launcher thread forms a
// window from the queue head whenever a window slot is free and launches it
// on the server's stream; the completer thread waits for windows in launch
// order, copies host-memory results out and publishes per-request results.
// One stream runs every window, so windows execute in launch order and the
// batch executor's per-window ordering gives arrival order per filter (or,
// under the phased ordering, the window's two phases: its inserts, then its
// lookups and deletes together).
// Window slots (two by default, window_slots in the options): one window runs
// on the device while the next ones are formed, staged and queued behind it.
// A third slot lets the host copy of one window overlap the upload of the one
// before it, which pays when requests come from host memory.  Host-memory keys
// are copied into the slot's pinned arena by a pool of helper threads when a
// window holds many of them, and go to the device on a copy stream, so that
// the upload of one window overlaps the kernels of the one before it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "capi_internal.hpp"

using namespace cfabi::internal;

namespace {

constexpr int kMaxSlots = 8;  // window slots: one window runs while the others are formed and staged
constexpr size_t kStageAlign = 256;
constexpr size_t kParallelCopyBytes = 8u << 20;  // host copies at least this large use several threads
constexpr unsigned kMaxCopyThreads = 4;  // 4 threads copy about 85 GB/s into pinned memory, 8 about 100

struct CopyJob {
    void* dst;
    const void* src;
    size_t bytes;
};

// Copies byte range [from, to) of the concatenation of jobs.
void copyRange(const std::vector<CopyJob>& jobs, size_t from, size_t to) {
    size_t at = 0;
    for (const CopyJob& j : jobs) {
        const size_t lo = std::max(from, at), hi = std::min(to, at + j.bytes);
        if (lo < hi)
            std::memcpy(static_cast<char*>(j.dst) + (lo - at), static_cast<const char*>(j.src) + (lo - at), hi - lo);
        at += j.bytes;
        if (at >= to) break;
    }
}

// Helper threads, started once, that split large host copies with the
// calling thread: each copies one contiguous share of the concatenated jobs.
// One pool serves one calling thread (run() is not reentrant).  Without
// helpers (none could be started) run() copies alone.
class CopyPool {
   public:
    explicit CopyPool(unsigned helpers) {
        try {
            for (unsigned k = 1; k <= helpers; ++k) threads_.emplace_back([this, k] { loop(k); });
        } catch (...) {
            stopAll();  // keep none rather than a partial set with gaps in the shares
        }
    }
    CopyPool(const CopyPool&) = delete;
    CopyPool& operator=(const CopyPool&) = delete;
    ~CopyPool() { stopAll(); }

    void run(const std::vector<CopyJob>& jobs) {
        size_t total = 0;
        for (const CopyJob& j : jobs) total += j.bytes;
        const unsigned parts = total < kParallelCopyBytes ? 1u : static_cast<unsigned>(threads_.size()) + 1u;
        if (parts == 1) {
            copyRange(jobs, 0, total);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_ = &jobs;
            total_ = total;
            parts_ = parts;
            pending_ = parts - 1;
            ++gen_;
        }
        cv_.notify_all();
        copyRange(jobs, 0, total / parts);  // share 0
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
        jobs_ = nullptr;
    }

   private:
    void loop(unsigned k) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            const std::vector<CopyJob>* jobs = jobs_;
            const size_t from = total_ * k / parts_, to = total_ * (k + 1) / parts_;
            lk.unlock();
            copyRange(*jobs, from, to);
            lk.lock();
            if (--pending_ == 0) done_.notify_one();
        }
    }
    void stopAll() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread& t : threads_) t.join();
        threads_.clear();
    }

    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::vector<CopyJob>* jobs_ = nullptr;
    size_t total_ = 0;
    unsigned parts_ = 1;
    unsigned pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

unsigned copyHelpers() {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    return std::min(kMaxCopyThreads, std::max(1u, hw / 2)) - 1;
}

size_t alignUp(size_t x) { return (x + kStageAlign - 1) & ~(kStageAlign - 1); }


uint64_t nowNs() {
    return static_cast<uint64_t>(
        std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
            .count());
}

// How long the server's threads poll an event by yielding before they fall
// back to sleeping.  A budget in time, not in polls: one yield takes 0.1 us
// on some hosts and a requested 20 us sleep 70 us or more, so a budget of
// 2000 yields ran out inside a window of a few hundred microseconds and the
// window's completion was then noticed up to a sleep late.
constexpr uint64_t kSpinNs = 2000000;

/// Waits until `ev` is no longer pending; returns its final status.
hipError_t waitEvent(hipEvent_t ev) {
    const uint64_t t0 = nowNs();
    hipError_t e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
        if (nowNs() - t0 > kSpinNs) {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        } else {
            std::this_thread::yield();
        }
    }
    return e;
}

struct Pending {
    cf_request req{};
    uint64_t ticket = 0;
    uint64_t arrivalNs = 0;
    size_t stageKeys = 0;    // staging bytes for keys (host-memory requests)
    size_t stageOut = 0;     // staging bytes for results (host-memory requests)
    cf_request_result* sink = nullptr;  // replay: result goes straight here
    uint64_t* sinkCount = nullptr;      // replay: completions counted here
};

struct Slot {
    std::vector<std::pair<const cfabi::Ops*, void*>> executors;  // one per configuration used
    void* dArena = nullptr;  // device staging
    void* hArena = nullptr;  // pinned host staging (same layout)
    hipEvent_t done = nullptr;
    hipEvent_t uploaded = nullptr;  // the window's keys are on the device (copy stream)
    hipEvent_t lastLevel = nullptr;  // the window has reached its last level
    // the window in this slot
    std::vector<Pending> reqs;
    std::vector<cuckoo::BatchOpDesc> descs;
    std::vector<size_t> keyOff, outOff;  // arena offsets per request (host-memory requests)
    const cfabi::Ops* ops = nullptr;
    void* state = nullptr;
    uint64_t id = 0;
    uint64_t startNs = 0;
    int status = CF_OK;
    uint32_t levels = 0;
};

}  // namespace

struct cf_server {
    cf_server_options opt{};
    size_t arenaBytes = 0;  // per slot
    hipStream_t stream = nullptr;
    hipStream_t copyStream = nullptr;  // host-to-device uploads of staged keys
    Slot slots[kMaxSlots];
    int numSlots = 2;

    std::mutex mu;
    std::condition_variable cvWork;      // launcher: queue or free slot changed
    std::condition_variable cvInflight;  // completer: a window was launched
    std::condition_variable cvDone;      // waiters: results published
    std::deque<Pending> queue;
    std::vector<int> freeSlots;
    std::deque<int> inflight;
    std::unordered_set<uint64_t> outstanding;
    std::unordered_map<uint64_t, cf_request_result> finished;
    uint64_t nextTicket = 1;
    uint64_t nextWindow = 0;
    uint64_t submitted = 0;
    uint64_t completed = 0;
    bool stopping = false;
    bool launcherDone = false;
    cf_server_stats stats{};

    CopyPool launcherCopies{copyHelpers()};   // stages keys into the pinned arena
    CopyPool completerCopies{copyHelpers()};  // copies results out of it
    std::thread launcher;
    std::thread completer;

    uint32_t maxLevels = 0;  // levels per window (0 in the options: unlimited)
    bool phased = false;     // CF_ORDER_PHASED
    bool formLate = false;   // form a window once the one before it has reached its last level
    int lastLaunched = -1;   // slot of the window launched last (launcher thread only)
    // launcher's planning scratch: (last mutation level, last level) per filter;
    // phased: the phases present per filter, as a bit mask
    std::unordered_map<const cf_filter*, std::pair<int32_t, int32_t>> depth;
    std::unordered_map<const cf_filter*, uint32_t> kindsSeen;

    void launcherLoop();
    void completerLoop();
    int launchWindow(Slot& sl);
};

namespace {

void* executorFor(Slot& sl, const cfabi::Ops* ops) {
    for (auto& [k, st] : sl.executors) {
        if (k == ops) return &st;
    }
    sl.executors.emplace_back(ops, nullptr);
    return &sl.executors.back().second;
}

}  // namespace

int cf_server::launchWindow(Slot& sl) {
    // host-memory requests: keys into the pinned arena, one upload
    size_t keyBytes = 0;
    sl.keyOff.assign(sl.reqs.size(), 0);
    sl.outOff.assign(sl.reqs.size(), 0);
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        if (sl.reqs[i].stageKeys == 0) continue;
        sl.keyOff[i] = keyBytes;
        keyBytes += sl.reqs[i].stageKeys;
    }
    size_t outBytes = keyBytes;
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        if (sl.reqs[i].stageOut == 0) continue;
        sl.outOff[i] = outBytes;
        outBytes += sl.reqs[i].stageOut;
    }
    auto* h = static_cast<unsigned char*>(sl.hArena);
    auto* d = static_cast<unsigned char*>(sl.dArena);
    sl.descs.resize(sl.reqs.size());
    std::vector<CopyJob> stage;
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        const cf_request& r = sl.reqs[i].req;
        const void* keys = r.keys;
        bool* out = r.output;
        if (r.flags & CF_REQUEST_HOST_MEMORY) {
            const size_t kb = static_cast<size_t>(r.n) * r.filter->cfg.key_bytes;
            if (kb != 0) stage.push_back({h + sl.keyOff[i], r.keys, kb});
            keys = d + sl.keyOff[i];
            out = sl.reqs[i].stageOut != 0 ? reinterpret_cast<bool*>(d + sl.outOff[i]) : nullptr;
        }
        sl.descs[i] = describe(r.filter, r.op, keys, r.n, out);
    }
    launcherCopies.run(stage);
    hipError_t e = hipSuccess;
    if (keyBytes != 0) {
        // upload on the copy stream; the window's kernels wait for it on the main stream
        e = hipMemcpyAsync(d, h, keyBytes, hipMemcpyHostToDevice, copyStream);
        if (e == hipSuccess) e = hipEventRecord(sl.uploaded, copyStream);
        if (e == hipSuccess) e = hipStreamWaitEvent(stream, sl.uploaded, 0);
    }
    void** st = static_cast<void**>(executorFor(sl, sl.ops));
    // Windows differ from one to the next: level by level, each level's grid
    // sized to its work.  Arrival-ordered windows (many small levels) take one
    // key per thread, so that a small request's keys run side by side (a level
    // lasts as long as its slowest workgroup); phased windows (two large
    // levels) take two, which ran them 14% faster (20.5 -> 23.3 Gkeys/s
    // saturated on the k-mer serving workload).
    if (e == hipSuccess) {
        e = sl.ops->batchLaunch(st, sl.descs.data(), sl.descs.size(), stream, false,
                                phased ? cuckoo::kBatchChunk : cuckoo::kBatchChunkLatency, phased, sl.lastLevel);
    }
    sl.state = *st;
    if (e == hipSuccess && outBytes > keyBytes)
        e = hipMemcpyAsync(h + keyBytes, d + keyBytes, outBytes - keyBytes, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipEventRecord(sl.done, stream);
    sl.levels = e == hipSuccess ? sl.ops->batchLevels(sl.state) : 0;
    if (e != hipSuccess) {
        const int rc = hipFail(e, "launching a window");
        (void)hipStreamSynchronize(copyStream);  // nothing of this window may still use the slot
        (void)hipStreamSynchronize(stream);
        return rc;
    }
    return CF_OK;
}

void cf_server::launcherLoop() {
    (void)hipSetDevice(opt.device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cvWork.wait(lk, [&] { return (stopping && queue.empty()) || (!queue.empty() && !freeSlots.empty()); });
        if (queue.empty()) break;  // stopping and drained
        if (formLate && lastLaunched >= 0) {
            // Wait until the window launched last has reached its last level,
            // so that the requests arriving until then join the next window
            // instead of one formed behind it.  Only this thread reuses a
            // slot, so the event stays the one that window recorded.
            const hipEvent_t ev = slots[lastLaunched].lastLevel;
            lk.unlock();
            (void)waitEvent(ev);  // an error surfaces at the window's completion
            lk.lock();
        }
        const int si = freeSlots.back();
        freeSlots.pop_back();
        Slot& sl = slots[si];
        sl.reqs.clear();
        sl.ops = queue.front().req.filter->ops;
        size_t keys = 0;
        size_t stage = 0;
        // the level each request would take in the window (the batch planner's
        // rule: a lookup one level after its filter's last insert / delete, an
        // insert / delete one level after its filter's last request; phased:
        // one level for the filter's inserts, one for its lookups and deletes)
        depth.clear();
        kindsSeen.clear();
        while (!queue.empty() && sl.reqs.size() < opt.max_window_requests) {
            const Pending& p = queue.front();
            if (p.req.filter->ops != sl.ops) break;  // one configuration per window
            if (!sl.reqs.empty() && keys + p.req.n > opt.max_window_keys) break;
            if (stage + p.stageKeys + p.stageOut > arenaBytes) break;  // waits for the arena
            auto [it, fresh] = depth.try_emplace(p.req.filter, std::pair<int32_t, int32_t>{-1, -1});
            (void)fresh;
            auto& [lastMutation, maxLevel] = it->second;
            const bool lookup = p.req.op == CF_OP_LOOKUP;
            if (phased) {
                // bit 0: inserts; bit 1: lookups or deletes
                uint32_t& seen = kindsSeen[p.req.filter];
                const uint32_t kinds = seen | (p.req.op == CF_OP_INSERT ? 1u : 2u);
                if (!sl.reqs.empty() && static_cast<uint32_t>(__builtin_popcount(kinds)) > maxLevels) break;
                seen = kinds;
            } else {
                const int32_t level = lookup ? lastMutation + 1 : maxLevel + 1;
                if (!sl.reqs.empty() && static_cast<uint32_t>(level) >= maxLevels) break;
                if (lookup) {
                    maxLevel = std::max(maxLevel, level);
                } else {
                    lastMutation = maxLevel = level;
                }
            }
            keys += p.req.n;
            stage += p.stageKeys + p.stageOut;
            sl.reqs.push_back(p);
            queue.pop_front();
        }
        sl.id = nextWindow++;
        lastLaunched = si;
        lk.unlock();
        sl.startNs = nowNs();
        sl.status = launchWindow(sl);
        lk.lock();
        stats.windows += 1;
        stats.requests += sl.reqs.size();
        stats.keys += keys;
        stats.levels += sl.levels;
        inflight.push_back(si);
        cvInflight.notify_one();
    }
    launcherDone = true;
    cvInflight.notify_one();
}

void cf_server::completerLoop() {
    (void)hipSetDevice(opt.device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cvInflight.wait(lk, [&] { return !inflight.empty() || launcherDone; });
        if (inflight.empty()) break;
        const int si = inflight.front();
        lk.unlock();
        Slot& sl = slots[si];
        int status = sl.status;
        if (status == CF_OK) {
            const hipError_t e = waitEvent(sl.done);
            if (e != hipSuccess) status = hipFail(e, "window completion");
        }
        const unsigned long long* occ = status == CF_OK ? sl.ops->batchResults(sl.state) : nullptr;
        const auto* h = static_cast<const unsigned char*>(sl.hArena);
        std::vector<cf_request_result> res(sl.reqs.size());
        if (status == CF_OK) {
            std::vector<CopyJob> back;
            for (size_t i = 0; i < sl.reqs.size(); ++i) {
                const Pending& p = sl.reqs[i];
                if (p.stageOut != 0 && p.req.n != 0) back.push_back({p.req.output, h + sl.outOff[i], p.req.n});
            }
            completerCopies.run(back);
        }
        for (size_t i = 0; i < sl.reqs.size(); ++i) {
            const Pending& p = sl.reqs[i];
            cf_request_result& r = res[i];
            r.status = status;
            r.window_levels = sl.levels;
            r.window = sl.id;
            r.occupied = occ != nullptr ? occ[i] : 0;
            r.arrival_ns = p.arrivalNs;
            r.start_ns = sl.startNs;
        }
        const uint64_t t = nowNs();
        lk.lock();
        for (size_t i = 0; i < sl.reqs.size(); ++i) {
            res[i].done_ns = t;
            const Pending& p = sl.reqs[i];
            if (p.sink != nullptr) {
                *p.sink = res[i];
                *p.sinkCount += 1;
            } else {
                outstanding.erase(p.ticket);
                finished.emplace(p.ticket, res[i]);
            }
        }
        completed += sl.reqs.size();
        sl.reqs.clear();
        inflight.pop_front();
        freeSlots.push_back(si);
        cvWork.notify_one();
        cvDone.notify_all();
    }
}

namespace {

int checkRequest(const cf_server* s, const cf_request* r, size_t arenaBytes, size_t* stageKeys, size_t* stageOut) {
    if (r == nullptr) return fail(CF_ERR_INVALID, "request is NULL");
    if (r->filter == nullptr || r->filter->ops == nullptr) return fail(CF_ERR_INVALID, "request filter is NULL");
    if (r->filter->device != s->opt.device)
        return fail(CF_ERR_INVALID, "request filter lives on device %d, the server on %d", r->filter->device,
                    s->opt.device);
    if (r->op > CF_OP_DELETE) return fail(CF_ERR_INVALID, "unknown operation %u", r->op);
    if (r->n != 0 && r->keys == nullptr) return fail(CF_ERR_INVALID, "key pointer is NULL");
    if (r->op == CF_OP_LOOKUP && r->n != 0 && r->output == nullptr)
        return fail(CF_ERR_INVALID, "lookup needs an output");
    *stageKeys = 0;
    *stageOut = 0;
    if (r->flags & ~CF_REQUEST_HOST_MEMORY) return fail(CF_ERR_INVALID, "unknown request flags 0x%x", r->flags);
    if (r->flags & CF_REQUEST_HOST_MEMORY) {
        *stageKeys = alignUp(static_cast<size_t>(r->n) * r->filter->cfg.key_bytes);
        const bool wantsOut = r->op == CF_OP_LOOKUP || (r->op == CF_OP_DELETE && r->output != nullptr);
        *stageOut = wantsOut ? alignUp(static_cast<size_t>(r->n)) : 0;
        if (*stageKeys + *stageOut > arenaBytes)
            return fail(CF_ERR_INVALID,
                        "host-memory request of %llu keys needs %zu staging bytes; a window has %zu (raise "
                        "staging_bytes or pass device memory)",
                        static_cast<unsigned long long>(r->n), *stageKeys + *stageOut, arenaBytes);
    }
    return CF_OK;
}

int enqueue(cf_server* s, const cf_request* r, uint64_t arrivalNs, cf_request_result* sink, uint64_t* sinkCount,
            uint64_t* ticket) {
    size_t sk = 0, so = 0;
    const int rc = checkRequest(s, r, s->arenaBytes, &sk, &so);
    if (rc != CF_OK) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return fail(CF_ERR_INVALID, "server is stopping");
    Pending p;
    p.req = *r;
    p.ticket = s->nextTicket++;
    p.arrivalNs = arrivalNs;
    p.stageKeys = sk;
    p.stageOut = so;
    p.sink = sink;
    p.sinkCount = sinkCount;
    if (sink == nullptr) s->outstanding.insert(p.ticket);
    s->queue.push_back(p);
    s->submitted += 1;
    s->stats.max_queue = std::max<uint64_t>(s->stats.max_queue, s->queue.size());
    if (ticket != nullptr) *ticket = p.ticket;
    s->cvWork.notify_one();
    return CF_OK;
}

void freeServer(cf_server* s) {
    (void)hipSetDevice(s->opt.device);
    for (Slot& sl : s->slots) {
        for (auto& [ops, st] : sl.executors) {
            if (st != nullptr) ops->batchRelease(st);
        }
        if (sl.dArena != nullptr) (void)hipFree(sl.dArena);
        if (sl.hArena != nullptr) (void)hipHostFree(sl.hArena);
        if (sl.done != nullptr) (void)hipEventDestroy(sl.done);
        if (sl.uploaded != nullptr) (void)hipEventDestroy(sl.uploaded);
        if (sl.lastLevel != nullptr) (void)hipEventDestroy(sl.lastLevel);
    }
    if (s->copyStream != nullptr) (void)hipStreamDestroy(s->copyStream);
    if (s->stream != nullptr) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // namespace

extern "C" {

uint64_t cf_server_now_ns(void) { return nowNs(); }

int cf_server_options_init(cf_server_options* o) {
    if (o == nullptr) return fail(CF_ERR_INVALID, "options is NULL");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    o->device = dev;
    o->max_window_requests = 4096;
    o->max_window_keys = 1ull << 22;
    o->staging_bytes = 0;
    o->window_slots = 2;
    o->max_window_levels = 0;
    o->ordering = CF_ORDER_ARRIVAL;
    o->form_late = 0;
    return CF_OK;
}

int cf_server_create(const cf_server_options* o, cf_server** out) {
    if (out == nullptr) return fail(CF_ERR_INVALID, "out is NULL");
    *out = nullptr;
    cf_server_options opt{};
    if (o != nullptr) {
        opt = *o;
    } else {
        cf_server_options_init(&opt);
    }
    if (opt.max_window_requests == 0 || opt.max_window_keys == 0)
        return fail(CF_ERR_INVALID, "window limits must be positive");
    if (opt.window_slots < 2 || opt.window_slots > kMaxSlots)
        return fail(CF_ERR_INVALID, "window_slots must be 2..%d (got %u)", kMaxSlots, opt.window_slots);
    if (opt.ordering != CF_ORDER_ARRIVAL && opt.ordering != CF_ORDER_PHASED)
        return fail(CF_ERR_INVALID, "unknown ordering %u", opt.ordering);
    auto* s = new (std::nothrow) cf_server();
    if (s == nullptr) return fail(CF_ERR_ALLOC, "host allocation failed");
    s->opt = opt;
    s->maxLevels = opt.max_window_levels != 0 ? opt.max_window_levels : UINT32_MAX;
    s->phased = opt.ordering == CF_ORDER_PHASED;
    s->formLate = opt.form_late != 0;
    DeviceGuard g(opt.device);
    hipError_t e = hipSuccess;
    size_t staging = opt.staging_bytes;
    if (staging == 0) {
        size_t freeB = 0, totalB = 0;
        e = hipMemGetInfo(&freeB, &totalB);
        if (e != hipSuccess) {
            delete s;
            return hipFail(e, "reading free device memory");
        }
        staging = std::min<size_t>(512ull << 20, freeB / 16);
    }
    s->numSlots = static_cast<int>(opt.window_slots);
    s->arenaBytes = std::max<size_t>(kStageAlign, (staging / s->numSlots) & ~(kStageAlign - 1));
    s->stats.staging_bytes = s->arenaBytes * s->numSlots;
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->copyStream, hipStreamNonBlocking);
    for (int i = 0; i < s->numSlots && e == hipSuccess; ++i) {
        Slot& sl = s->slots[i];
        e = hipMalloc(&sl.dArena, s->arenaBytes);
        if (e == hipSuccess) e = hipHostMalloc(&sl.hArena, s->arenaBytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.uploaded, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.lastLevel, hipEventDisableTiming);
        s->freeSlots.push_back(s->numSlots - 1 - i);
    }
    if (e != hipSuccess) {
        const int rc = hipFail(e, "server set-up");
        freeServer(s);
        return rc;
    }
    try {
        s->launcher = std::thread([s] { s->launcherLoop(); });
        s->completer = std::thread([s] { s->completerLoop(); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stopping = true;
            if (!s->launcher.joinable()) s->launcherDone = true;
        }
        s->cvWork.notify_all();
        if (s->launcher.joinable()) s->launcher.join();
        s->cvInflight.notify_all();
        if (s->completer.joinable()) s->completer.join();
        freeServer(s);
        return fail(CF_ERR_ALLOC, "starting the server threads failed");
    }
    *out = s;
    return CF_OK;
}

int cf_server_destroy(cf_server* s) {
    if (s == nullptr) return CF_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cvWork.notify_all();
    s->launcher.join();   // drains the queue into windows
    s->completer.join();  // retires every launched window
    freeServer(s);
    return CF_OK;
}

int cf_server_submit(cf_server* s, const cf_request* r, uint64_t* ticket) {
    if (s == nullptr) return fail(CF_ERR_INVALID, "server is NULL");
    return enqueue(s, r, nowNs(), nullptr, nullptr, ticket);
}

int cf_server_wait(cf_server* s, uint64_t ticket, cf_request_result* out) {
    if (s == nullptr) return fail(CF_ERR_INVALID, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->finished.count(ticket) == 0 && s->outstanding.count(ticket) == 0)
        return fail(CF_ERR_INVALID, "unknown ticket %llu", static_cast<unsigned long long>(ticket));
    s->cvDone.wait(lk, [&] { return s->finished.count(ticket) != 0; });
    auto it = s->finished.find(ticket);
    if (out != nullptr) *out = it->second;
    s->finished.erase(it);
    return CF_OK;
}

int cf_server_poll(cf_server* s, uint64_t ticket, cf_request_result* out) {
    if (s == nullptr) return fail(CF_ERR_INVALID, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->finished.find(ticket);
    if (it == s->finished.end()) {
        if (s->outstanding.count(ticket) == 0)
            return fail(CF_ERR_INVALID, "unknown ticket %llu", static_cast<unsigned long long>(ticket));
        return 0;
    }
    if (out != nullptr) *out = it->second;
    s->finished.erase(it);
    return 1;
}

int cf_server_drain(cf_server* s) {
    if (s == nullptr) return fail(CF_ERR_INVALID, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    const uint64_t target = s->submitted;
    s->cvDone.wait(lk, [&] { return s->completed >= target; });
    return CF_OK;
}

int cf_server_stats_get(cf_server* s, cf_server_stats* out) {
    if (s == nullptr || out == nullptr) return fail(CF_ERR_INVALID, "server or out is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return CF_OK;
}

int cf_server_stats_reset(cf_server* s) {
    if (s == nullptr) return fail(CF_ERR_INVALID, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    const uint64_t staging = s->stats.staging_bytes;
    s->stats = cf_server_stats{};
    s->stats.staging_bytes = staging;
    return CF_OK;
}

int cf_server_replay(cf_server* s, const cf_request* reqs, const double* arrival_s, size_t count,
                     cf_request_result* results) {
    if (s == nullptr) return fail(CF_ERR_INVALID, "server is NULL");
    if (count == 0) return CF_OK;
    if (reqs == nullptr || arrival_s == nullptr || results == nullptr)
        return fail(CF_ERR_INVALID, "reqs, arrival_s and results must not be NULL");
    for (size_t i = 0; i < count; ++i) {
        if (!(arrival_s[i] >= 0) || (i > 0 && arrival_s[i] < arrival_s[i - 1]))
            return fail(CF_ERR_INVALID, "arrival_s must be non-negative and non-decreasing (index %zu)", i);
        size_t sk, so;
        const int rc = checkRequest(s, &reqs[i], s->arenaBytes, &sk, &so);
        if (rc != CF_OK) return fail(rc, "request %zu: %s", i, g_lastError.c_str());
    }
    uint64_t done = 0;  // guarded by s->mu
    const uint64_t t0 = nowNs();
    for (size_t i = 0; i < count; ++i) {
        const uint64_t at = t0 + static_cast<uint64_t>(arrival_s[i] * 1e9);
        // sleep only toward arrivals further off than a sleep can overshoot
        for (uint64_t t = nowNs(); t < at; t = nowNs()) {
            if (at - t > 2000000) std::this_thread::sleep_for(std::chrono::nanoseconds(at - t - 1000000));
        }
        const int rc = enqueue(s, &reqs[i], at, &results[i], &done, nullptr);
        if (rc != CF_OK) {
            // the ones already queued still write into results / done: wait for them
            std::unique_lock<std::mutex> lk(s->mu);
            s->cvDone.wait(lk, [&] { return done >= i; });
            return rc;
        }
    }
    std::unique_lock<std::mutex> lk(s->mu);
    s->cvDone.wait(lk, [&] { return done >= count; });
    return CF_OK;
}

}  // extern "C"
