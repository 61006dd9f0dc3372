"""Benchmark of the cuckoo-filter hot path on one MI355X.

Workload (the plain GPU drop-in at HBM scale): a 2^28-slot filter (2^24
buckets x 16 slots x 16-bit tags = 512 MiB, uint64 words, XOR alt-bucket
policy, BFS eviction, maxEvictions 500, 256-thread workgroups) and uniform
seeded uint64 keys resident in HBM.  One step is one pass of the hot path:

    clear() -> insertMany(255,013,683 keys: 0 -> 95% load)
            -> containsMany(2^27 keys, 50% inserted / 50% never inserted)
            -> deleteMany(25% of the inserted keys)

through the C ABI (each call synchronises and reads the counter back, as the
reference's API does).  `value` is the whole step's key throughput in Mkeys/s;
per-operation rates come from HIP events on the launch stream.  `roofline`
is containsKernel's (the step's longest single launch), `roofline_insert` the
insert pipeline's; both count algorithmic bytes per key from the sequential
oracle at the same load trajectory (BYTES_PER_KEY below), and `traffic` is the
HBM bytes per launch read from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
summary in profiles/hbm_traffic_2p28.md.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--no-load-sweep]
                    [--no-load-factor] [--no-capacity-sweep] [--no-eviction-sweep]
                    [--no-policy-sweep] [--cpu-cuckoo-log2 20,24]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "cuckoo-filter_amd"), os.path.join(ROOT, "oracle")]

METRIC = "insert/lookup Mkeys/s at 95% load, 2^28 slots; batched vs plain; serving p99"
CAPACITY = 1 << 28
LOAD = 0.95
N_LOOKUP = 1 << 27
SEED = 42
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_2p28.md")

# Algorithmic bytes per key (uint64 keys, 32-byte buckets, 8-byte words), with
# bucket reads / word writes per key from the sequential oracle filling a
# 2^22-slot table from empty to 95% with BFS (the same load trajectory):
#   insert : 8 + 32 * 1.2118 + 8 * 1.0468        (reads 1.2118 buckets, writes 1.0468 words)
#   lookup : 8 + 1 + 32 * 1.5279                  (50/50 positives/negatives)
#   delete : 8 + 1 + 32 * 1.0550 + 8 * 1.0       (deleting inserted keys)
BYTES_PER_KEY = {
    "insert": 8 + 32 * 1.2118 + 8 * 1.0468,
    "lookup": 8 + 1 + 32 * 1.5279,
    "delete": 8 + 1 + 32 * 1.0550 + 8 * 1.0,
}


def sources_digest():
    """SHA-256 (first 16 hex digits) of the HIP sources the library is built
    from, in path order: what a profile's counters were measured on."""
    import hashlib

    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "cuckoo-filter_amd", "csrc")
    files = []
    for d, _, names in os.walk(csrc):
        files += [os.path.join(d, x) for x in names if x.endswith((".hip", ".hpp", ".h"))]
    for path in sorted(files, key=lambda p: os.path.relpath(p, csrc)):
        h.update(os.path.relpath(path, csrc).encode())
        with open(path, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def _traffic_profile():
    try:
        with open(TRAFFIC_PROFILE) as fh:
            return fh.read()
    except OSError:
        return ""


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE + WRITE_SIZE
    summary (counters scaled by a plain copy of known size), or None."""
    m = re.search(rf"^traffic {re.escape(kernel)} (\d+)\s*$", _traffic_profile(), re.M)
    return int(m.group(1)) if m else None


def traffic_stamp():
    """The git revision and source digest the traffic summary was measured at,
    and whether the sources are still those."""
    text = _traffic_profile()
    rev = re.search(r"^revision (\S+)\s*$", text, re.M)
    dig = re.search(r"^sources (\S+)\s*$", text, re.M)
    return {"traffic_revision": rev.group(1) if rev else None,
            "traffic_sources": dig.group(1) if dig else None,
            "traffic_current": bool(dig) and dig.group(1) == sources_digest()}


def _roofline(kernel, bytes_per_key, n, ms):
    achieved = bytes_per_key * n / (ms / 1e3) / 1e9
    traffic = pmc_traffic(kernel.split()[0])
    return {
        "bound": "hbm",
        "kernel": kernel,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        # physical HBM bytes per second: the committed per-launch traffic over
        # this run's launch time (a random 32-byte read moves a 128-byte line)
        "traffic_gbs": round(traffic / (ms / 1e3) / 1e9, 1) if traffic else None,
        "algorithmic_bytes": int(bytes_per_key * n),
        "bytes_per_key": round(bytes_per_key, 2),
        "ms_per_launch": round(ms, 3),
        **traffic_stamp(),
    }


_T0 = time.perf_counter()


def log(msg):
    print(f"[bench {time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def _events(torch):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def cpu_baseline(n_ins, n_look, n_del, budget_s=25.0, cpu_cuckoo_log2=(20, 24)):
    """The oracle's lock-free multi-threaded CAS insert / lookup / delete on the
    host cores, over a bounded sample: the same 2^28-slot table filled to the
    same 95% load, but only a 1/8 slice of the lookups and deletes."""
    import numpy as np
    import oracle

    threads = int(os.environ.get("CUCKOO_CPU_THREADS", min(16, os.cpu_count() or 1)))
    cfg = oracle.make_config(eviction=oracle.EVICT_BFS)
    f = oracle.Filter(cfg, CAPACITY)
    keys = oracle.gen_keys(SEED, n_ins + N_LOOKUP // 2)
    t0 = time.perf_counter()
    f.insert_many(keys[:n_ins], threads=threads)
    t_ins = time.perf_counter() - t0
    nq = n_look // 8
    q = np.concatenate([keys[: nq // 2], keys[n_ins: n_ins + nq // 2]])
    t0 = time.perf_counter()
    f.contains_many(q, threads=threads)
    t_look = time.perf_counter() - t0
    nd = n_del // 8
    t0 = time.perf_counter()
    f.delete_many(keys[:nd], threads=threads)
    t_del = time.perf_counter() - t0
    keys_done = n_ins + nq + nd
    total = t_ins + t_look + t_del
    del f
    # the CPU-runnable configuration: capacity 2^20 (2^16 buckets), 996,147
    # inserts (95%), 2^20 lookups (50/50), then half of the inserted keys
    # deleted; the sequential restatement and the threaded one
    c1 = {}
    for label, thr in (("sequential", 1), (f"{threads}_threads", threads)):
        g = oracle.Filter(cfg, 1 << 20)
        n1 = int(g.capacity() * 0.95)
        k1 = oracle.gen_keys(SEED + 1, n1 + (1 << 19))
        q1 = np.concatenate([k1[: 1 << 19], k1[n1:]])
        t0 = time.perf_counter()
        g.insert_many(k1[:n1], threads=thr)
        ti = time.perf_counter() - t0
        t0 = time.perf_counter()
        g.contains_many(q1, threads=thr)
        tl = time.perf_counter() - t0
        t0 = time.perf_counter()
        g.delete_many(k1[: n1 // 2], threads=thr)
        td = time.perf_counter() - t0
        c1[label] = {"insert_mkeys_s": round(n1 / ti / 1e6, 2), "lookup_mkeys_s": round(q1.size / tl / 1e6, 2),
                     "delete_mkeys_s": round((n1 // 2) / td / 1e6, 2),
                     "mkeys_s": round((n1 + q1.size + n1 // 2) / (ti + tl + td) / 1e6, 2)}
    # a restatement of the CPU filter the reference compares against
    # (cuckoofilter::CuckooFilter<uint64_t, 16>, benchmark/cuckoo_cpu_vs_gpu.cu:
    # 52-155), on that harness's workload: insert n = 95% of the capacity,
    # look the same n keys up, delete them all; one thread
    cc = {}
    for lg in cpu_cuckoo_log2:
        label, cap = f"capacity_2p{lg}", 1 << lg
        g = oracle.CpuCuckoo(cap)
        n1 = int(cap * 0.95)
        k1 = oracle.gen_keys(SEED + 2, n1)
        t0 = time.perf_counter()
        added = g.add_many(k1)
        ti = time.perf_counter() - t0
        t0 = time.perf_counter()
        found = g.contain_many(k1)
        tl = time.perf_counter() - t0
        t0 = time.perf_counter()
        gone = g.delete_many(k1)
        td = time.perf_counter() - t0
        cc[label] = {"keys": n1, "table_bytes": g.size_bytes(), "added": added, "found": found, "deleted": gone,
                     "insert_mkeys_s": round(n1 / ti / 1e6, 2), "lookup_mkeys_s": round(n1 / tl / 1e6, 2),
                     "delete_mkeys_s": round(n1 / td / 1e6, 2), "mkeys_s": round(3 * n1 / (ti + tl + td) / 1e6, 2)}
        del g
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": round(keys_done / total / 1e6, 2),
        "unit": "Mkeys/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle lock-free CAS filter (cpu, {threads} threads): fill a 2^28-slot table to 95% "
                   f"({n_ins} inserts), then {nq} lookups (50/50) and {nd} deletes"),
        "insert_mkeys_s": round(n_ins / t_ins / 1e6, 2),
        "lookup_mkeys_s": round(nq / t_look / 1e6, 2),
        "delete_mkeys_s": round(nd / t_del / 1e6, 2),
        "seconds": round(total, 2),
        "capacity_2p20": c1,
        "cpu_cuckoo_restatement": {
            "what": ("single-threaded restatement of the published cuckoo filter design behind "
                     "cuckoofilter::CuckooFilter<uint64_t,16> (4 x 16-bit slots per bucket, victim cache), "
                     "not upstream code; insert 95% of capacity, look them up, delete them"),
            **cc},
        "host": {"nproc": os.cpu_count(), "model": model},
    }


def batched_section(cf, torch, workloads, dev, n_filters=256, n_batches=8192, seed=7):
    """Batched many-filter workload: n_filters filters with capacities
    log-uniform in 2^14..2^26 slots, each filled to a uniform-random load in
    0.3..0.95, driven by n_batches batches of log-uniform 32..2^20 keys, ops
    40% insert / 50% lookup (50% positive) / 10% delete, target filter
    uniform.  Inserts are clipped so that no filter passes 95% load.  The same
    window runs as plain per-batch calls and as one cf_batch_execute, each
    from the same starting filter states, first with 16-bit tags, bucket 16,
    u64 words and BFS, then with the configuration the autotuner picks for
    this op mix under the same false-positive budget."""
    import autotune

    default = dict(key_bytes=8, policy="xor", bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="bfs",
                   block_size=256)
    out = autotune.batched_trial(default, n_filters, n_batches, seed, device=dev)
    # stage 1 prunes on a scaled-down copy of the workload (64 filters, 1024
    # batches); stage 2 reruns the shortlist and the default on the full size
    load = 0.95
    budget = autotune.fpr_closed_form(16, 16, load)
    best, results, confirmed = autotune.choose_batched(budget, load=load, n_filters=64, n_batches=1024,
                                                       baseline=default, confirm=(n_filters, n_batches))
    tuned = best if best is not None else default
    axes = ("bits_per_tag", "bucket_size", "word_bytes", "eviction")
    out["autotune"] = {
        "trial": ("stage 1: 64 filters, 1024 batches of the same distributions, every candidate meeting the "
                  "default's FPR at 0.95; stage 2: the 4 fastest and the default on the full window (another "
                  "seed, best of 3), the default kept unless beaten by 3%"),
        "fpr_budget": budget,
        "points": len(results),
        "chosen": {k: tuned[k] for k in axes},
        "stage1_top": [{**{k: kw[k] for k in axes}, "batched_mkeys_s": r["batched_mkeys_s"]}
                       for kw, r in sorted(results, key=lambda x: -x[1]["batched_mkeys_s"])[:5]],
        "stage2": [{**{k: kw[k] for k in axes}, "batched_mkeys_s": r["batched_mkeys_s"]} for kw, r in confirmed],
    }
    if tuned != default:
        t = autotune.batched_trial(tuned, n_filters, n_batches, seed, device=dev)
        out["autotuned"] = {k: t[k] for k in ("plain_mkeys_s", "batched_mkeys_s", "speedup",
                                              "occupancies_match_plain")}
    return out


def serving_section(cf, torch, workloads, dev, n_requests=20000, seed=5, loads=(0.5, 0.7, 1.0),
                    batched_mkeys_s=None, capped_levels=(), host_burst_run=True, orderings=("arrival", "phased")):
    """Serving workload: canonical 31-mers of a 2^28-base ACGT sequence (10%
    of it in duplicated regions), 2-bit packed on the device, in one 2^30-slot
    filter, plus 64 tenant filters of 2^20..2^24 slots (log-uniform, filled to
    U(0.3, 0.8)); all 16-bit tags, bucket 16, BFS.  n_requests requests:
    sizes log-uniform 1..2^18 keys, target the k-mer filter with probability
    1/2 (else a uniform tenant), ops 80% lookup / 15% insert / 5% delete.
    Lookups read a per-filter pool that is half Zipf(1.1)-popular present keys
    and half absent keys; inserts add fresh keys (tenant inserts clipped at
    95% load); deletes remove the oldest present keys.  Keys are resident in
    HBM.  The trace first runs as one burst (every request queued at t = 0)
    to find the saturated throughput, then open-loop with Poisson arrivals at
    each fraction of it in `loads`; every run starts from the same filter
    states.  Latency = completion - scheduled arrival.  The runs repeat for
    each of `orderings`: "arrival" (default server options: results as if
    one by one in arrival order; its figures are the top-level keys) and
    "phased" (a window's requests concurrent: per filter its inserts, then
    its lookups and deletes together; windows formed once the one before has
    reached its last level and capped at 2^24 keys; under the "phased"
    key)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    cfg = cf.make_config(key_bytes=8, bits_per_tag=16, bucket_size=16, eviction="bfs")
    K = 31
    seq_len = 1 << 28
    seq = workloads.dna_sequence(seq_len, dup_fraction=0.1, regions=1024, seed=seed, device=dev)
    kmers = cf.encode_kmers(seq, K, canonical=True)
    del seq
    n_kmers = kmers.numel()
    initial = n_kmers * 3 // 4  # inserted up front; the rest feeds insert requests
    big = cf.CuckooFilter(1 << 30, config=cfg)
    big.insertMany(kmers[:initial])
    absent_mask = (1 << 62) - 1  # 2-bit packed 31-mers use the low 62 bits

    def lookup_pool(present, n_present, size, s):
        hot = present[workloads.zipf_ranks(size // 2, n_present, s=1.1, seed=s, device=dev)]
        cold = workloads.gen_keys_torch(s, size - size // 2, device=dev) & absent_mask
        mix = torch.cat([hot, cold])
        return mix[torch.randperm(size, device=dev)]

    n_tenants = 64
    caps = (2 ** rng.uniform(20, 24, n_tenants)).astype(np.int64)
    filters, pools, lpools, hi, lo, cap_n = [big], [kmers], [], [initial], [0], [n_kmers]
    lpools.append(lookup_pool(kmers, initial, 1 << 24, seed))
    for t in range(n_tenants):
        f = cf.CuckooFilter(int(caps[t]), config=cfg)
        cap = f.capacity()
        pool = workloads.gen_keys_torch(7000 + t, cap, device=dev)
        fill = int(cap * rng.uniform(0.3, 0.8))
        f.insertMany(pool[:fill])
        filters.append(f)
        pools.append(pool)
        lpools.append(lookup_pool(pool, fill, 1 << 20, 7000 + t))
        hi.append(fill)
        lo.append(0)
        cap_n.append(int(0.95 * cap))
    snapshots = []
    for f in filters:
        c = cf.CuckooFilter(f.capacity(), config=cfg)
        c.copyFrom(f)
        snapshots.append(c)
    ring = torch.empty(1 << 27, dtype=torch.bool, device=dev)
    ring_at = 0
    reqs = np.zeros(n_requests, cf.REQUEST_DTYPE)
    kinds = rng.choice(3, size=n_requests, p=[0.15, 0.8, 0.05])
    sizes = (2 ** rng.uniform(0, 18, n_requests)).astype(np.int64)
    target = np.where(rng.random(n_requests) < 0.5, 0, 1 + rng.integers(0, n_tenants, n_requests))
    clipped = 0
    total_keys = 0
    origin = []  # (source tensor, start, n) of each request's keys
    for j in range(n_requests):
        i, kind, n = int(target[j]), int(kinds[j]), int(sizes[j])
        out = 0
        if kind == 1:  # lookup (cf.OP_LOOKUP)
            pool = lpools[i]
            start = int(rng.integers(0, pool.numel() - n + 1))
            keys = pool.data_ptr() + 8 * start
            origin.append((pool, start, n))
            if ring_at + n > ring.numel():
                ring_at = 0
            out = ring.data_ptr() + ring_at
            ring_at += n
        elif kind == 0:  # insert
            m = max(0, min(n, cap_n[i] - hi[i])) if i else min(n, cap_n[0] - hi[0])
            clipped += n - m
            n = m
            keys = pools[i].data_ptr() + 8 * hi[i]
            origin.append((pools[i], hi[i], n))
            hi[i] += n
        else:  # delete the oldest present keys
            n = min(n, hi[i] - lo[i])
            keys = pools[i].data_ptr() + 8 * lo[i]
            origin.append((pools[i], lo[i], n))
            lo[i] += n
        reqs[j] = (filters[i]._h.value, kind, 0, keys, n, out)
        total_keys += n
    op_keys = {name: int(reqs["n"][reqs["op"] == k].sum()) for k, name in enumerate(("insert", "lookup", "delete"))}

    def restore():
        for f, c in zip(filters, snapshots):
            f.copyFrom(c)
        torch.cuda.synchronize()

    def run(srv, arrivals):
        restore()
        srv.reset_stats()
        res = srv.replay(reqs, arrivals)
        assert (res["status"] == 0).all(), "serving request failed"
        st = srv.stats()
        t_first = int(res["arrival_ns"].min())
        span = (int(res["done_ns"].max()) - t_first) / 1e9
        lat = (res["done_ns"] - res["arrival_ns"]).astype(np.float64) / 1e6
        return {
            "mkeys_s": round(total_keys / span / 1e6, 1),
            "requests_s": round(n_requests / span, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "p999_ms": round(float(np.percentile(lat, 99.9)), 3),
            "max_ms": round(float(lat.max()), 3),
            "windows": st["windows"],
            "mean_window_keys": round(st["keys"] / max(st["windows"], 1)),
            "mean_window_levels": round(st["levels"] / max(st["windows"], 1), 2),
            "max_queue": st["max_queue"],
        }

    def host_burst(srv, count=2000):
        """The first `count` requests with keys and results in pageable host
        memory, replayed as one burst from the same starting states: the rate
        includes the staging copies and PCIe in both directions.  Result
        buffers are written once before the burst, as a client reusing its
        buffers would have them, so first-touch page faults stay out of it."""
        sub = reqs[:count].copy()
        keep = []
        for j in range(count):
            pool, start, n = origin[j]
            if n == 0:
                continue
            hk = np.ascontiguousarray(pool[start:start + n].cpu().numpy().view(np.uint64))
            keep.append(hk)
            sub["flags"][j] = cf.CF_REQUEST_HOST_MEMORY
            sub["keys"][j] = hk.ctypes.data
            if int(sub["op"][j]) == 1:
                o = np.empty(n, dtype=bool)
                o.fill(False)
                keep.append(o)
                sub["output"][j] = o.ctypes.data
            else:
                sub["output"][j] = 0
        keys_total = int(sub["n"].sum())
        restore()
        srv.replay(sub, np.zeros(count))  # warm-up: the server's arenas and executors
        restore()
        res = srv.replay(sub, np.zeros(count))
        assert (res["status"] == 0).all(), "host-memory request failed"
        span = (int(res["done_ns"].max()) - int(res["arrival_ns"].min())) / 1e9
        return {"requests": count, "keys": keys_total, "mkeys_s": round(keys_total / span / 1e6, 1),
                "gb_s_over_pcie": round(keys_total * 8 / span / 1e9, 2)}

    out = {
        "workload": ("canonical 31-mers of a 2^28-base sequence (10% duplicated regions) in a 2^30-slot filter "
                     "+ 64 tenants of 2^20..2^24 slots; requests log-uniform 1..2^18 keys, 80/15/5 "
                     "lookup/insert/delete, lookups Zipf(1.1) present + 50% absent; keys resident in HBM"),
        "requests": n_requests,
        "keys": total_keys,
        "keys_by_op": op_keys,
        "insert_keys_clipped": clipped,
        "kmers": n_kmers,
    }
    mean_keys = total_keys / n_requests

    def poisson(mkeys_s):
        rate = mkeys_s * 1e6 / mean_keys  # requests per second
        arr = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
        return arr - arr[0]

    def sweep(srv, into):
        run(srv, np.zeros(n_requests))  # warm-up: plans, graphs, arenas
        sat = run(srv, np.zeros(n_requests))
        into["saturated"] = sat
        for frac in loads:
            offered = frac * sat["mkeys_s"]
            into[f"load_{int(frac * 100)}pct"] = dict(run(srv, poisson(offered)), offered_mkeys_s=round(offered, 1))
        if batched_mkeys_s:
            # the load the serving target names: 70% of the batched window's throughput
            offered = 0.7 * batched_mkeys_s
            into["load_70pct_of_batched"] = dict(run(srv, poisson(offered)), offered_mkeys_s=round(offered, 1))

    if "arrival" in orderings:
        with cf.Server() as srv:
            sweep(srv, out)
    if "phased" in orderings:
        opts = dict(ordering="phased", form_late=True, max_window_keys=1 << 24)
        part = {"options": {k: v for k, v in opts.items()}}
        with cf.Server(**opts) as srv:
            sweep(srv, part)
        out["phased"] = part
    for cap in capped_levels:
        # windows capped at `cap` levels: shorter windows, so a request waits
        # for less work ahead of it
        part = {"max_window_levels": cap}
        with cf.Server(max_window_levels=cap) as srv:
            sweep(srv, part)
        out[f"max_window_levels_{cap}"] = part
    if host_burst_run:
        with cf.Server(window_slots=3) as srv3:  # a third slot overlaps host copies with uploads
            out["host_memory_burst"] = dict(host_burst(srv3), window_slots=3)
    for x in filters + snapshots:
        x.close()
    del kmers, pools, lpools, ring
    return out


def big_table_section(cf, torch, workloads, dev, log2_slots=30, reps=3):
    """The plain step on a table eight times the memory-side cache: 2^30 slots
    (2^26 buckets x 32 B = 2 GiB), 16-bit tags, bucket 16, u64 words, BFS.
    insertMany to 95% load (1,020,054,732 keys), containsMany of 2^27 keys
    (50/50) and deleteMany of a quarter of the inserted keys, each the mean
    of `reps` timed runs after one untimed, with HIP events on the launch
    stream.  Algorithmic bytes per key as for the 2^28-slot step."""
    cap = 1 << log2_slots
    f = cf.CuckooFilter(cap, key_bytes=8, bits_per_tag=16, bucket_size=16, word_bytes=8, policy="xor",
                        eviction="bfs", max_evictions=500, block_size=256)
    n = int(cap * LOAD)
    nd = n // 4
    log(f"2^{log2_slots}-slot table: {n} inserts, {N_LOOKUP} lookups, {nd} deletes")
    keys = workloads.gen_keys_torch(SEED + 1, n + N_LOOKUP // 2, device=dev)
    query = torch.cat([keys[: N_LOOKUP // 2], keys[n:]])
    out = torch.empty(N_LOOKUP, dtype=torch.bool, device=dev)
    dout = torch.empty(nd, dtype=torch.bool, device=dev)
    t = {"insert": [], "lookup": [], "delete": []}
    occ = rem = 0
    for i in range(reps + 1):
        f.clear()
        ev = {k: _events(torch) for k in t}
        ev["insert"][0].record()
        occ = f.insertMany(keys[:n])
        ev["insert"][1].record()
        ev["lookup"][0].record()
        f.containsMany(query, out=out)
        ev["lookup"][1].record()
        ev["delete"][0].record()
        rem = f.deleteMany(keys[:nd], out=dout)
        ev["delete"][1].record()
        torch.cuda.synchronize()
        if i:
            for k in t:
                t[k].append(ev[k][0].elapsed_time(ev[k][1]))
    half = N_LOOKUP // 2
    res = {"slots": cap, "table_bytes": f.sizeInBytes(), "keys_inserted": n, "lookups": N_LOOKUP,
           "deletes": nd, "placed": occ, "after_delete": rem,
           "positives_found": int(out[:half].sum().item()), "false_positives": int(out[half:].sum().item()),
           "deletes_ok": bool(dout.all().item())}
    units = {"insert": n, "lookup": N_LOOKUP, "delete": nd}
    for k, v in t.items():
        ms = sum(v) / len(v)
        res[f"{k}_ms"] = round(ms, 3)
        res[f"{k}_mkeys_s"] = round(units[k] / ms / 1e3, 1)
        res[f"{k}_algorithmic_gbs"] = round(BYTES_PER_KEY[k] * units[k] / (ms / 1e3) / 1e9, 1)
    # one more fill, untimed, with the eviction-chain histogram on: swaps made
    # -> inserts placed with that many (bin 0 = placed without eviction)
    f.clear()
    f.setEvictionHistogram(True)
    placed = f.insertMany(keys[:n])
    hist = f.evictionHistogram()
    f.setEvictionHistogram(False)
    evicted = int(hist[1:-1].sum())
    res["eviction_chains"] = {
        "placed": placed, "failed": int(hist[-1]),
        "histogram": {"0": placed - evicted, **{str(c): int(hist[c]) for c in range(1, hist.size - 1) if hist[c]}},
    }
    del keys, query, out, dout
    f.close()
    torch.cuda.empty_cache()
    return res


def capacity_sweep_section(cf, torch, keys, dev, log2_caps=range(16, 29), iters=10):
    """The reference's benchmark matrix (benchmark/benchmark_common.cuh:247-384,
    benchmark/cuckoo_sorted_vs_unsorted.cu:18-106): CuckooConfig<uint64_t, 16,
    500, 256, 16, XorAltBucketPolicy> with the config's default eviction (DFS)
    at capacities 2^16..2^28, n = floor(capacity x 0.95) keys.  Per capacity,
    each figure is the mean of `iters` timed calls after one untimed, HIP
    events around the synchronous call:
      insert            clear (untimed), insertMany(n keys)
      insert_sorted     clear (untimed), insertManySorted(n keys)
      insert_presorted  clear (untimed), insertKernelSorted over packed tags
                        computed and sorted once beforehand
      query             containsMany of the same n keys on the filled table
      delete            clear + insertMany (untimed), deleteMany of the n keys
                        with per-key results
    """
    n_max = int((1 << max(log2_caps)) * LOAD)
    res = torch.empty(n_max, dtype=torch.bool, device=dev)
    rows = {}
    for lg in log2_caps:
        cap = 1 << lg
        n = int(cap * LOAD)
        k = keys[:n]
        f = cf.CuckooFilter(cap, key_bytes=8, bits_per_tag=16, bucket_size=16, word_bytes=8, policy="xor",
                            eviction="dfs", max_evictions=500, block_size=256)
        packed = f.sortPackedTags(f.computePackedTags(k))
        out = res[:n]

        def timed(call, prep):
            ts = []
            for i in range(iters + 1):
                prep()
                torch.cuda.synchronize()
                ev = _events(torch)
                ev[0].record()
                call()
                ev[1].record()
                torch.cuda.synchronize()
                if i:
                    ts.append(ev[0].elapsed_time(ev[1]))
            return sum(ts) / len(ts)

        def fill():
            f.clear()
            f.insertMany(k)

        ms = {
            "insert": timed(lambda: f.insertMany(k), f.clear),
            "insert_sorted": timed(lambda: f.insertManySorted(k), f.clear),
            "insert_presorted": timed(lambda: f.insertPackedSorted(packed), f.clear),
        }
        fill()
        ms["query"] = timed(lambda: f.containsMany(k, out=out), lambda: None)
        found = int(out.sum().item())
        ms["delete"] = timed(lambda: f.deleteMany(k, out=out), fill)
        row = {"keys": n, "table_bytes": f.sizeInBytes(), "query_found": found}
        for op, t in ms.items():
            row[f"{op}_ms"] = round(t, 4)
            row[f"{op}_mkeys_s"] = round(n / t / 1e3, 1)
        rows[str(lg)] = row
        del packed
        f.close()
    del res
    torch.cuda.empty_cache()
    return {"config": "CuckooConfig<uint64_t, 16, 500, 256, 16, XorAltBucketPolicy> (DFS), load 0.95",
            "iterations": iters, "by_log2_capacity": rows}


def load_factor_section(cf, torch, keys, dev, loads=range(5, 100, 5), iters=3):
    """The reference's load-factor benchmark (benchmark/load_factor_benchmark.cu:
    36, 44, 345-401): CuckooConfig<uint64_t, 16, 500, 128, 16,
    XorAltBucketPolicy> (DFS) at 2^28 slots, n = floor(2^28 x load) keys for
    load = 0.05..0.95 in steps of 0.05.  At this n the reference's
    adaptiveInsert (benchmark/benchmark_common.cuh:114-126) is insertMany.
    Per load, each figure is the mean of `iters` timed calls after one
    untimed, HIP events around the synchronous call:
      insert          clear (untimed), insertMany(n keys)
      query           containsMany of the same n keys on the filled table
      query_negative  containsMany of n keys never inserted (the next n keys
                      of the same bijective stream, so disjoint from the n)
      delete          clear + insertMany (untimed), deleteMany of the n keys
    """
    n_max = int(CAPACITY * max(loads) / 100)
    assert keys.numel() >= n_max
    import workloads

    neg_all = workloads.gen_keys_torch(SEED, n_max, start=keys.numel(), device=dev)
    res = torch.empty(n_max, dtype=torch.bool, device=dev)
    f = cf.CuckooFilter(CAPACITY, key_bytes=8, bits_per_tag=16, bucket_size=16, word_bytes=8, policy="xor",
                        eviction="dfs", max_evictions=500, block_size=128)
    rows = {}
    for pct in loads:
        n = int(CAPACITY * pct / 100)
        k, neg, out = keys[:n], neg_all[:n], res[:n]

        def timed(call, prep):
            ts = []
            for i in range(iters + 1):
                prep()
                torch.cuda.synchronize()
                ev = _events(torch)
                ev[0].record()
                call()
                ev[1].record()
                torch.cuda.synchronize()
                if i:
                    ts.append(ev[0].elapsed_time(ev[1]))
            return sum(ts) / len(ts)

        def fill():
            f.clear()
            f.insertMany(k)

        placed = []
        ms = {"insert": timed(lambda: placed.append(f.insertMany(k)), f.clear)}
        fill()
        ms["query"] = timed(lambda: f.containsMany(k, out=out), lambda: None)
        found = int(out.sum().item())
        ms["query_negative"] = timed(lambda: f.containsMany(neg, out=out), lambda: None)
        false_pos = int(out.sum().item())
        ms["delete"] = timed(lambda: f.deleteMany(k, out=out), fill)
        row = {"keys": n, "placed": placed[-1], "query_found": found,
               "false_positive_rate": round(false_pos / n, 7)}
        for op, t in ms.items():
            row[f"{op}_ms"] = round(t, 4)
            row[f"{op}_mkeys_s"] = round(n / t / 1e3, 1)
        rows[f"{pct / 100:.2f}"] = row
    f.close()
    del neg_all, res
    torch.cuda.empty_cache()
    return {"config": "CuckooConfig<uint64_t, 16, 500, 128, 16, XorAltBucketPolicy> (DFS), 2^28 slots",
            "iterations": iters, "by_load": rows}


def policy_sweep_section(cf, torch, keys, n_pos, dev, iters=3):
    """The reference's alternate-bucket policy comparison
    (benchmark/bucket_policy_benchmark.cu): CuckooConfig<uint64_t, 16, 500,
    128, 16, Xor | AddSub | Offset> at 2^28 slots and 95% load.  Per policy:
    insertMany of the n keys (clear untimed), containsMany of the same n keys
    (all present) and of 2^26 keys never inserted, and deleteMany of the n keys
    (refill untimed); each the mean of `iters` timed calls after one untimed.
    At 2^28 slots every policy gets 2^24 buckets, so the load is 95% for
    each."""
    n = n_pos
    neg = keys[n:n + (1 << 26)]
    res = torch.empty(n, dtype=torch.bool, device=dev)
    rows = {}
    for pol in ("xor", "addsub", "offset"):
        f = cf.CuckooFilter(CAPACITY, key_bytes=8, bits_per_tag=16, bucket_size=16, word_bytes=8, policy=pol,
                            eviction="dfs", max_evictions=500, block_size=128)
        k = keys[:n]

        def timed(call, prep):
            ts = []
            for i in range(iters + 1):
                prep()
                torch.cuda.synchronize()
                ev = _events(torch)
                ev[0].record()
                call()
                ev[1].record()
                torch.cuda.synchronize()
                if i:
                    ts.append(ev[0].elapsed_time(ev[1]))
            return sum(ts) / len(ts)

        def fill():
            f.clear()
            f.insertMany(k)

        ms = {"insert": timed(lambda: f.insertMany(k), f.clear)}
        fill()
        placed = f.occupiedSlots()
        ms["query"] = timed(lambda: f.containsMany(k, out=res), lambda: None)
        found = int(res.sum().item())
        ms["query_negative"] = timed(lambda: f.containsMany(neg, out=res[:neg.numel()]), lambda: None)
        fpr = int(res[:neg.numel()].sum().item()) / neg.numel()
        ms["delete"] = timed(lambda: f.deleteMany(k, out=res), fill)
        units = {"insert": n, "query": n, "query_negative": neg.numel(), "delete": n}
        row = {"buckets": f.getNumBuckets(), "load": round(n / (f.getNumBuckets() * 16), 4), "placed": placed,
               "query_found": found, "false_positive_rate": round(fpr, 7)}
        for op, t in ms.items():
            row[f"{op}_mkeys_s"] = round(units[op] / t / 1e3, 1)
        rows[pol] = row
        f.close()
    del res
    torch.cuda.empty_cache()
    return {"config": "CuckooConfig<uint64_t, 16, 500, 128, 16, Xor | AddSub | Offset> (DFS), 2^28 slots",
            "keys": n, "iterations": iters, "by_policy": rows}


EVICTION_LOADS = (0.76, 0.78, 0.80, 0.82, 0.84, 0.86, 0.88, 0.90, 0.92, 0.94, 0.95, 0.96, 0.97, 0.98, 0.99,
                  0.995, 0.999)


def eviction_sweep_section(cf, torch, keys, dev, iters=3):
    """The reference's eviction study (benchmark/eviction_benchmark.cu:14-166):
    CuckooConfig<uint64_t, 16, 500, 256, 16, XorAltBucketPolicy, BFS | DFS> at
    2^28 slots; per load factor L of 0.76..0.999, 75% of the L x capacity keys
    are inserted untimed, the eviction counter is reset, and the other 25% are
    inserted by one timed insertMany.  Per point: that insert's Mkeys/s, the
    evictions it made (swaps, as the reference counts them) and per key, and
    the keys it failed to place; the mean of `iters` runs after one untimed."""
    rows = {}
    for ev in ("bfs", "dfs"):
        f = cf.CuckooFilter(CAPACITY, key_bytes=8, bits_per_tag=16, bucket_size=16, word_bytes=8, policy="xor",
                            eviction=ev, max_evictions=500, block_size=256)
        for load in EVICTION_LOADS:
            total = int(CAPACITY * load)
            pre = int(total * 0.75)
            measured = total - pre
            ts, evs, fails = [], [], []
            for i in range(iters + 1):
                f.clear()
                occ_pre = f.insertMany(keys[:pre])
                f.resetEvictionCount()
                torch.cuda.synchronize()
                e = _events(torch)
                e[0].record()
                occ = f.insertMany(keys[pre:total])
                e[1].record()
                torch.cuda.synchronize()
                if i:
                    ts.append(e[0].elapsed_time(e[1]))
                    evs.append(f.evictionCount())
                    fails.append(measured - (occ - occ_pre))
            ms = sum(ts) / len(ts)
            mean_ev = sum(evs) / len(evs)
            rows.setdefault(ev, {})[str(load)] = {
                "measured_keys": measured, "mkeys_s": round(measured / ms / 1e3, 1), "ms": round(ms, 3),
                "evictions": int(mean_ev), "evictions_per_insert": round(mean_ev / measured, 4),
                "failed": int(sum(fails) / len(fails)),
            }
        f.close()
    torch.cuda.empty_cache()
    return {"config": "CuckooConfig<uint64_t, 16, 500, 256, 16, XorAltBucketPolicy, BFS | DFS>, 2^28 slots, "
                      "prefill 75% of each load untimed, the other 25% timed",
            "iterations": iters, "by_policy": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cuckoo-log2", default="20,24",
                    help="log2 capacities for the single-threaded cpu-cuckoo baseline (28 takes about a minute)")
    ap.add_argument("--no-load-sweep", action="store_true",
                    help="skip the insertMany / insertManySorted / lookup / delete sweep over load factors")
    ap.add_argument("--no-roofline", action="store_true", help="skip the measured gather/CAS roofline")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched many-filter comparison")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop latency runs")
    ap.add_argument("--no-big-table", action="store_true", help="skip the 2^30-slot (2 GiB) table")
    ap.add_argument("--no-load-factor", action="store_true",
                    help="skip the reference's load-factor benchmark (loads 0.05..0.95, 2^28 slots)")
    ap.add_argument("--no-capacity-sweep", action="store_true",
                    help="skip the reference's benchmark matrix over capacities 2^16..2^28")
    ap.add_argument("--no-eviction-sweep", action="store_true",
                    help="skip the BFS / DFS eviction study over load factors 0.76..0.999")
    ap.add_argument("--no-policy-sweep", action="store_true",
                    help="skip the XOR / AddSub / Offset alternate-bucket policy comparison")
    ap.add_argument("--big-table-only", action="store_true",
                    help="run only the 2^30-slot table's insert / lookup / delete (for profiling it alone)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the cuckoo-filter hot path has no multi-GPU leg in this build "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    import torch

    import cuckoo_amd as cf
    import workloads

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.big_table_only:
        print(json.dumps({"table_2gib": big_table_section(cf, torch, workloads, dev)}), flush=True)
        return 0
    n_ins = int(CAPACITY * LOAD)
    n_del = n_ins // 4
    f = cf.CuckooFilter(CAPACITY, key_bytes=8, bits_per_tag=16, bucket_size=16, word_bytes=8,
                        policy="xor", eviction="bfs", max_evictions=500, block_size=256)
    assert f.getNumBuckets() == 1 << 24
    keys = workloads.gen_keys_torch(SEED, n_ins + N_LOOKUP // 2, device=dev)
    pos = keys[:n_ins]
    query = torch.cat([keys[: N_LOOKUP // 2], keys[n_ins:]])
    dels = keys[:n_del]
    out = torch.empty(N_LOOKUP, dtype=torch.bool, device=dev)
    dout = torch.empty(n_del, dtype=torch.bool, device=dev)
    torch.cuda.synchronize()

    times = {"insert": [], "lookup": [], "delete": []}
    checks = {}

    def step(record):
        f.clear()
        e = {k: _events(torch) for k in times}
        e["insert"][0].record()
        occ = f.insertMany(pos)
        e["insert"][1].record()
        e["lookup"][0].record()
        f.containsMany(query, out=out)
        e["lookup"][1].record()
        e["delete"][0].record()
        rem = f.deleteMany(dels, out=dout)
        e["delete"][1].record()
        if record:
            torch.cuda.synchronize()
            for k in times:
                times[k].append(e[k][0].elapsed_time(e[k][1]))
        checks["occupied"] = occ
        checks["remaining"] = rem

    log(f"keys resident ({n_ins + N_LOOKUP // 2} u64); warmup {args.warmup}, timed {args.steps}")
    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the last step's lookups (on the filled table): first half inserted keys, second half never inserted
    half = N_LOOKUP // 2
    checks["positives_found"] = int(out[:half].sum().item())
    checks["false_positives"] = int(out[half:].sum().item())

    ms_step = wall * 1e3 / args.steps
    keys_per_step = n_ins + N_LOOKUP + n_del
    value = keys_per_step / (ms_step / 1e3) / 1e6
    avg = {k: sum(v) / len(v) for k, v in times.items()}
    per_op = {
        "insert_mkeys_s": n_ins / avg["insert"] / 1e3,
        "lookup_mkeys_s": N_LOOKUP / avg["lookup"] / 1e3,
        "delete_mkeys_s": n_del / avg["delete"] / 1e3,
        "insert_ms": avg["insert"], "lookup_ms": avg["lookup"], "delete_ms": avg["delete"],
    }
    units = {"insert": n_ins, "lookup": N_LOOKUP, "delete": n_del}
    for op in ("insert", "lookup", "delete"):
        gbs = BYTES_PER_KEY[op] * units[op] / (avg[op] / 1e3) / 1e9
        per_op[f"{op}_algorithmic_gbs"] = round(gbs, 1)
    # containsKernel is the step's longest single launch (one launch per
    # containsMany); insertMany is a pipeline of several launches and is
    # reported beside it with the same byte accounting.
    roofline = _roofline("containsKernel", BYTES_PER_KEY["lookup"], N_LOOKUP, avg["lookup"])
    roofline_insert = _roofline("insertMany pipeline", BYTES_PER_KEY["insert"], n_ins, avg["insert"])
    roofline_insert["kernels"] = ("two-level partition (wide-tile count + scatter over hashed keys, then "
                                  "per-partition count + scatter), binInsertKernel, secondaryKernel, "
                                  "bfsStageKernel, dfsStageKernel")

    log(f"plain steps done: {ms_step:.2f} ms/step")
    # The batched window and the serving loop run next, before the load sweep
    # and the 2^30-slot table: after those two sections' allocations (about
    # 14 GiB of keys and tables) serving latency below saturation rose several
    # fold in the same process (p99 at half load 0.42 -> 0.84-4.1 ms) while
    # every throughput stayed the same.
    batched = serving = None
    if not args.no_batched:
        log("batched many-filter window vs plain calls")
        batched = batched_section(cf, torch, workloads, dev)
        torch.cuda.empty_cache()
    if not args.no_serving:
        log("serving loop: saturated and open-loop latency")
        batched_rate = (batched or {}).get("batched_mkeys_s")
        serving = serving_section(cf, torch, workloads, dev, batched_mkeys_s=batched_rate)
        torch.cuda.empty_cache()
    load_sweep = None
    if not args.no_load_sweep:
        # The same table filled from empty to each load factor, by insertMany and
        # then by insertManySorted (each the mean of 3 timed calls after one
        # untimed), then 2^27 lookups (half inserted keys, half never inserted)
        # and a delete of a quarter of the inserted keys, on the filled table.
        log("load sweep: alpha 0.5 / 0.8 / 0.9 / 0.95")
        load_sweep = {}
        negatives = keys[n_ins:]
        for alpha in (0.5, 0.8, 0.9, 0.95):
            n_a = int(CAPACITY * alpha)
            ka = keys[:n_a]
            row = {"keys": n_a}
            for name, ins in (("insert", f.insertMany), ("insert_sorted", f.insertManySorted)):
                ts = []
                for i in range(4):
                    f.clear()
                    ev = _events(torch)
                    ev[0].record()
                    occ_a = ins(ka)
                    ev[1].record()
                    torch.cuda.synchronize()
                    if i:
                        ts.append(ev[0].elapsed_time(ev[1]))
                row[f"{name}_mkeys_s"] = round(n_a / (sum(ts) / len(ts)) / 1e3, 1)
                row[f"{name}_placed"] = occ_a
            q = torch.cat([ka[: N_LOOKUP // 2], negatives])
            ev = _events(torch)
            ev[0].record()
            f.containsMany(q, out=out)
            ev[1].record()
            torch.cuda.synchronize()
            row["lookup_mkeys_s"] = round(N_LOOKUP / ev[0].elapsed_time(ev[1]) / 1e3, 1)
            row["lookup_positives_found"] = bool(out[: N_LOOKUP // 2].all().item())
            nd_a = n_a // 4
            ev[0].record()
            f.deleteMany(ka[:nd_a], out=dout[:nd_a])
            ev[1].record()
            torch.cuda.synchronize()
            row["delete_mkeys_s"] = round(nd_a / ev[0].elapsed_time(ev[1]) / 1e3, 1)
            load_sweep[str(alpha)] = row
        per_op["insert_sorted_mkeys_s"] = load_sweep["0.95"]["insert_sorted_mkeys_s"]

    load_factor = None
    if not args.no_load_factor:
        log("load factor: loads 0.05..0.95, insert / query / query negative / delete of all n keys")
        load_factor = load_factor_section(cf, torch, keys, dev)

    capacity_sweep = None
    if not args.no_capacity_sweep:
        log("capacity sweep: 2^16..2^28 slots, insert / sorted / presorted / query / delete")
        capacity_sweep = capacity_sweep_section(cf, torch, keys, dev)

    eviction_sweep = None
    if not args.no_eviction_sweep:
        log("eviction sweep: BFS / DFS at loads 0.76..0.999, 2^28 slots")
        eviction_sweep = eviction_sweep_section(cf, torch, keys, dev)

    policy_sweep = None
    if not args.no_policy_sweep:
        log("policy sweep: XOR / AddSub / Offset at 2^28 slots")
        policy_sweep = policy_sweep_section(cf, torch, keys, n_ins, dev)

    big_table = None
    if not args.no_big_table:
        big_table = big_table_section(cf, torch, workloads, dev)

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": ("plain drop-in at HBM scale: 2^28 slots (2^24 buckets x 16 x 16-bit tags, u64 words, "
                         "XOR, BFS, maxEvictions 500, block 256); per step clear + insertMany 255,013,683 u64 "
                         "keys (0->95% load) + containsMany 2^27 (50/50) + deleteMany 25%"),
            "capacity_slots": CAPACITY,
            "load_factor": LOAD,
            "keys_per_step": keys_per_step,
        },
        "breakdown": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in per_op.items()},
        "roofline": roofline,
        "roofline_insert": roofline_insert,
        "checks": {"occupied_after_insert": checks.get("occupied"), "expected": n_ins,
                   "occupied_after_delete": checks.get("remaining"),
                   "lookup_positives_found": checks.get("positives_found"), "lookup_positives": N_LOOKUP // 2,
                   "false_positive_rate": round(checks.get("false_positives", 0) / (N_LOOKUP // 2), 7),
                   # 2 * bucketSize * alpha / 2^bitsPerTag (src/main.cu:114-115)
                   "fpr_closed_form": round(2 * 16 * (checks.get("occupied") or 0) / CAPACITY / 2**16, 7)},
    }
    if load_sweep is not None:
        result["load_sweep"] = load_sweep
    if load_factor is not None:
        result["load_factor"] = load_factor
    if capacity_sweep is not None:
        result["capacity_sweep"] = capacity_sweep
    if eviction_sweep is not None:
        result["eviction_sweep"] = eviction_sweep
    if policy_sweep is not None:
        result["policy_sweep"] = policy_sweep
    if big_table is not None:
        result["table_2gib"] = big_table
    if batched is not None:
        result["batched"] = batched
    if serving is not None:
        result["serving"] = serving
    del keys, pos, query, dels, out, dout
    f.close()
    torch.cuda.empty_cache()
    if not args.no_roofline:
        import roofline as rl

        log("measuring the random-gather roofline")
        m = rl.measure()
        result["measured_roofline"] = m
        g = m.get("gather32_512MiB_gbs")
        if g:
            for op in ("insert", "lookup", "delete"):
                result["breakdown"][f"{op}_frac_of_gather32"] = round(per_op[f"{op}_algorithmic_gbs"] / g, 3)
            for r in (roofline, roofline_insert):
                r["frac_of_measured_gather32"] = round(r["achieved"] / g, 3)
        g2 = m.get("gather32_2048MiB_gbs")
        if g2 and big_table is not None:
            for op in ("insert", "lookup", "delete"):
                big_table[f"{op}_frac_of_gather32_2048MiB"] = round(big_table[f"{op}_algorithmic_gbs"] / g2, 3)
    if not args.no_cpu_baseline:
        log("timing the CPU baseline")
        result["cpu_baseline"] = cpu_baseline(n_ins, N_LOOKUP, n_del,
                                              cpu_cuckoo_log2=[int(x) for x in args.cpu_cuckoo_log2.split(",") if x])
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
