"""Serving loop (cf_server_*) against the plain calls and the oracle.

The serving contract: every request gets the result it would get if the
requests ran one by one, in arrival order, through the plain API.
  * lookups are pure: on given images they are bit-exact with the oracle
    whether a request runs alone (submit + wait) or in the mix (a replayed
    burst that the server batches into shared windows);
  * inserts / deletes in a mixed stream: each request's post-op occupancy
    equals the plain call's, deletes of present keys all succeed, every
    present key is found at its point in the order, and the false-positive
    count of negatives agrees with the plain run within a binomial band;
  * host-memory requests go through the staging arena; an arena smaller than
    the stream makes admission wait, without changing results.
"""
import numpy as np
import pytest

from conftest import assert_table_count

pytestmark = pytest.mark.gpu


def _oracle_filters(cf, O, caps, seed, load, kws=None):
    """Filters with oracle-built images; filter i takes configuration
    kws[i % len(kws)] (default: 16-bit tags, bucket 16, u64 words, BFS)."""
    kws = kws or [dict(eviction="bfs")]
    fs, ofs, keysets = [], [], []
    for i, cap in enumerate(caps):
        c = cf.make_config(**kws[i % len(kws)])
        oc = O.make_config(**c.as_dict())
        of = O.Filter(oc, cap)
        keys = O.gen_keys(seed + i, int(of.capacity() * load))
        occ, _ = of.insert_many(keys)
        f = cf.CuckooFilter(cap, config=c)
        f.importTable(of.table_bytes(), occ)
        fs.append(f)
        ofs.append(of)
        keysets.append(keys)
    return fs, ofs, keysets


def _lookup_stream(cf, O, fs, keysets, rng, count, max_log2=13, host_every=3):
    import torch

    reqs, expect_idx, queries = [], [], []
    for j in range(count):
        i = int(rng.integers(len(fs)))
        n = int(2 ** rng.uniform(0, max_log2))
        pos = keysets[i][rng.integers(0, keysets[i].size, n // 2)]
        q = np.concatenate([pos, O.gen_keys(50_000 + j, n - n // 2)])
        host = j % host_every == 0
        if host:
            keys, out = q, np.zeros(n, dtype=bool)
        else:
            keys, out = cf.keys_to_device(q), torch.zeros(n, dtype=torch.bool, device="cuda")
        reqs.append((fs[i], cf.OP_LOOKUP, keys, out, host))
        expect_idx.append(i)
        queries.append(q)
    return reqs, expect_idx, queries


def _out_np(out):
    return out if isinstance(out, np.ndarray) else out.cpu().numpy()


@pytest.mark.parametrize("ordering", ["arrival", "phased"])
def test_server_lookups_alone_and_in_mix_bit_exact(cf, oracle_mod, ordering):
    O = oracle_mod
    rng = np.random.default_rng(3)
    caps = [int(2 ** x) for x in rng.uniform(12, 18, size=10)]
    fs, ofs, keysets = _oracle_filters(cf, O, caps, seed=300, load=0.85)
    reqs, idx, queries = _lookup_stream(cf, O, fs, keysets, rng, 300)
    expect = [ofs[i].contains_many(q) for i, q in zip(idx, queries)]
    with cf.Server(ordering=ordering, form_late=ordering == "phased") as srv:
        # alone: one request in flight at a time
        for (f, op, keys, out, host), e in zip(reqs[:60], expect[:60]):
            t = srv.submit(f, op, keys, out, host)
            r = srv.wait(t)
            assert r["status"] == 0
            np.testing.assert_array_equal(_out_np(out), e)
            assert r["occupied"] == f.occupiedSlots()
        # in the mix: the whole stream as one burst
        for _, _, _, out, _ in reqs:
            if isinstance(out, np.ndarray):
                out[:] = False
            else:
                out.zero_()
        arr = srv.make_requests(reqs)
        res = srv.replay(arr, np.zeros(len(reqs)))
        assert (res["status"] == 0).all()
        st = srv.stats()
        assert st["requests"] == 60 + len(reqs)
        assert st["windows"] < 60 + len(reqs)  # the burst shares windows
        for (f, _, _, out, _), e, r in zip(reqs, expect, res):
            np.testing.assert_array_equal(_out_np(out), e)
            assert int(r["occupied"]) == f.occupiedSlots()
        assert (res["done_ns"] >= res["start_ns"]).all() and (res["start_ns"] >= res["arrival_ns"]).all()


@pytest.mark.parametrize("ordering", ["arrival", "phased"])
def test_server_filters_of_several_configurations(cf, oracle_mod, ordering):
    """A window holds requests of one configuration: a stream that
    alternates between filters of four configurations (tag widths, word
    sizes, policies) is cut into windows wherever the configuration changes,
    and every lookup still equals the oracle's, alone and in a burst."""
    O = oracle_mod
    rng = np.random.default_rng(11)
    caps = [int(2 ** x) for x in rng.uniform(12, 17, size=8)]
    kws = [dict(eviction="bfs"), dict(eviction="bfs", bits_per_tag=8, word_bytes=4),
           dict(eviction="dfs", bits_per_tag=32, bucket_size=8), dict(eviction="bfs", policy="offset")]
    fs, ofs, keysets = _oracle_filters(cf, O, caps, seed=700, load=0.85, kws=kws)
    reqs, idx, queries = _lookup_stream(cf, O, fs, keysets, rng, 200, max_log2=11)
    expect = [ofs[i].contains_many(q) for i, q in zip(idx, queries)]
    with cf.Server(ordering=ordering) as srv:
        for (f, op, keys, out, host), e in zip(reqs[:40], expect[:40]):
            r = srv.wait(srv.submit(f, op, keys, out, host))
            assert r["status"] == 0
            np.testing.assert_array_equal(_out_np(out), e)
        for _, _, _, out, _ in reqs:
            if isinstance(out, np.ndarray):
                out[:] = False
            else:
                out.zero_()
        res = srv.replay(srv.make_requests(reqs), np.zeros(len(reqs)))
        assert (res["status"] == 0).all()
        for (f, _, _, out, _), e, r in zip(reqs, expect, res):
            np.testing.assert_array_equal(_out_np(out), e)
            assert int(r["occupied"]) == f.occupiedSlots()
        changes = sum(reqs[j][0].config.as_dict() != reqs[j - 1][0].config.as_dict() for j in range(1, len(reqs)))
        assert srv.stats()["windows"] >= 40 + changes + 1


@pytest.mark.parametrize("max_levels", [None, 1, 2])
def test_server_keeps_arrival_order_per_filter(cf, oracle_mod, max_levels):
    """Also with windows capped at one and two levels: the cap splits the
    stream into more windows, never reorders it."""
    import torch

    f = cf.CuckooFilter(1 << 16, key_bytes=8, bits_per_tag=16, eviction="bfs")
    g = cf.CuckooFilter(1 << 16, key_bytes=8, bits_per_tag=16, eviction="bfs")
    k = cf.keys_to_device(oracle_mod.gen_keys(78, 30000))
    outs = [torch.zeros(k.numel(), dtype=torch.bool, device="cuda") for _ in range(4)]
    reqs = [(f, cf.OP_LOOKUP, k, outs[0], False), (f, cf.OP_INSERT, k, None, False),
            (g, cf.OP_INSERT, k, None, False), (f, cf.OP_LOOKUP, k, outs[1], False),
            (f, cf.OP_LOOKUP, k, outs[2], False), (f, cf.OP_DELETE, k, None, False),
            (f, cf.OP_LOOKUP, k, outs[3], False)]
    with cf.Server(max_window_levels=max_levels) as srv:
        res = srv.replay(srv.make_requests(reqs), np.zeros(len(reqs)))
        windows = srv.stats()["windows"]
    assert (res["status"] == 0).all()
    if max_levels is not None:
        assert int(res["window_levels"].max()) <= max_levels
        assert windows >= {1: 5, 2: 3}[max_levels]
    assert not outs[0].any().item()
    assert outs[1].all().item() and outs[2].all().item()
    assert not outs[3].any().item()
    assert list(res["occupied"]) == [0, 30000, 30000, 30000, 30000, 0, 0]
    assert f.occupiedSlots() == 0 and g.occupiedSlots() == 30000


@pytest.mark.parametrize("staging,max_levels,late", [(None, None, False), (1 << 20, None, False), (None, 3, False),
                                                     (None, None, True)])
def test_server_mixed_stream_matches_plain(cf, oracle_mod, staging, max_levels, late):
    import torch

    O = oracle_mod
    rng = np.random.default_rng(11 if staging else 12)
    caps = [int(2 ** x) for x in rng.uniform(13, 19, size=16)]
    plain, _, keysets = _oracle_filters(cf, O, caps, seed=700, load=0.5)
    served = []
    for a in plain:
        b = cf.CuckooFilter(a.capacity(), config=a.config)
        b.copyFrom(a)
        served.append(b)
    present = [set(map(int, ks)) for ks in keysets]
    count = [a.occupiedSlots() for a in plain]
    plan = []
    for j in range(300):
        i = int(rng.integers(len(caps)))
        kind = int(rng.choice([0, 1, 2], p=[0.3, 0.55, 0.15]))
        n = int(2 ** rng.uniform(0, 11))
        if kind == 0:
            n = max(0, min(n, int(0.9 * plain[i].capacity()) - count[i]))
            keys = O.gen_keys(120_000 + j, n)
            present[i].update(map(int, keys))
            count[i] += n
            want = None
        elif kind == 1:
            have = np.fromiter(present[i], dtype=np.uint64)
            pos = have[rng.integers(0, have.size, n // 2)] if have.size else np.zeros(0, np.uint64)
            keys = np.concatenate([pos, O.gen_keys(150_000 + j, n - pos.size)])
            want = np.array([int(x) in present[i] for x in keys], dtype=bool)
        else:
            have = np.fromiter(present[i], dtype=np.uint64)
            keys = rng.choice(have, size=min(n, have.size), replace=False) if have.size else np.zeros(0, np.uint64)
            for x in keys:
                present[i].discard(int(x))
            count[i] -= keys.size
            want = None
        plan.append((i, kind, keys, want, j % 2 == 0))
    # plain, one by one
    occ_plain, out_plain = [], []
    for i, kind, keys, _, _ in plan:
        dk = cf.keys_to_device(keys) if keys.size else torch.zeros(0, dtype=torch.int64, device="cuda")
        if kind == 0:
            occ_plain.append(plain[i].insertMany(dk))
            out_plain.append(None)
        elif kind == 1:
            o = plain[i].containsMany(dk) if keys.size else torch.zeros(0, dtype=torch.bool, device="cuda")
            out_plain.append(o.cpu().numpy())
            occ_plain.append(plain[i].occupiedSlots())
        else:
            o = torch.zeros(max(keys.size, 1), dtype=torch.bool, device="cuda")
            occ_plain.append(plain[i].deleteMany(dk, out=o[: keys.size]))
            out_plain.append(o[: keys.size].cpu().numpy())
    # served: Poisson arrivals over ~30 ms, half the requests in host memory
    reqs = []
    for i, kind, keys, _, host in plan:
        if host:
            k = keys.copy()
            out = np.zeros(keys.size, dtype=bool) if kind else None
        else:
            k = cf.keys_to_device(keys) if keys.size else torch.zeros(0, dtype=torch.int64, device="cuda")
            out = torch.zeros(max(keys.size, 1), dtype=torch.bool, device="cuda") if kind else None
        reqs.append((served[i], kind, k, out, host))
    arrivals = np.cumsum(rng.exponential(1e-4, size=len(reqs)))
    with cf.Server(staging_bytes=staging, max_window_levels=max_levels, form_late=late) as srv:
        res = srv.replay(srv.make_requests(reqs), arrivals)
        st = srv.stats()
    assert (res["status"] == 0).all()
    if max_levels is not None:
        assert int(res["window_levels"].max()) <= max_levels
    np.testing.assert_array_equal(res["occupied"], np.array(occ_plain, dtype=np.uint64))
    for (i, kind, keys, want, _), (_, _, _, out, _), op in zip(plan, reqs, out_plain):
        if not keys.size or kind == 0:
            continue
        got = _out_np(out)[: keys.size]
        if kind == 2:
            assert got.all() and op.all()
        else:
            assert got[want].all()
            fp_s, fp_p = int(got[~want].sum()), int(op[~want].sum())
            assert abs(fp_s - fp_p) <= 3 * np.sqrt(max(fp_s + fp_p, 1)) + 2
    for a, b in zip(plain, served):
        assert a.occupiedSlots() == b.occupiedSlots()
        assert_table_count(b, b.occupiedSlots(), b.capacity())
    if staging:
        assert st["staging_bytes"] <= staging


def _phase_model(plan, windows, start_counts, start_present):
    """Expected results of a phased run, given the window each request ran in:
    per window and filter, the inserts (all of them), then the lookups and
    deletes together.  plan[j] = (filter index, kind, keys); returns per
    request the expected occupancy (its filter's after its phase) and, for
    lookups, the keys that must be found and the keys a delete of the same
    window removes (either answer is allowed for those)."""
    count = list(start_counts)
    present = [set(p) for p in start_present]
    occ = [None] * len(plan)
    must = [None] * len(plan)
    maybe = [None] * len(plan)
    for w in np.unique(windows):
        idx = [j for j in range(len(plan)) if windows[j] == w]
        for j in idx:  # phase 1: inserts (cf.OP_INSERT == 0)
            i, k, keys = plan[j]
            if k == 0:
                present[i].update(map(int, keys))
                count[i] += keys.size
        for j in idx:
            if plan[j][1] == 0:
                occ[j] = count[plan[j][0]]
        gone = {}  # phase 2: keys deleted in this window, per filter (cf.OP_DELETE == 2)
        for j in idx:
            i, k, keys = plan[j]
            if k == 2:
                gone.setdefault(i, set()).update(map(int, keys))
        for j in idx:
            i, k, keys = plan[j]
            if k == 1:
                g = gone.get(i, set())
                maybe[j] = np.array([int(x) in g for x in keys], dtype=bool)
                must[j] = np.array([int(x) in present[i] for x in keys], dtype=bool) & ~maybe[j]
        for j in idx:
            i, k, keys = plan[j]
            if k == 2:
                present[i].difference_update(map(int, keys))
                count[i] -= keys.size
        for j in idx:
            if plan[j][1] != 0:
                occ[j] = count[plan[j][0]]
    return occ, must, maybe, count


def test_server_phased_order_in_one_filter(cf, oracle_mod):
    """The phased ordering on a short stream: whatever windows the server forms,
    each window runs a filter's inserts, then its lookups and deletes."""
    import torch

    f = cf.CuckooFilter(1 << 16, key_bytes=8, bits_per_tag=16, eviction="bfs")
    g = cf.CuckooFilter(1 << 16, key_bytes=8, bits_per_tag=16, eviction="bfs")
    kh = oracle_mod.gen_keys(78, 30000)
    k = cf.keys_to_device(kh)
    outs = [torch.zeros(k.numel(), dtype=torch.bool, device="cuda") for _ in range(4)]
    reqs = [(f, cf.OP_LOOKUP, k, outs[0], False), (f, cf.OP_INSERT, k, None, False),
            (g, cf.OP_INSERT, k, None, False), (f, cf.OP_LOOKUP, k, outs[1], False),
            (f, cf.OP_LOOKUP, k, outs[2], False), (f, cf.OP_DELETE, k, None, False),
            (f, cf.OP_LOOKUP, k, outs[3], False)]
    plan = [(0 if r[0] is f else 1, r[1], kh) for r in reqs]
    with cf.Server(ordering="phased") as srv:
        res = srv.replay(srv.make_requests(reqs), np.zeros(len(reqs)))
    assert (res["status"] == 0).all()
    assert int(res["window_levels"].max()) <= 2
    occ, must, maybe, count = _phase_model(plan, res["window"], [0, 0], [set(), set()])
    assert list(res["occupied"]) == occ
    look = [j for j, r in enumerate(reqs) if r[1] == cf.OP_LOOKUP]
    for o, j in zip(outs, look):
        got = o.cpu().numpy()
        assert got[must[j]].all()
        assert got[~must[j] & ~maybe[j]].sum() <= 60  # absent keys: false positives only
    assert f.occupiedSlots() == count[0] == 0 and g.occupiedSlots() == count[1] == 30000


@pytest.mark.parametrize("max_levels,late", [(None, False), (1, False), (None, True)])
def test_server_phased_stream_matches_model(cf, oracle_mod, max_levels, late):
    """A mixed stream of 300 requests on 16 filters under the phased ordering,
    Poisson arrivals: occupancies equal the phase model's for the windows the
    server formed, every key the model holds present is found, false positives
    stay within the closed-form bound, deletes of present keys succeed, and the
    final tables hold exactly the model's counts."""
    import torch

    O = oracle_mod
    rng = np.random.default_rng(21)
    caps = [int(2 ** x) for x in rng.uniform(13, 19, size=16)]
    served, _, keysets = _oracle_filters(cf, O, caps, seed=900, load=0.5)
    present0 = [set(map(int, ks)) for ks in keysets]
    count0 = [f.occupiedSlots() for f in served]
    present = [set(p) for p in present0]
    count = list(count0)
    plan = []
    for j in range(300):
        i = int(rng.integers(len(caps)))
        kind = int(rng.choice([0, 1, 2], p=[0.3, 0.55, 0.15]))
        n = int(2 ** rng.uniform(0, 11))
        if kind == 0:
            n = max(0, min(n, int(0.9 * served[i].capacity()) - count[i]))
            keys = O.gen_keys(320_000 + j, n)
            present[i].update(map(int, keys))
            count[i] += n
        elif kind == 1:
            have = np.fromiter(present[i], dtype=np.uint64)
            pos = have[rng.integers(0, have.size, n // 2)] if have.size else np.zeros(0, np.uint64)
            keys = np.concatenate([pos, O.gen_keys(350_000 + j, n - pos.size)])
        else:
            have = np.fromiter(present[i], dtype=np.uint64)
            keys = rng.choice(have, size=min(n, have.size), replace=False) if have.size else np.zeros(0, np.uint64)
            for x in keys:
                present[i].discard(int(x))
            count[i] -= keys.size
        plan.append((i, kind, keys))
    reqs = []
    for j, (i, kind, keys) in enumerate(plan):
        host = j % 2 == 0
        if host:
            k = keys.copy()
            out = np.zeros(keys.size, dtype=bool) if kind else None
        else:
            k = cf.keys_to_device(keys) if keys.size else torch.zeros(0, dtype=torch.int64, device="cuda")
            out = torch.zeros(max(keys.size, 1), dtype=torch.bool, device="cuda") if kind else None
        reqs.append((served[i], kind, k, out, host))
    arrivals = np.cumsum(rng.exponential(1e-4, size=len(reqs)))
    with cf.Server(ordering="phased", max_window_levels=max_levels, form_late=late) as srv:
        res = srv.replay(srv.make_requests(reqs), arrivals)
        st = srv.stats()
    assert (res["status"] == 0).all()
    assert int(res["window_levels"].max()) <= (max_levels or 2)
    assert st["windows"] < len(reqs)  # requests shared windows
    occ, must, maybe, final = _phase_model(plan, res["window"], count0, present0)
    np.testing.assert_array_equal(res["occupied"], np.array(occ, dtype=np.uint64))
    fp, neg = 0, 0
    for j, ((i, kind, keys), (_, _, _, out, _)) in enumerate(zip(plan, reqs)):
        if not keys.size or kind == 0:
            continue
        got = _out_np(out)[: keys.size]
        if kind == 2:
            assert got.all()
        else:
            assert got[must[j]].all()
            absent = ~must[j] & ~maybe[j]
            fp += int(got[absent].sum())
            neg += int(absent.sum())
    # 2 * bucketSize * alpha / 2^bitsPerTag at alpha <= 0.9 (src/main.cu:114-115)
    p = 2 * 16 * 0.9 / 2 ** 16
    assert fp <= neg * p + 4 * np.sqrt(neg * p) + 3
    for f, n in zip(served, final):
        assert f.occupiedSlots() == n
        assert_table_count(f, n, f.capacity())


def test_server_ticket_api_and_errors(cf, oracle_mod):
    import torch

    f = cf.CuckooFilter(1 << 14, key_bytes=8, bits_per_tag=16)
    keys = oracle_mod.gen_keys(5, 1000)
    with cf.Server(staging_bytes=64 << 10) as srv:
        t1 = srv.submit(f, cf.OP_INSERT, keys, None, host=True)
        out = np.zeros(keys.size, dtype=bool)
        t2 = srv.submit(f, cf.OP_LOOKUP, keys, out, host=True)
        srv.drain()
        r2 = srv.poll(t2)
        assert r2 is not None and r2["status"] == 0 and out.all()
        assert srv.wait(t1)["occupied"] == 1000
        with pytest.raises(cf.CuckooError):
            srv.wait(t1)  # already collected
        with pytest.raises(cf.CuckooError):
            srv.poll(10 ** 9)  # never issued
        big = oracle_mod.gen_keys(6, 10000)  # 80 KB of keys > a 32 KiB window arena
        with pytest.raises(cf.CuckooError, match="staging"):
            srv.submit(f, cf.OP_LOOKUP, big, np.zeros(big.size, dtype=bool), host=True)
        with pytest.raises(cf.CuckooError):
            srv.submit(f, cf.OP_LOOKUP, keys, None, host=True)  # lookup without output
        # device memory is not staged: the same request passes on the device
        dk = cf.keys_to_device(big)
        dout = torch.zeros(big.size, dtype=torch.bool, device="cuda")
        assert srv.wait(srv.submit(f, cf.OP_LOOKUP, dk, dout))["status"] == 0


def test_server_destroy_finishes_pending_requests(cf, oracle_mod):
    import torch

    O = oracle_mod
    f = cf.CuckooFilter(1 << 18, key_bytes=8, bits_per_tag=16, eviction="bfs")
    keys = O.gen_keys(900, 200_000)
    dk = cf.keys_to_device(keys)
    outs = []
    srv = cf.Server(max_window_requests=4)
    for j in range(0, keys.size, 2000):
        srv.submit(f, cf.OP_INSERT, dk[j:j + 2000])
    for j in range(0, keys.size, 5000):
        o = torch.zeros(5000, dtype=torch.bool, device="cuda")
        srv.submit(f, cf.OP_LOOKUP, dk[j:j + 5000], o)
        outs.append(o)
    srv.close()  # returns once every queued request has run
    assert f.occupiedSlots() == keys.size
    assert all(bool(o.all().item()) for o in outs)


@pytest.mark.parametrize("slots", [2, 3])
def test_server_large_host_memory_windows(cf, oracle_mod, slots):
    """Requests of more than 8 MB of host-memory keys each, so that every
    window holding one is staged by several threads and uploaded on the copy
    stream: inserts then lookups through host memory give the plain calls'
    occupancies and the oracle's lookup bits."""
    O = oracle_mod
    c = cf.make_config(eviction="bfs")
    oc = O.make_config(**c.as_dict())
    f = cf.CuckooFilter(1 << 24, config=c)
    of = O.Filter(oc, 1 << 24)
    n_each, parts = 3 << 19, 6  # 12 MiB of keys per request
    keys = O.gen_keys(880, n_each * parts)
    reqs = [(f, cf.OP_INSERT, keys[p * n_each:(p + 1) * n_each].copy(), None, True) for p in range(parts)]
    with cf.Server(window_slots=slots) as srv:
        res = srv.replay(srv.make_requests(reqs), np.zeros(parts))
        assert (res["status"] == 0).all()
        np.testing.assert_array_equal(res["occupied"], np.arange(1, parts + 1, dtype=np.uint64) * n_each)
        # the oracle holds the served image; lookups through host memory must match it bit for bit
        of.load_table(f.exportTable(), f.occupiedSlots())
        q = np.concatenate([keys[::3], O.gen_keys(881, keys.size // 3)])  # 8 MiB per lookup request
        chunks = np.array_split(q, parts)
        outs = [np.zeros(ch.size, dtype=bool) for ch in chunks]
        lreqs = [(f, cf.OP_LOOKUP, ch, o, True) for ch, o in zip(chunks, outs)]
        res = srv.replay(srv.make_requests(lreqs), np.zeros(parts))
        assert (res["status"] == 0).all()
    got = np.concatenate(outs)
    np.testing.assert_array_equal(got, of.contains_many(q))
    assert got[: keys[::3].size].all()
    plain = f.containsMany(cf.keys_to_device(q)).cpu().numpy()
    np.testing.assert_array_equal(got, plain)
