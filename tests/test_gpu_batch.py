"""Batched many-filter execution (cf_batch_execute) against plain per-batch calls.

A window must give the results of running its batches one by one through the
plain API (cf_insert_many / cf_contains_many / cf_delete_many):
  * lookups on given images are bit-exact with the plain calls and with the
    oracle, also when the captured graph is replayed with eager work between
    replays (each replay checked);
  * per-filter order holds inside a window (lookup -> insert -> lookup ->
    delete -> lookup on one filter);
  * mixed windows over many filters return the same post-op occupancies,
    the same delete results, and find every key that is present.
"""
import numpy as np
import pytest

from conftest import assert_table_count

pytestmark = pytest.mark.gpu


def _dev(cf, keys):
    return cf.keys_to_device(np.ascontiguousarray(keys))


# configurations the batched executor is run with: tag widths, word sizes,
# the three alternate-bucket policies, both eviction policies
BATCH_CFGS = [dict(eviction="bfs"), dict(eviction="bfs", bits_per_tag=8, word_bytes=4),
              dict(eviction="dfs", bits_per_tag=32, bucket_size=8), dict(eviction="bfs", policy="addsub"),
              dict(eviction="dfs", policy="offset")]
CFG_IDS = ["-".join(f"{k}{v}" for k, v in d.items()) for d in BATCH_CFGS]


def _make_filters(cf, O, caps, seed, load=0.8, kw=None):
    """Filters with oracle-built images, and identical twins."""
    c = cf.make_config(**(kw or dict(eviction="bfs")))
    oc = O.make_config(**c.as_dict())
    plain, batched, keysets = [], [], []
    for i, cap in enumerate(caps):
        of = O.Filter(oc, cap)
        keys = O.gen_keys(seed + i, int(of.capacity() * load))
        occ, _ = of.insert_many(keys)
        a = cf.CuckooFilter(cap, config=c)
        a.importTable(of.table_bytes(), occ)
        b = cf.CuckooFilter(cap, config=c)
        b.copyFrom(a)
        plain.append(a)
        batched.append(b)
        keysets.append(keys)
    return plain, batched, keysets


def _run_plain(ops):
    occ = []
    for f, op, keys, out in ops:
        if op == 0:
            occ.append(f.insertMany(keys))
        elif op == 1:
            f.containsMany(keys, out=out)
            occ.append(f.occupiedSlots())
        else:
            occ.append(f.deleteMany(keys, out=out))
    return np.array(occ, dtype=np.uint64)


@pytest.mark.parametrize("kw", BATCH_CFGS, ids=CFG_IDS)
def test_batch_lookups_bit_exact(cf, oracle_mod, kw):
    import torch

    O = oracle_mod
    rng = np.random.default_rng(1)
    caps = [int(2 ** x) for x in rng.uniform(12, 18, size=12)]
    plain, batched, keysets = _make_filters(cf, O, caps, seed=100, kw=kw)
    ops_p, ops_b, expect = [], [], []
    for j in range(200):
        i = int(rng.integers(len(caps)))
        n = int(2 ** rng.uniform(0, 14))
        pos = keysets[i][rng.integers(0, keysets[i].size, n // 2)]
        neg = O.gen_keys(10_000 + j, n - n // 2)
        q = np.concatenate([pos, neg])
        dq = _dev(cf, q)
        ops_p.append((plain[i], 1, dq, torch.zeros(n, dtype=torch.bool, device="cuda")))
        ops_b.append((batched[i], 1, dq, torch.zeros(n, dtype=torch.bool, device="cuda")))
        of = O.Filter(O.make_config(**plain[i].config.as_dict()), caps[i])
        of.load_table(plain[i].exportTable(), plain[i].occupiedSlots())
        expect.append(of.contains_many(q))
    _run_plain(ops_p)
    b = cf.Batch()
    arr = b.make_ops(ops_b)
    for replay in range(4):
        for _, _, _, out in ops_b:
            out.zero_()
        b.execute(arr)
        for (_, _, _, op), (_, _, _, ob), e in zip(ops_p, ops_b, expect):
            got = ob.cpu().numpy()
            np.testing.assert_array_equal(got, op.cpu().numpy(), err_msg=f"replay {replay}")
            np.testing.assert_array_equal(got, e)
        # eager work and a synchronisation between replays
        plain[0].containsMany(ops_p[0][2])
        torch.cuda.synchronize()
    b.close()


def test_batch_keeps_per_filter_order(cf, oracle_mod):
    import torch

    f = cf.CuckooFilter(1 << 16, key_bytes=8, bits_per_tag=16, eviction="bfs")
    g = cf.CuckooFilter(1 << 16, key_bytes=8, bits_per_tag=16, eviction="bfs")
    k = _dev(cf, oracle_mod.gen_keys(77, 40000))
    outs = [torch.zeros(k.numel(), dtype=torch.bool, device="cuda") for _ in range(3)]
    ops = [(f, 1, k, outs[0]), (f, 0, k, None), (f, 1, k, outs[1]), (f, 2, k, None), (f, 1, k, outs[2]),
           (g, 0, k, None)]
    b = cf.Batch()
    occ = b.execute(ops)
    assert not outs[0].any().item()            # empty filter
    assert outs[1].all().item()                # after the insert
    assert not outs[2].any().item()            # after deleting everything again
    assert list(occ) == [0, 40000, 40000, 0, 0, 40000]
    assert f.occupiedSlots() == 0 and g.occupiedSlots() == 40000
    b.close()


@pytest.mark.parametrize("chunk_sizes", [[3000] * 11, [1, 700, 5000, 2, 0, 1300, 1, 9000, 4, 60, 513, 511]])
def test_batch_consecutive_inserts_report_sequential_occupancies(cf, oracle_mod, chunk_sizes):
    """Runs of inserts on one filter, some longer than eight, broken by a
    lookup or a delete, with a second filter and empty inserts between: each
    insert reports the occupancy after itself.  Without insert failures
    (under 50% load) every reported occupancy equals the plain calls'
    exactly, and the lookup sees every insert before it."""
    import torch

    O = oracle_mod
    f = cf.CuckooFilter(1 << 18, key_bytes=8, bits_per_tag=16, eviction="bfs")
    g = cf.CuckooFilter(1 << 18, key_bytes=8, bits_per_tag=16, eviction="bfs")
    pf = cf.CuckooFilter(1 << 18, key_bytes=8, bits_per_tag=16, eviction="bfs")
    pg = cf.CuckooFilter(1 << 18, key_bytes=8, bits_per_tag=16, eviction="bfs")
    keys = O.gen_keys(901, sum(chunk_sizes) + 1)
    parts, at = [], 0
    for n in chunk_sizes:
        parts.append(keys[at:at + n])
        at += n
    plan = [(0, 0, p) for p in parts[:9]]                       # nine inserts in a row
    plan += [(0, 1, parts[0]), (1, 0, parts[9]), (0, 0, parts[9])]  # a lookup breaks the run; g in between
    plan += [(0, 0, parts[10]), (0, 2, parts[1]), (0, 0, parts[-1][:1])]  # insert, delete, insert
    def run(fs):
        ops, outs = [], []
        for i, kind, k in plan:
            dk = _dev(cf, k) if k.size else torch.zeros(0, dtype=torch.int64, device="cuda")
            out = torch.zeros(max(k.size, 1), dtype=torch.bool, device="cuda") if kind else None
            ops.append((fs[i], kind, dk, out))
            outs.append(out)
        return ops, outs
    ops_b, outs_b = run([f, g])
    ops_p, outs_p = run([pf, pg])
    occ_p = _run_plain(ops_p)
    b = cf.Batch()
    occ_b = b.execute(ops_b)
    b.close()
    np.testing.assert_array_equal(occ_b, occ_p)
    assert outs_b[9][: parts[0].size].all().item()  # the lookup sees every insert before it
    assert f.occupiedSlots() == pf.occupiedSlots() and g.occupiedSlots() == pg.occupiedSlots()
    assert_table_count(f, f.occupiedSlots(), f.capacity())


@pytest.mark.parametrize("kw", BATCH_CFGS, ids=CFG_IDS)
def test_batch_mixed_window_matches_plain(cf, oracle_mod, kw):
    import torch

    O = oracle_mod
    rng = np.random.default_rng(7)
    caps = [int(2 ** x) for x in rng.uniform(13, 20, size=24)]
    plain, batched, keysets = _make_filters(cf, O, caps, seed=500, load=0.5, kw=kw)
    present = [set(map(int, ks)) for ks in keysets]
    count = [plain[i].occupiedSlots() for i in range(len(caps))]
    ops_p, ops_b, checks = [], [], []
    fresh = 0
    for j in range(400):
        i = int(rng.integers(len(caps)))
        kind = rng.choice([0, 1, 2], p=[0.4, 0.5, 0.1])
        n = int(2 ** rng.uniform(0, 12))
        if kind == 0:
            room = int(0.9 * plain[i].capacity()) - count[i]
            n = max(0, min(n, room))
            keys = O.gen_keys(90_000 + fresh, n)
            fresh += 1
            present[i].update(map(int, keys))
            count[i] += n
        elif kind == 1:
            have = np.fromiter(present[i], dtype=np.uint64) if present[i] else np.zeros(0, np.uint64)
            pos = have[rng.integers(0, have.size, n // 2)] if have.size else np.zeros(0, np.uint64)
            keys = np.concatenate([pos, O.gen_keys(70_000 + j, n - pos.size)])
        else:
            have = np.fromiter(present[i], dtype=np.uint64)
            keys = rng.choice(have, size=min(n, have.size), replace=False) if have.size else np.zeros(0, np.uint64)
            for x in keys:
                present[i].discard(int(x))
            count[i] -= keys.size
        dk = _dev(cf, keys) if keys.size else torch.zeros(0, dtype=torch.int64, device="cuda")
        out_p = torch.zeros(max(keys.size, 1), dtype=torch.bool, device="cuda") if kind else None
        out_b = torch.zeros(max(keys.size, 1), dtype=torch.bool, device="cuda") if kind else None
        ops_p.append((plain[i], int(kind), dk, out_p))
        ops_b.append((batched[i], int(kind), dk, out_b))
        want = np.array([int(x) in present[i] for x in keys], dtype=bool) if kind == 1 else None
        checks.append((int(kind), keys, want))
    # a third twin of every filter runs the window through cf_run_plain_calls (the C loop)
    looped = []
    for b in batched:
        t = cf.CuckooFilter(b.capacity(), config=b.config)
        t.copyFrom(b)
        looped.append(t)
    by_filter = {id(b): t for b, t in zip(batched, looped)}
    ops_c = [(by_filter[id(f)], kind, dk, torch.zeros_like(ob) if ob is not None else None)
             for f, kind, dk, ob in ops_b]
    occ_p = _run_plain(ops_p)
    occ_b = cf.Batch().execute(ops_b)
    occ_c = cf.run_plain_calls(ops_c)
    np.testing.assert_array_equal(occ_b, occ_p)
    np.testing.assert_array_equal(occ_c, occ_p)
    for (_, kind, dk, oc_), (_, _, _, op) in zip(ops_c, ops_p):
        if kind == 2 and dk.numel():
            assert oc_[: dk.numel()].all().item()
    for (f, kind, dk, op), (_, _, _, ob), (_, keys, want_pos) in zip(ops_p, ops_b, checks):
        if kind == 2 and keys.size:
            assert ob[: keys.size].all().item() and op[: keys.size].all().item()
        if kind == 1 and keys.size:
            got = ob[: keys.size].cpu().numpy()
            assert got[want_pos].all()  # every present key is found
            # negatives: false-positive counts agree within a binomial band
            fp_b, fp_p = int(got[~want_pos].sum()), int(op[: keys.size].cpu().numpy()[~want_pos].sum())
            assert abs(fp_b - fp_p) <= 3 * np.sqrt(max(fp_b + fp_p, 1)) + 2
    for a, b in zip(plain, batched):
        assert a.occupiedSlots() == b.occupiedSlots()
        assert_table_count(b, b.occupiedSlots(), b.capacity())
