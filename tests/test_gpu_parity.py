"""Parity of the HIP kernels (through the C ABI) with the oracle.

Bit-exact where the reference is deterministic:
  * device getCandidateBuckets / computePackedTags vs the golden vectors;
  * containsMany on a given table image (oracle-built and hand-built) vs the
    oracle's lookups;
  * one-key-per-call insertMany / insertManySorted / deleteMany -- each call
    is a whole kernel launch, so the calls are strictly ordered and the table
    image after every call must equal the sequential oracle's, eviction paths
    and exhausted chains included.
Statistical / invariant where the reference itself races (concurrent batch
inserts, CuckooFilter.cuh:1119-1141): occupancy == non-empty slots, stored
tags == the keys' fingerprints, every stored tag sits in one of its key's
buckets, no false negatives, failure count and FPR within binomial bounds of
the oracle at the same load.
"""
import numpy as np
import pytest

from conftest import golden_cases, assert_table_count

pytestmark = pytest.mark.gpu

TAG_DT = {8: np.uint8, 16: np.uint16, 32: np.uint32}


def _dev_keys(cf, keys):
    return cf.keys_to_device(np.ascontiguousarray(keys))


def _oracle_cfg(O, c):
    d = c.as_dict()
    return O.make_config(**d)


def _tags(image, c):
    return np.frombuffer(image.tobytes(), dtype=TAG_DT[c.bits_per_tag]).reshape(-1, c.bucket_size)


# ---------------------------------------------------------------------------
# hashing
# ---------------------------------------------------------------------------

def test_device_candidates_and_packed_tags_match_golden(cf, golden):
    for name in golden_cases(golden):
        pol, kb, bits, nb = (int(x) for x in golden[name + "/meta"])
        c = cf.make_config(key_bytes=kb, bits_per_tag=bits, policy=pol, bucket_size=8, word_bytes=8)
        f = cf.CuckooFilter(nb * 8, config=c)
        assert f.getNumBuckets() == nb, name
        k = _dev_keys(cf, golden[name + "/keys"])
        i1, i2, fp = f.computeCandidates(k)
        np.testing.assert_array_equal(cf.to_numpy_u(i1, np.uint64), golden[name + "/i1"], err_msg=name)
        np.testing.assert_array_equal(cf.to_numpy_u(i2, np.uint64), golden[name + "/i2"], err_msg=name)
        np.testing.assert_array_equal(cf.to_numpy_u(fp, np.uint32), golden[name + "/fp"], err_msg=name)
        packed = f.computePackedTags(k)
        pdt = np.uint32 if bits <= 8 else np.uint64
        np.testing.assert_array_equal(cf.to_numpy_u(packed, pdt).astype(np.uint64), golden[name + "/packed"],
                                      err_msg=name)
        f.close()


# ---------------------------------------------------------------------------
# lookup on a given image: bit-exact
# ---------------------------------------------------------------------------

LOOKUP_CFGS = [
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8),
    dict(bits_per_tag=16, bucket_size=16, word_bytes=4),
    dict(bits_per_tag=8, bucket_size=32, word_bytes=8),
    dict(bits_per_tag=8, bucket_size=4, word_bytes=4),
    dict(bits_per_tag=32, bucket_size=32, word_bytes=8),   # 128-byte buckets
    dict(bits_per_tag=32, bucket_size=4, word_bytes=4),
    dict(bits_per_tag=16, bucket_size=4, word_bytes=4),    # the reference's undefined 128-bit case
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, policy="addsub"),
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, policy="offset"),
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, key_bytes=4),
]


@pytest.mark.parametrize("kw", LOOKUP_CFGS, ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_contains_bit_exact_on_oracle_image(cf, oracle_mod, kw):
    O = oracle_mod
    c = cf.make_config(eviction="bfs", **kw)
    oc = _oracle_cfg(O, c)
    cap = 1 << 16
    of = O.Filter(oc, cap)
    f = cf.CuckooFilter(cap, config=c)
    assert f.getNumBuckets() == of.num_buckets and f.sizeInBytes() == of.size_bytes
    n = int(of.capacity() * 0.9)
    keys = O.gen_keys(11, n + 60000, c.key_bytes)
    occ, _ = of.insert_many(keys[:n])
    f.importTable(of.table_bytes(), occ)
    assert f.occupiedSlots() == occ
    q = np.concatenate([keys[: n // 2], keys[n:]])
    got = f.containsMany(_dev_keys(cf, q)).cpu().numpy()
    np.testing.assert_array_equal(got, of.contains_many(q))
    np.testing.assert_array_equal(f.exportTable(), of.table_bytes())


def test_contains_bit_exact_on_hand_built_images(cf, oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(3)
    for kw in LOOKUP_CFGS:
        c = cf.make_config(**kw)
        oc = _oracle_cfg(O, c)
        of = O.Filter(oc, 512 * c.bucket_size)
        nb = of.num_buckets
        keys = O.gen_keys(21, 20000, c.key_bytes)
        i1, i2, fp = O.candidates(oc, nb, keys)
        tmax = (1 << c.bits_per_tag) - 1
        tags = rng.integers(0, tmax + 1, size=(nb, c.bucket_size)).astype(np.uint64)
        tags[rng.random(tags.shape) < 0.3] = 0
        tags[rng.random(tags.shape) < 0.05] = 1
        tags[::5] = tags[::5, :1]  # duplicated tags / full buckets
        for j in rng.choice(keys.size, 2000, replace=False):
            b, t = (i1[j], fp[j]) if rng.random() < 0.5 else (i2[j], O.alt_bucket(oc, nb, int(i1[j]), int(fp[j]))[1])
            tags[b, rng.integers(0, c.bucket_size)] = t
        image = tags.astype(TAG_DT[c.bits_per_tag]).view(np.uint8).reshape(-1)
        of.load_table(image, int((tags != 0).sum()))
        f = cf.CuckooFilter(512 * c.bucket_size, config=c)
        f.importTable(image, int((tags != 0).sum()))
        got = f.containsMany(_dev_keys(cf, keys)).cpu().numpy()
        want = of.contains_many(keys)
        np.testing.assert_array_equal(got, want, err_msg=str(kw))
        assert want.any() and (~want).any()
        f.close()


# ---------------------------------------------------------------------------
# strictly ordered single-key calls: bit-exact table images
# ---------------------------------------------------------------------------

SEQ_CFGS = [
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="bfs", max_evictions=40),
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="dfs", max_evictions=40),
    dict(bits_per_tag=16, bucket_size=4, word_bytes=8, eviction="bfs", max_evictions=25),
    dict(bits_per_tag=8, bucket_size=16, word_bytes=4, eviction="dfs", max_evictions=30),
    dict(bits_per_tag=32, bucket_size=8, word_bytes=8, eviction="bfs", max_evictions=30),
    dict(bits_per_tag=32, bucket_size=32, word_bytes=4, eviction="bfs", max_evictions=30),
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="bfs", policy="addsub", max_evictions=30),
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="dfs", policy="offset", max_evictions=30),
    dict(bits_per_tag=8, bucket_size=8, word_bytes=8, eviction="bfs", policy="offset", max_evictions=30),
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="dfs", key_bytes=4, max_evictions=30),
]


@pytest.mark.parametrize("path", ["direct", "binned"])
@pytest.mark.parametrize("kw", SEQ_CFGS, ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_single_key_calls_bit_exact(cf, oracle_mod, kw, path, monkeypatch):
    """insertMany's and deleteMany's two device paths (one thread per key with
    device CAS; the bucket-binned LDS path) both reproduce the sequential
    oracle call by call."""
    monkeypatch.setenv("CUCKOO_INSERT_PATH", path)
    monkeypatch.setenv("CUCKOO_DELETE_PATH", path)
    O = oracle_mod
    c = cf.make_config(**kw)
    oc = _oracle_cfg(O, c)
    cap = 64 * c.bucket_size
    of = O.Filter(oc, cap)
    f = cf.CuckooFilter(cap, config=c)
    f.setEvictionHistogram(True)
    n = int(of.capacity() * 1.06)
    keys = O.gen_keys(31, n, c.key_bytes)
    dk = _dev_keys(cf, keys)
    fails = 0
    want_hist = np.zeros(c.max_evictions + 2, np.uint64)
    for i in range(n):
        occ_gpu = f.insertMany(dk[i:i + 1])
        occ_cpu, ok, chain = of.insert_many(keys[i:i + 1], want_chain=True)
        fails += int(not ok[0])
        if not ok[0]:
            want_hist[-1] += 1
        elif chain[0]:
            want_hist[chain[0]] += 1
        assert occ_gpu == occ_cpu, (i, occ_gpu, occ_cpu)
    np.testing.assert_array_equal(f.exportTable(), of.table_bytes())
    assert fails > 0  # exhausted chains were exercised
    assert f.evictionCount() == of.evictions()
    # the eviction-chain histogram, bin for bin
    np.testing.assert_array_equal(f.evictionHistogram(), want_hist)
    # ordered single-key deletes of every other key
    out = __import__("torch").zeros(1, dtype=__import__("torch").bool, device=dk.device)
    for i in range(0, n, 2):
        occ_gpu = f.deleteMany(dk[i:i + 1], out=out)
        occ_cpu, ok = of.delete_many(keys[i:i + 1])
        assert occ_gpu == occ_cpu and bool(out.item()) == bool(ok[0]), i
    np.testing.assert_array_equal(f.exportTable(), of.table_bytes())


@pytest.mark.parametrize("path", ["binned", "auto"])
def test_single_key_sorted_calls_bit_exact(cf, oracle_mod, path, monkeypatch):
    """insertManySorted one key per call, down the binned path (forced) and
    the default one (single keys go one thread per key)."""
    monkeypatch.setenv("CUCKOO_INSERT_PATH", path)
    O = oracle_mod
    c = cf.make_config(bits_per_tag=16, eviction="bfs", max_evictions=40)
    oc = _oracle_cfg(O, c)
    of = O.Filter(oc, 1024)
    f = cf.CuckooFilter(1024, config=c)
    keys = O.gen_keys(41, 1080)
    dk = _dev_keys(cf, keys)
    for i in range(keys.size):
        assert f.insertManySorted(dk[i:i + 1]) == of.insert_many_sorted(keys[i:i + 1])
    np.testing.assert_array_equal(f.exportTable(), of.table_bytes())


# ---------------------------------------------------------------------------
# concurrent batches: invariants and binomial parity (2^20 slots, alpha 0.95)
# ---------------------------------------------------------------------------

def _check_invariants(O, oc, image, keys, occupied, nb):
    tags = _tags(image, oc)
    b, s = np.nonzero(tags)
    stored = tags[b, s].astype(np.uint64)
    assert stored.size == occupied
    i1, i2, fp = O.candidates(oc, nb, keys)
    fp2 = fp if oc.policy != 2 else np.array([O.alt_bucket(oc, nb, int(x), int(y))[1] for x, y in zip(i1, fp)],
                                               np.uint32)
    pairs = np.union1d((i1 << np.uint64(32)) | fp, (i2 << np.uint64(32)) | fp2)
    assert np.isin((b.astype(np.uint64) << np.uint64(32)) | stored, pairs).all()
    return stored, fp


BATCH_CFGS = [
    dict(eviction="bfs"),
    dict(eviction="dfs"),
    dict(eviction="bfs", block_size=64),
    dict(eviction="bfs", block_size=512),
    dict(eviction="bfs", word_bytes=4),
    dict(eviction="bfs", bits_per_tag=8, word_bytes=4, bucket_size=16),
    dict(eviction="bfs", bits_per_tag=32, bucket_size=8),
    dict(eviction="bfs", policy="addsub"),
    dict(eviction="dfs", policy="offset"),
]


@pytest.mark.parametrize("mode", ["direct", "binned", "sorted", "presorted"])
@pytest.mark.parametrize("kw", BATCH_CFGS, ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_batch_insert_invariants_and_binomial_parity(cf, oracle_mod, kw, mode, monkeypatch):
    """mode: insertMany's direct and binned paths, insertManySorted's binned
    path, and insertManySorted down the reference's own presorted route
    (packed tags, full sort, insertKernelSorted)."""
    sorted_insert = mode in ("sorted", "presorted")
    monkeypatch.setenv("CUCKOO_INSERT_PATH", "binned" if mode == "sorted" else mode)
    O = oracle_mod
    c = cf.make_config(**kw)
    oc = _oracle_cfg(O, c)
    cap = 1 << 20
    of = O.Filter(oc, cap)
    f = cf.CuckooFilter(cap, config=c)
    nb = f.getNumBuckets()
    n = int(f.capacity() * 0.95)
    keys = O.gen_keys(42, n + (1 << 19))
    pos, neg = keys[:n], keys[n:]
    dpos = _dev_keys(cf, pos)
    occ = f.insertManySorted(dpos) if sorted_insert else f.insertMany(dpos)
    occ_cpu, ok_cpu = of.insert_many(pos)
    image = f.exportTable()
    table = assert_table_count(f, occ, n)
    stored, fp = _check_invariants(O, oc, image, pos, table, nb)
    fails_gpu, fails_cpu = n - occ, int((~ok_cpu).sum())
    # failures: both essentially zero at 95% load; allow a binomial 99% band
    band = 3 * np.sqrt(max(fails_cpu, 1)) + 3
    assert abs(fails_gpu - fails_cpu) <= band, (fails_gpu, fails_cpu)
    if fails_gpu == 0:
        # the Offset policy stores the tag with its choice bit flipped in i2
        pure = np.uint64((1 << (c.bits_per_tag - 1)) - 1) if c.policy == 2 else np.uint64(~0 & ((1 << 64) - 1))
        # every inserted fingerprint is stored; extras are stray duplicates of stored tags
        s_vals, s_cnt = np.unique(stored & pure, return_counts=True)
        f_vals, f_cnt = np.unique(fp.astype(np.uint64) & pure, return_counts=True)
        assert np.isin(s_vals, f_vals).all()
        at = np.searchsorted(s_vals, f_vals)
        assert (at < s_vals.size).all() and (s_vals[np.minimum(at, s_vals.size - 1)] == f_vals).all()
        extra = s_cnt[at] - f_cnt
        assert (extra >= 0).all() and int(extra.sum()) == table - occ
        assert f.containsMany(dpos).all().item()
    # FPR: two-sample binomial test against the oracle at the same load
    fp_gpu = int(f.containsMany(_dev_keys(cf, neg)).sum().item())
    fp_cpu = int(of.contains_many(neg).sum())
    pooled = (fp_gpu + fp_cpu) / (2 * neg.size)
    sd = np.sqrt(2 * neg.size * pooled * (1 - pooled)) + 1
    assert abs(fp_gpu - fp_cpu) < 3.3 * sd, (fp_gpu, fp_cpu)
    # delete half: every delete of an inserted key succeeds
    if fails_gpu == 0:
        out = __import__("torch").zeros(n // 2, dtype=__import__("torch").bool, device=dpos.device)
        rem = f.deleteMany(dpos[: n // 2], out=out)
        assert out.all().item() and rem == n - n // 2
        assert_table_count(f, rem, n)
        assert f.containsMany(dpos[n // 2:]).all().item()
    f.close()


# Tables of 2^9 slices and more split the slice index into two digits.  Where
# the bucket bits below the first digit fit beside the tag in 32 bits, the
# binned insert partitions in two levels (MSD) with 32-bit entries; otherwise
# the first level keeps whole 64-bit entries, and the second writes 32-bit
# entries where the bucket's place in its slice fits beside the tag (16-bit
# tags in large tables) or 64-bit ones (32-bit tags).
TWO_LEVEL_CFGS = [
    (dict(eviction="bfs"), 1 << 23),                                      # 512 slices: digits 8 + 1
    (dict(eviction="dfs"), 1 << 24),                                      # 8 + 2
    (dict(eviction="bfs", bits_per_tag=8, word_bytes=4, bucket_size=16), 1 << 24),  # 16-byte buckets
    (dict(eviction="bfs", bits_per_tag=8, bucket_size=4, word_bytes=4), 1 << 24),   # 4-byte buckets, 2^13-bucket slices
    (dict(eviction="bfs", bits_per_tag=32, bucket_size=8), 1 << 23),      # 64-bit entries in both levels
    # 16-bit tags, 8-byte buckets, 2^25 buckets in 2^13 slices: 12 + 5 + 16 bits do
    # not fit in 32, so the first level keeps 64-bit entries and the second packs to 32
    (dict(eviction="bfs", bucket_size=4, word_bytes=8), 1 << 27),
    (dict(eviction="bfs", policy="addsub"), 1 << 23),
    # non-power-of-two bucket counts: 781,250 buckets in 763 slices (the last
    # partial), 393,988 in 385; some first-level partitions stay empty
    (dict(eviction="bfs", policy="addsub"), 12_500_000),
    (dict(eviction="dfs", policy="offset"), 6_303_801),
]


@pytest.mark.parametrize("partition", ["default", "lsd"])
@pytest.mark.parametrize("mode", ["binned", "sorted"])
@pytest.mark.parametrize("kw,cap", TWO_LEVEL_CFGS, ids=lambda x: "-".join(f"{k}{v}" for k, v in x.items())
                         if isinstance(x, dict) else f"cap{x}")
def test_partition_levels_insert_invariants(cf, oracle_mod, kw, cap, mode, partition, monkeypatch):
    import torch

    monkeypatch.setenv("CUCKOO_INSERT_PATH", "binned")
    if partition == "lsd":
        monkeypatch.setenv("CUCKOO_PARTITION", "lsd")
    c = cf.make_config(**kw)
    f = cf.CuckooFilter(cap, config=c)
    load = 0.95 if c.bucket_size >= 8 else 0.85
    n = int(f.capacity() * load)
    host = oracle_mod.gen_keys(77, n + (1 << 20), c.key_bytes)
    keys = _dev_keys(cf, host)
    pos, neg = keys[:n], keys[n:]
    f.setEvictionHistogram(True)
    occ = f.insertManySorted(pos) if mode == "sorted" else f.insertMany(pos)
    fails = n - occ
    # every key that was not placed is an exhausted eviction chain: the
    # partition lost no entry
    assert fails == int(f.evictionHistogram()[-1])
    if fails:
        # exhausted chains are rare (about 1e-6 of the keys here) and their
        # count depends on the concurrent schedule; it stays of the order of
        # the sequential oracle's
        of = oracle_mod.Filter(_oracle_cfg(oracle_mod, c), cap)
        fails_cpu = int((~of.insert_many(host[:n])[1]).sum())
        assert fails <= 4 * fails_cpu + 10, (fails, fails_cpu)
    assert_table_count(f, occ, n)
    assert n - int(f.containsMany(pos).sum().item()) <= fails
    # false positives: binomial band around the closed form at the reached load
    p = 1 - (1 - 1 / (2 ** c.bits_per_tag - 1)) ** (2 * c.bucket_size * occ / f.capacity())
    m = neg.numel()
    hits = int(f.containsMany(neg).sum().item())
    assert abs(hits - m * p) < 5 * np.sqrt(m * p * (1 - p)) + 3, (hits, m * p)
    out = torch.zeros(n // 3, dtype=torch.bool, device=pos.device)
    rem = f.deleteMany(pos[: n // 3], out=out)
    deleted = int(out.sum().item())
    assert deleted >= n // 3 - fails and rem == occ - deleted
    assert n - n // 3 - int(f.containsMany(pos[n // 3:]).sum().item()) <= fails
    f.close()


@pytest.mark.parametrize("mode", ["insert", "sorted"])
@pytest.mark.parametrize("n", [1, 7, 4095, 4097, 12287, 12289, 40961, 100003])
@pytest.mark.parametrize("kw", [dict(eviction="bfs"), dict(eviction="bfs", bits_per_tag=8, word_bytes=4)],
                         ids=["t16", "t8"])
def test_two_level_partition_small_and_ragged_batches(cf, oracle_mod, kw, n, mode, monkeypatch):
    """Batches far smaller than the table through the two-level partition
    (forced binned path, 2^24 slots: 1024 slices or more): single keys,
    sizes around the 4096-value tile and the 12288-value first-pass tile, and
    partitions that stay empty.  At this load nothing fails, every stored tag
    sits in one of its key's two buckets, and every key is found, then
    deleted."""
    import torch

    monkeypatch.setenv("CUCKOO_INSERT_PATH", "binned")
    O = oracle_mod
    c = cf.make_config(**kw)
    oc = _oracle_cfg(O, c)
    f = cf.CuckooFilter(1 << 24, config=c)
    nb = f.getNumBuckets()
    insert = f.insertManySorted if mode == "sorted" else f.insertMany
    host = O.gen_keys(900 + n, n)
    keys = _dev_keys(cf, host)
    occ = insert(keys)
    assert occ == n
    table = assert_table_count(f, occ, n)
    _check_invariants(O, oc, f.exportTable(), host, table, nb)
    assert f.containsMany(keys).all().item()
    # a second batch on top (the partition's scratch is reused)
    host2 = O.gen_keys(1900 + n, n)
    keys2 = _dev_keys(cf, host2)
    assert insert(keys2) == 2 * n
    assert f.containsMany(keys2).all().item() and f.containsMany(keys).all().item()
    out = torch.zeros(n, dtype=torch.bool, device=keys.device)
    assert f.deleteMany(keys, out=out) == n and out.all().item()
    assert f.containsMany(keys2).all().item()
    f.close()


@pytest.mark.parametrize("path", ["direct", "binned"])
def test_eviction_histogram_accounts_for_every_insert(cf, oracle_mod, path, monkeypatch):
    monkeypatch.setenv("CUCKOO_INSERT_PATH", path)
    c = cf.make_config(eviction="bfs", max_evictions=500)
    f = cf.CuckooFilter(1 << 20, config=c)
    f.setEvictionHistogram(True)
    n = int(f.capacity() * 0.98)
    keys = _dev_keys(cf, oracle_mod.gen_keys(5, n))
    occ = f.insertMany(keys)
    h = f.evictionHistogram()
    fails = int(h[-1])
    assert occ == n - fails
    swaps = int(sum(i * int(v) for i, v in enumerate(h[:-1]))) + fails * 500
    assert swaps == f.evictionCount()
    assert h[1] > 0  # BFS moves happen at 98% load


def test_empty_and_ragged_batches(cf, oracle_mod):
    import torch

    f = cf.CuckooFilter(10000, key_bytes=8, bits_per_tag=16)
    empty = torch.empty(0, dtype=torch.int64, device="cuda")
    assert f.insertMany(empty) == 0
    assert f.insertManySorted(empty) == 0
    assert f.deleteMany(empty) == 0
    assert f.containsMany(empty).numel() == 0
    for n in (1, 63, 64, 65, 255, 257, 1000):
        keys = oracle_mod.gen_keys(100 + n, n)
        f.clear()
        assert f.insertMany(_dev_keys(cf, keys)) == n
        assert f.containsMany(_dev_keys(cf, keys)).all().item()


@pytest.mark.parametrize("block", [1, 33, 100, 1000, 1024])
def test_direct_paths_with_any_block_size(cf, oracle_mod, block, monkeypatch):
    """The config's block size is a run-time value (1..1024) for the
    one-thread-per-key kernels, and workgroups that are not a multiple of 64
    leave a partial last wave (blockCountAdd then skips the wave reduction).
    insertMany and deleteMany down the direct paths and containsMany, at
    2^16 slots and 90% load: the counter matches the table, every key is
    found, lookups on the exported image equal the oracle's bit for bit, and
    every key's delete succeeds, bringing the counter back to 0."""
    import torch

    monkeypatch.setenv("CUCKOO_INSERT_PATH", "direct")
    monkeypatch.setenv("CUCKOO_DELETE_PATH", "direct")
    O = oracle_mod
    c = cf.make_config(eviction="bfs", block_size=block)
    oc = _oracle_cfg(O, c)
    cap = 1 << 16
    f = cf.CuckooFilter(cap, config=c)
    n = int(f.capacity() * 0.9)
    keys = O.gen_keys(700 + block, n + 8192)
    dk = _dev_keys(cf, keys)
    occ = f.insertMany(dk[:n])
    assert occ == n
    assert_table_count(f, occ, n)
    got = f.containsMany(dk).cpu().numpy()
    assert got[:n].all()
    of = O.Filter(oc, cap)
    of.load_table(f.exportTable(), occ)
    np.testing.assert_array_equal(got, of.contains_many(keys))
    out = torch.empty(n, dtype=torch.bool, device=dk.device)
    assert f.deleteMany(dk[:n], out=out) == 0
    assert out.all().item()
    f.close()
