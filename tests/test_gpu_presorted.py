"""The reference's presorted insert route on the GPU, through the C ABI.

insertManySorted in the reference computes packed tags
(computePackedTagsKernel, include/CuckooFilter.cuh:1185-1207), sorts them
with cub::DeviceRadixSort (:485-498) and inserts them with insertKernelSorted
(:1209-1266); its InsertPresorted benchmark launches those three steps itself
(benchmark/cuckoo_sorted_vs_unsorted.cu:48-86).  Here the same steps are
cf_compute_packed_tags, cf_sort_packed_tags and cf_insert_packed_sorted, and
insertManySorted takes them end to end under CUCKOO_INSERT_PATH=presorted.

  * the sort is bit-exact against numpy's sort of the oracle's packed tags;
  * one-key-per-call inserts down the route give the sequential oracle's
    table image after every call, eviction chains and exhausted chains
    included;
  * a whole batch keeps the invariants of the concurrent insert (tests in
    test_gpu_parity.py run the "presorted" mode over nine configurations).
"""
import numpy as np
import pytest

from conftest import assert_table_count

pytestmark = pytest.mark.gpu

PK_DT = {True: np.uint32, False: np.uint64}  # PackedTagType: u32 when bits_per_tag <= 8


def _packed_host(cf, t, bits):
    return cf.to_numpy_u(t, PK_DT[bits <= 8]).astype(np.uint64)


# (config, capacity): the LSD passes over bits_per_tag + log2(numBuckets)
# bits number 3 (8-bit tags, 2^16 buckets: 24 bits), 4 (16-bit, 2^16: 32), 5
# (16-bit, 2^24: 40) and 6 (32-bit, 2^16: 48), so the result lands in the
# workspace (odd counts, then copied back) and in place (even counts); 8-bit
# tags at 2^26 buckets keep only the 24 bucket bits a u32 packed tag holds
SORT_CASES = [
    (dict(bits_per_tag=8, word_bytes=4), 1 << 20),
    (dict(bits_per_tag=16), 1 << 20),
    (dict(bits_per_tag=16), 1 << 28),
    (dict(bits_per_tag=32), 1 << 20),
    (dict(bits_per_tag=8, word_bytes=4), 1 << 30),
    (dict(bits_per_tag=16, policy="addsub"), 12_500_000),
]


@pytest.mark.parametrize("n", [0, 1, 2, 4095, 4097, 100_003, 1 << 20])
@pytest.mark.parametrize("kw,cap", SORT_CASES, ids=lambda x: str(x))
def test_sort_packed_tags_is_the_full_ascending_order(cf, oracle_mod, kw, cap, n):
    O = oracle_mod
    c = cf.make_config(**kw)
    f = cf.CuckooFilter(cap, config=c)
    nb = f.getNumBuckets()
    keys = O.gen_keys(7 + n, max(n, 1), c.key_bytes)[:n]
    want = np.sort(O.packed_tags(O.make_config(**c.as_dict()), nb, keys))
    import torch
    dk = cf.keys_to_device(keys) if n else torch.empty(0, dtype=torch.int64, device="cuda")
    packed = f.computePackedTags(dk)
    np.testing.assert_array_equal(np.sort(_packed_host(cf, packed, c.bits_per_tag)), want)
    f.sortPackedTags(packed)
    np.testing.assert_array_equal(_packed_host(cf, packed, c.bits_per_tag), want)
    f.close()


SEQ_CFGS = [
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="bfs", max_evictions=40),
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="dfs", max_evictions=40),
    dict(bits_per_tag=8, bucket_size=16, word_bytes=4, eviction="dfs", max_evictions=30),
    dict(bits_per_tag=32, bucket_size=8, word_bytes=8, eviction="bfs", max_evictions=30),
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="bfs", policy="addsub", max_evictions=30),
    dict(bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="dfs", policy="offset", max_evictions=30),
    dict(bits_per_tag=8, bucket_size=8, word_bytes=8, eviction="bfs", policy="offset", max_evictions=30),
]


@pytest.mark.parametrize("entry", ["insertManySorted", "insertPackedSorted"])
@pytest.mark.parametrize("kw", SEQ_CFGS, ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_presorted_single_key_calls_bit_exact(cf, oracle_mod, kw, entry, monkeypatch):
    """insertKernelSorted, one key per call, against the sequential oracle:
    occupancy after every call, then the table image and eviction count."""
    monkeypatch.setenv("CUCKOO_INSERT_PATH", "presorted")
    O = oracle_mod
    c = cf.make_config(**kw)
    oc = O.make_config(**c.as_dict())
    cap = 64 * c.bucket_size
    of = O.Filter(oc, cap)
    f = cf.CuckooFilter(cap, config=c)
    n = int(of.capacity() * 1.06)
    keys = O.gen_keys(53, n, c.key_bytes)
    dk = cf.keys_to_device(keys)
    packed = f.computePackedTags(dk) if entry == "insertPackedSorted" else None
    fails = 0
    for i in range(n):
        if packed is None:
            occ_gpu = f.insertManySorted(dk[i:i + 1])
        else:
            occ_gpu = f.insertPackedSorted(packed[i:i + 1])
        before = of.occupied()
        occ_cpu = of.insert_many_sorted(keys[i:i + 1])
        fails += int(occ_cpu == before)
        assert occ_gpu == occ_cpu, (i, occ_gpu, occ_cpu)
    np.testing.assert_array_equal(f.exportTable(), of.table_bytes())
    assert fails > 0  # exhausted chains were exercised
    assert f.evictionCount() == of.evictions()
    f.close()


@pytest.mark.parametrize("kw", [dict(eviction="bfs"), dict(eviction="dfs"), dict(bits_per_tag=8, word_bytes=4),
                                dict(eviction="bfs", policy="offset")],
                         ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_presorted_flow_batch(cf, oracle_mod, kw):
    """The InsertPresorted benchmark's three steps on a 2^20-slot filter at
    95% load: sorted packed tags equal numpy's sort of the oracle's, then
    insertKernelSorted keeps the table count and loses no key."""
    import torch

    O = oracle_mod
    c = cf.make_config(**kw)
    oc = O.make_config(**c.as_dict())
    f = cf.CuckooFilter(1 << 20, config=c)
    nb = f.getNumBuckets()
    n = int(f.capacity() * 0.95)
    keys = O.gen_keys(61, n, c.key_bytes)
    dk = cf.keys_to_device(keys)
    packed = f.sortPackedTags(f.computePackedTags(dk))
    np.testing.assert_array_equal(_packed_host(cf, packed, c.bits_per_tag), np.sort(O.packed_tags(oc, nb, keys)))
    occ = f.insertPackedSorted(packed)
    assert occ >= n - 2, (occ, n)
    table = assert_table_count(f, occ, n)
    missing = int((~f.containsMany(dk)).sum().item())
    # Offset + BFS: after a lost swap the reference's clean-up removes the
    # unflipped candidate tag (include/CuckooFilter.cuh:1011-1012), which can
    # be another key's copy; each lost swap leaves a stray or such a removal
    quirk = table - occ if kw.get("policy") == "offset" and kw.get("eviction") == "bfs" else 0
    assert missing <= n - occ + quirk, (missing, n - occ, quirk)
    if occ == n and missing == 0:
        out = torch.zeros(n // 2, dtype=torch.bool, device=dk.device)
        assert f.deleteMany(dk[: n // 2], out=out) == n - n // 2 and out.all().item()
    f.close()


# insertKernelSorted places the tags of consecutive lanes that share a primary
# bucket together (one bucket read, one CAS per changed word); these cases
# cover runs split over waves, partial waves (block sizes that are not a
# multiple of 64), buckets of 16 to 64 bytes and the three policies
RUN_CFGS = [dict(), dict(bits_per_tag=8, word_bytes=4), dict(bits_per_tag=32, bucket_size=16),
            dict(bits_per_tag=16, bucket_size=32), dict(bits_per_tag=16, bucket_size=8),
            dict(policy="offset"), dict(policy="addsub")]


@pytest.mark.parametrize("block", [64, 100, 256, 1024])
@pytest.mark.parametrize("kw", RUN_CFGS, ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()) or "default")
def test_presorted_runs_hold_each_bucket_s_tags(cf, oracle_mod, kw, block):
    """With no bucket receiving more primary tags than it has slots, no insert
    leaves its primary bucket, whatever the interleaving: after one batch
    every bucket holds exactly the fingerprints of the keys whose primary it
    is (as a multiset: the order inside a bucket depends on which run of a
    bucket split over two waves commits first)."""
    O = oracle_mod
    c = cf.make_config(block_size=block, **kw)
    if not cf.config_supported(c):
        pytest.skip("configuration not built")
    oc = O.make_config(**c.as_dict())
    cap = 1 << 18
    f = cf.CuckooFilter(cap, config=c)
    nb = f.getNumBuckets()
    bs = c.bucket_size
    # a quarter of the slots; fewer for small buckets and for AddSub, whose
    # primary buckets are not uniform (half of them receive no key)
    n = cap // 10 if bs < 16 else (cap // 16 if kw.get("policy") == "addsub" else cap // 4)
    keys = O.gen_keys(97 + block, n, c.key_bytes)
    i1, _, fp = O.candidates(oc, nb, keys)
    i1 = i1.astype(np.int64)
    per = np.bincount(i1, minlength=nb)
    assert per.max() <= bs, "the keys must not overfill a primary bucket"
    dk = cf.keys_to_device(keys)
    packed = f.sortPackedTags(f.computePackedTags(dk))
    assert f.insertPackedSorted(packed) == n
    got = np.sort(O.table_tags(f.exportTable(), oc), axis=1)
    # expected rows, sorted: the bucket's empty slots (0) first, then its tags
    order = np.lexsort((fp, i1))
    b = i1[order]
    start = np.concatenate([[0], np.cumsum(per)])[:-1]
    col = bs - per[b] + (np.arange(n) - start[b])
    want = np.zeros_like(got)
    want[b, col] = fp[order].astype(got.dtype)
    mism = np.nonzero((got != want).any(axis=1))[0]
    assert mism.size == 0, f"{mism.size} buckets differ, first {mism[:5]}"
    assert f.countOccupiedSlots() == n
    f.close()


@pytest.mark.parametrize("block", [100, 1024])
@pytest.mark.parametrize("eviction", ["bfs", "dfs"])
def test_presorted_duplicate_runs_keep_every_copy(cf, oracle_mod, eviction, block):
    """Keys repeated up to 20 times give runs longer than a 16-slot bucket,
    many of them crossing a wave: the copies beyond a full primary bucket go
    on to the secondary bucket and eviction, one lane each.  Twenty copies of
    one fingerprint crowd its two buckets, so some inserts fail in the
    sequential oracle too (3-12 of these 15,500); the batch may fail a few
    more, the table agrees with the counter, and every failure loses at most
    one stored copy (the tag it drops), so the copies missing across all
    keys number no more than the failures."""
    O = oracle_mod
    c = cf.make_config(eviction=eviction, block_size=block)
    oc = O.make_config(**c.as_dict())
    cap = 1 << 16
    f = cf.CuckooFilter(cap, config=c)
    nb = f.getNumBuckets()
    rng = np.random.default_rng(5 + block)
    base = O.gen_keys(73, 1500, c.key_bytes)
    reps = rng.integers(1, 21, size=base.size)
    keys = np.repeat(base, reps)
    rng.shuffle(keys)
    n = keys.size
    of = O.Filter(oc, cap)
    occ_cpu, _ = of.insert_many(keys[np.argsort(O.packed_tags(oc, nb, keys), kind="stable")])
    dk = cf.keys_to_device(keys)
    packed = f.sortPackedTags(f.computePackedTags(dk))
    occ = f.insertPackedSorted(packed)
    assert n - occ <= 3 * (n - occ_cpu) + 10, (n - occ, n - occ_cpu)
    assert_table_count(f, occ, n)
    i1, i2, fp = O.candidates(oc, nb, base)
    tags = O.table_tags(f.exportTable(), oc)
    fpc = fp.astype(tags.dtype)[:, None]
    copies = (tags[i1.astype(np.int64)] == fpc).sum(axis=1)
    copies += np.where(i2 != i1, (tags[i2.astype(np.int64)] == fpc).sum(axis=1), 0)
    missing = int(np.maximum(reps - copies, 0).sum())
    assert missing <= n - occ, (missing, n - occ)
    f.close()


# Inserts whose primary attempt fails continue in the wave's queue with one
# DFS swap per round; a chain that is not done goes back into the queue with
# its carried tag, bucket and swap count.  Short eviction bounds exhaust many
# chains part-way through the queue.  Each insert must be accounted exactly
# once: the histogram's failure bin equals the failed inserts, the swaps the
# histogram records equal the eviction counter, and no placed key is missing.
# The failure count stays in the band of the sequential oracle's over the same
# sorted order (BFS 7682 / 5319 / 0, DFS 28130 / 15133 / 0 for 1 / 3 / 500).
@pytest.mark.parametrize("block", [256, 1024])
@pytest.mark.parametrize("max_ev", [1, 3, 500])
@pytest.mark.parametrize("eviction", ["bfs", "dfs"])
def test_presorted_paused_chains_account_every_insert(cf, oracle_mod, eviction, max_ev, block):
    O = oracle_mod
    c = cf.make_config(eviction=eviction, max_evictions=max_ev, block_size=block)
    f = cf.CuckooFilter(1 << 20, config=c)
    f.setEvictionHistogram(True)
    n = int(f.capacity() * 0.98)
    keys = O.gen_keys(83 + max_ev, n, c.key_bytes)
    dk = cf.keys_to_device(keys)
    occ = f.insertPackedSorted(f.sortPackedTags(f.computePackedTags(dk)))
    failed = n - occ
    hist = f.evictionHistogram().astype(np.int64)
    assert int(hist[max_ev + 1]) == failed, (int(hist[max_ev + 1]), failed)
    swaps = int((np.arange(1, max_ev + 1) * hist[1:max_ev + 1]).sum()) + max_ev * int(hist[max_ev + 1])
    assert swaps == f.evictionCount(), (swaps, f.evictionCount())
    assert_table_count(f, occ, n)
    missing = int((~f.containsMany(dk)).sum().item())
    assert missing <= failed, (missing, failed)
    of = O.Filter(O.make_config(**c.as_dict()), 1 << 20)
    want = n - of.insert_many_sorted(keys)
    assert failed <= 2 * want + 100, (failed, want)
    if want > 1000:
        assert failed >= want // 2, (failed, want)
    f.close()
