"""Parity at the HBM-scale workload: 2^28 slots (2^24 buckets x 16 x 16-bit
tags, 512 MiB), uint64 keys, XOR, BFS, maxEvictions 500.

  * test_oracle_image_lookups_and_deletes: the sequential oracle fills the
    table to 95% (255,013,683 keys) and the GPU imports its image.
    containsMany on 2^27 keys (half inserted) equals the oracle's lookups bit
    for bit.  deleteMany of a quarter of the keys then passes the
    delete-parity check of tests/delete_parity.py: every delete that shares no
    (bucket, fingerprint) with another has the oracle's result, every bucket
    outside those shared pairs ends byte-identical to the oracle's, each group
    of interacting deletes ends as some order of them does, and the counter
    falls by the successes.
  * test_gpu_fill_parity: the GPU fills the table to alpha in {0.5, 0.8, 0.9,
    0.95} with insertMany and with insertManySorted (and at 0.95 down the
    direct path and the reference's presorted route; plus the 2^27-slot,
    8-byte-bucket table whose partition keeps 64-bit entries in its first
    level).  The concurrent fill is checked by its properties (counter against
    non-empty slots, no inserted key missing, FPR in a binomial band); its
    exported image then gets the same bit-exact lookups and delete check.
  * test_sweep_point_at_scale: the 16-bit / bucket-16 / u64 sweep point at
    98% load on 2^28 and 2^30 slots, BFS and DFS: counter against the table,
    the eviction-chain histogram against the counter, no missing key, FPR.

Reference: include/CuckooFilter.cuh:1088-1116 (contains / remove),
:1143-1183 (containsKernel / deleteKernel); benchmark/benchmark_common.cuh:247-333.
"""
import os

import numpy as np
import pytest

from conftest import assert_table_count
from delete_parity import check_delete_parity

pytestmark = pytest.mark.gpu

SLOTS = 1 << 28
NQ = 1 << 26  # positives and negatives per lookup batch (2^27 lookups)
THREADS = max(1, min(16, os.cpu_count() or 1))


def _lookup_and_delete_parity(cf, O, f, oc, cap, keys, dk, n, occ):
    """Bit-exact lookups on the filter's current image, then the delete check."""
    import torch

    image = f.exportTable()
    npos = min(NQ, n)
    q = torch.cat([dk[:npos], dk[n:n + NQ]])
    got = f.containsMany(q).cpu().numpy()
    of = O.Filter(oc, cap)
    of.load_table(image, occ)
    want = of.contains_many(np.concatenate([keys[:npos], keys[n:n + NQ]]), threads=THREADS)
    mism = np.nonzero(got != want)[0]
    assert mism.size == 0, f"{mism.size} lookups differ from the oracle, first at {mism[0]}"
    if occ == n:
        assert got[:npos].all()
    fps = int(got[npos:].sum())
    p = 2 * oc.bucket_size * (occ / cap) / (1 << oc.bits_per_tag)
    assert abs(fps - NQ * p) < 4.5 * np.sqrt(NQ * p) + 2, (fps, NQ * p)
    del of
    nd = n // 4
    out = torch.zeros(nd, dtype=torch.bool, device=dk.device)
    rem = f.deleteMany(dk[:nd], out=out)
    st = check_delete_parity(O, oc, cap, image, occ, keys[:nd], out.cpu().numpy(), f.exportTable(), rem)
    print(f"lookups {q.numel()} bit-exact ({fps} false positives); deletes {st}", flush=True)
    if occ == n:
        assert st["successes"] == nd
        assert f.containsMany(dk[nd:n]).all().item()


@pytest.mark.timeout(900)
def test_oracle_image_lookups_and_deletes(cf, oracle_mod):
    O = oracle_mod
    c = cf.make_config(key_bytes=8, bits_per_tag=16, eviction="bfs")
    oc = O.make_config(**c.as_dict())
    n = int(SLOTS * 0.95)
    keys = O.gen_keys(42, n + NQ)
    of = O.Filter(oc, SLOTS)
    occ, ok = of.insert_many(keys[:n])
    print(f"oracle filled {occ} of {n}", flush=True)
    assert of.num_buckets == 1 << 24 and occ == n
    f = cf.CuckooFilter(SLOTS, config=c)
    f.importTable(of.table_bytes(), occ)
    del of
    dk = cf.keys_to_device(keys)
    _lookup_and_delete_parity(cf, O, f, oc, SLOTS, keys, dk, n, occ)
    f.close()


FILL_CASES = [(a, p) for a in (0.5, 0.8, 0.9, 0.95) for p in ("insertMany", "insertManySorted")] + [
    (0.95, "direct"), (0.95, "presorted"), (0.95, "bucket4-2^27")]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("alpha,path", FILL_CASES, ids=lambda x: str(x))
def test_gpu_fill_parity(cf, oracle_mod, alpha, path, monkeypatch):
    O = oracle_mod
    cap = SLOTS
    kw = dict(key_bytes=8, bits_per_tag=16, eviction="bfs")
    if path == "bucket4-2^27":
        cap = 1 << 27
        kw.update(bucket_size=4)
    if path in ("direct", "presorted"):
        monkeypatch.setenv("CUCKOO_INSERT_PATH", path)
    c = cf.make_config(**kw)
    oc = O.make_config(**c.as_dict())
    f = cf.CuckooFilter(cap, config=c)
    n = int(cap * alpha)
    keys = O.gen_keys(1000 + int(alpha * 100), n + NQ)
    dk = cf.keys_to_device(keys)
    occ = f.insertManySorted(dk[:n]) if path in ("insertManySorted", "presorted") else f.insertMany(dk[:n])
    print(f"{path} at {alpha}: {occ} of {n} placed", flush=True)
    if c.bucket_size == 16:
        assert occ == n  # BFS with a DFS fallback of 500 places every key at these loads
    else:
        assert occ >= n * 0.99
    assert_table_count(f, occ, n)
    _lookup_and_delete_parity(cf, O, f, oc, cap, keys, dk, n, occ)
    f.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("log2_slots", [28, 30])
@pytest.mark.parametrize("eviction", ["bfs", "dfs"])
def test_sweep_point_at_scale(cf, log2_slots, eviction):
    import torch
    import workloads

    cap = 1 << log2_slots
    f = cf.CuckooFilter(cap, key_bytes=8, bits_per_tag=16, bucket_size=16, word_bytes=8, eviction=eviction)
    f.setEvictionHistogram(True)
    n = int(cap * 0.98)
    nq = 1 << 26
    keys = workloads.gen_keys_torch(77 + log2_slots, n + nq)
    occ = f.insertMany(keys[:n])
    hist = f.evictionHistogram()
    failed = int(hist[-1])
    print(f"2^{log2_slots} {eviction} at 0.98: {occ} of {n}, {failed} failed, "
          f"{int(hist[1:-1].sum())} evicted", flush=True)
    assert occ + failed == n
    assert failed <= n // 100000
    assert_table_count(f, occ, n)
    if failed == 0:
        for lo in range(0, n, 1 << 28):
            assert f.containsMany(keys[lo:min(n, lo + (1 << 28))]).all().item()
    fps = int(f.containsMany(keys[n:]).sum().item())
    p = 2 * 16 * (occ / cap) / 65536
    assert abs(fps - nq * p) < 4.5 * np.sqrt(nq * p), (fps, nq * p)
    del keys
    torch.cuda.empty_cache()
    f.close()


@pytest.mark.timeout(300)
def test_table_past_2p32_buckets(cf, oracle_mod):
    """2^33 buckets of four 16-bit slots (a 64 GiB table, XOR): past 2^32
    buckets the binned paths do not apply, so insertMany and deleteMany run
    one thread per key and insertManySorted takes the presorted route (49-bit
    packed tags, insertKernelSorted).  i1 comes from the hash's low word
    (bucket_policies.cuh:56-57), so it stays below 2^32, while half of the
    alternate buckets lie past it.  The keys are chosen so that no two share a
    bucket; each goes in five times, so four copies fill its primary bucket
    and one lands in its alternate whatever the interleaving.  The device's
    (i1, i2, fp) equal the oracle's, the counter counts every copy, every key
    is found, the five deletes per key all succeed and leave the table empty,
    and the presorted route does the same."""
    import torch

    O = oracle_mod
    c = cf.make_config(policy="xor", bits_per_tag=16, bucket_size=4, word_bytes=8, eviction="bfs")
    oc = O.make_config(**c.as_dict())
    cap = 1 << 35
    nb = O.num_buckets(oc, cap)
    assert nb == 1 << 33
    pool = O.gen_keys(3301, 1 << 18, c.key_bytes)
    i1, i2, fp = O.candidates(oc, nb, pool)
    buckets, counts = np.unique(np.concatenate([i1, i2]), return_counts=True)
    once = set(buckets[counts == 1].tolist())
    keep = np.array([a in once and b in once for a, b in zip(i1.tolist(), i2.tolist())])
    keys, i1, i2, fp = pool[keep], i1[keep], i2[keep], fp[keep]
    n = keys.size
    assert n > 1 << 17 and i1.max() < 1 << 32 and (i2 >= 1 << 32).sum() > n // 4
    negatives = O.gen_keys(3302, n, c.key_bytes)
    f = cf.CuckooFilter(cap, config=c)
    assert f.getNumBuckets() == nb
    dk = cf.keys_to_device(keys)
    d1, d2, dfp = f.computeCandidates(dk)
    np.testing.assert_array_equal(d1.cpu().numpy().astype(np.uint64), i1)
    np.testing.assert_array_equal(d2.cpu().numpy().astype(np.uint64), i2)
    np.testing.assert_array_equal(dfp.cpu().numpy().astype(np.uint32), fp)
    five = dk.repeat(5)
    out = torch.empty(five.numel(), dtype=torch.bool, device=dk.device)
    for name, insert in (("insertMany", f.insertMany), ("insertManySorted", f.insertManySorted)):
        print(f"{name}: {five.numel()} inserts into 2^33 buckets", flush=True)
        assert insert(five) == 5 * n, name
        assert f.containsMany(dk).all().item(), name
        assert int(f.containsMany(cf.keys_to_device(negatives)).sum().item()) <= 5, name
        assert f.deleteMany(five, out=out) == 0, name
        assert out.all().item(), name
        assert not f.containsMany(dk).any().item(), name
    f.close()
