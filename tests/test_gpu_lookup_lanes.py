"""containsMany on 64- and 128-byte buckets, read by several lanes per key.

containsKernel reads a 128-byte bucket, and a 64-byte bucket of a table
past 2 GiB, with one lane per 32 bytes (two or four lanes per key, a ballot
over the group) when the block size is a multiple of that lane count, and
with one lane per key otherwise (csrc/cuckoo/kernels.hpp).  Either way the
result must be the reference's lookup (include/CuckooFilter.cuh:1088-1098,
1143-1155) bit for bit:

  * on oracle-built images of every 64/128-byte-bucket configuration (and the
    AddSub and Offset policies), at block sizes that take the grouped path
    (4, 64, 256, 1024) and ones that do not (100, 250), with a ragged number
    of keys;
  * on 4 GiB tables (2^30 slots, 32-bit tags; buckets of 32 slots in u64
    words, of 16 in u64 and u32 words) the GPU filled to 95%: 2^24 lookups,
    half of them inserted keys, against the oracle on the exported image;
    then 2^24 deletes (the binned path with 64-bit partition entries) through
    the delete-parity check of tests/delete_parity.py.
"""
import numpy as np
import pytest

from conftest import assert_table_count
from delete_parity import check_delete_parity

pytestmark = pytest.mark.gpu

CONFIGS = [
    dict(bits_per_tag=16, bucket_size=32, word_bytes=4),
    dict(bits_per_tag=16, bucket_size=32, word_bytes=8),
    dict(bits_per_tag=32, bucket_size=16, word_bytes=4),
    dict(bits_per_tag=32, bucket_size=16, word_bytes=8),
    dict(bits_per_tag=32, bucket_size=32, word_bytes=4),
    dict(bits_per_tag=32, bucket_size=32, word_bytes=8),
    dict(bits_per_tag=32, bucket_size=32, word_bytes=8, policy="addsub"),
    dict(bits_per_tag=32, bucket_size=16, word_bytes=8, policy="offset"),
    dict(bits_per_tag=16, bucket_size=32, word_bytes=8, policy="offset"),
]


@pytest.mark.parametrize("block", [4, 64, 256, 1024, 100, 250])
@pytest.mark.parametrize("kw", CONFIGS, ids=lambda k: "-".join(f"{v}" for v in k.values()))
def test_grouped_lookup_on_oracle_image(cf, oracle_mod, kw, block):
    O = oracle_mod
    c = cf.make_config(key_bytes=8, eviction="bfs", block_size=block, **kw)
    if not cf.config_supported(c):
        pytest.skip("configuration not built")
    oc = O.make_config(**c.as_dict())
    cap = 1 << 16
    of = O.Filter(oc, cap)
    n = int(of.capacity() * 0.95)
    m = n + 12_345
    keys = O.gen_keys(900 + block, n + m)
    occ, _ = of.insert_many(keys[:n])
    f = cf.CuckooFilter(cap, config=c)
    f.importTable(of.table_bytes(), occ)
    q = np.concatenate([keys[:n], keys[n:n + m]])[: n + m - 7]  # ragged: not a multiple of any group
    got = f.containsMany(cf.keys_to_device(q)).cpu().numpy()
    want = of.contains_many(q)
    mism = np.nonzero(got != want)[0]
    assert mism.size == 0, f"{mism.size} lookups differ, first at {mism[0]}"
    if occ == n:
        assert got[:n].all()
    f.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("bucket,word", [(32, 8), (16, 8), (16, 4)])
def test_grouped_lookup_on_4gib_table(cf, oracle_mod, bucket, word):
    import torch
    import workloads

    O = oracle_mod
    c = cf.make_config(key_bytes=8, bits_per_tag=32, bucket_size=bucket, word_bytes=word, eviction="bfs")
    oc = O.make_config(**c.as_dict())
    cap = 1 << 30
    f = cf.CuckooFilter(cap, config=c)
    assert f.sizeInBytes() == 4 << 30
    n = int(cap * 0.95)
    nq = 1 << 23
    keys = workloads.gen_keys_torch(4096 + bucket, n + nq)
    occ = f.insertMany(keys[:n])
    print(f"4 GiB table: {occ} of {n} placed", flush=True)
    assert occ == n
    assert_table_count(f, occ, n)
    q = torch.cat([keys[:nq], keys[n:]])
    got = f.containsMany(q).cpu().numpy()
    qh = q.cpu().numpy().view(np.uint64)
    nd = 1 << 24  # a binned delete on 2^25 buckets (its 64-bit-entry partition)
    dkeys = keys[:nd].clone()
    del keys, q
    torch.cuda.empty_cache()
    image = f.exportTable()
    of = O.Filter(oc, cap)
    of.load_table(image, occ)
    want = of.contains_many(qh, threads=16)
    del of
    mism = np.nonzero(got != want)[0]
    assert mism.size == 0, f"{mism.size} lookups differ, first at {mism[0]}"
    assert got[:nq].all()
    print(f"2^24 lookups bit-exact on the 4 GiB image ({int(got[nq:].sum())} false positives)", flush=True)
    out = torch.zeros(nd, dtype=torch.bool, device=dkeys.device)
    rem = f.deleteMany(dkeys, out=out)
    st = check_delete_parity(O, oc, cap, image, occ, dkeys.cpu().numpy().view(np.uint64), out.cpu().numpy(),
                             f.exportTable(), rem)
    print(f"2^24 deletes on the 4 GiB image: {st}", flush=True)
    assert st["successes"] == nd
    f.close()
