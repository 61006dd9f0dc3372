"""cf_insert_many_noalloc against cf_insert_many, through the C ABI.

The no-allocation entry point runs the same insert with its scratch in the
library's fixed per-device array.  Two calls in a row on two filters, the
second with fewer keys, both down the bucket-binned path (the one that uses
scratch):
  * at 25% load of 32-slot buckets no bucket overflows, so every key lands in
    its primary bucket and each bucket's tag multiset is fixed by the keys:
    the two filters' images must agree bucket by bucket (slots within a
    bucket may be taken in another order by the concurrent insert);
  * at 95% load of 16-slot buckets the concurrent insert's invariants hold
    for both: occupancy equal, counter equal to the non-empty slots up to BFS
    strays, no inserted key missing;
  * more keys than the limit are refused with CF_ERR_INVALID.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_table_count

pytestmark = pytest.mark.gpu


def _bucket_multisets(image, bucket_size):
    return np.sort(image.view(np.uint16).reshape(-1, bucket_size), axis=1)


@pytest.mark.parametrize("load,bucket", [(0.25, 32), (0.95, 16)])
def test_noalloc_insert_matches_insert_many(cf, oracle_mod, load, bucket, monkeypatch):
    monkeypatch.setenv("CUCKOO_INSERT_PATH", "binned")
    c = cf.make_config(key_bytes=8, bits_per_tag=16, bucket_size=bucket, eviction="bfs")
    cap = (1 << 23) if load <= 0.25 else (1 << 21)
    a = cf.CuckooFilter(cap, config=c)   # cf_insert_many_noalloc
    b = cf.CuckooFilter(cap, config=c)   # cf_insert_many
    total = int(a.capacity() * load)
    n1 = total * 7 // 8
    n2 = total - n1
    assert n1 <= cf.CF_NOALLOC_MAX_KEYS
    keys = oracle_mod.gen_keys(2024 + bucket, n1 + n2 + (1 << 16))
    dk = cf.keys_to_device(keys)
    total = 0
    for lo, hi in ((0, n1), (n1, n1 + n2)):
        occ_a = a.insertManyNoAlloc(dk[lo:hi])
        occ_b = b.insertMany(dk[lo:hi])
        total += hi - lo
        print(f"call [{lo}, {hi}): occupancy {occ_a} (no allocation) / {occ_b}", flush=True)
        assert occ_a == occ_b == total
        assert_table_count(a, occ_a, total)
        assert_table_count(b, occ_b, total)
        assert a.containsMany(dk[:hi]).all().item() and b.containsMany(dk[:hi]).all().item()
        if load <= 0.25:
            np.testing.assert_array_equal(_bucket_multisets(a.exportTable(), bucket),
                                          _bucket_multisets(b.exportTable(), bucket))
    fa = int(a.containsMany(dk[n1 + n2:]).sum().item())
    fb = int(b.containsMany(dk[n1 + n2:]).sum().item())
    assert abs(fa - fb) <= 5 * np.sqrt(fa + fb + 1)
    a.close()
    b.close()


def test_noalloc_insert_refuses_too_many_keys(cf):
    import torch

    f = cf.CuckooFilter(1 << 24, key_bytes=8, bits_per_tag=16)
    keys = torch.zeros((1 << 21) + 1, dtype=torch.int64, device="cuda")
    with pytest.raises(cf.CuckooError) as e:
        f.insertManyNoAlloc(keys)
    assert e.value.code == cf.CF_ERR_INVALID and "at most" in str(e.value)
    assert f.occupiedSlots() == 0
    f.close()
