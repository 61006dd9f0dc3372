"""deleteMany's two device paths against the oracle, through the C ABI.

A large deleteMany runs bucket-binned: the keys' (primary bucket,
fingerprint) entries are partitioned by table slice, each slice's first-bucket
removals happen in LDS, the rest try the secondary bucket with device CAS, and
the per-key results come from a hash table of the failed deletes
(csrc/cuckoo/binned.hpp).  A small one runs one thread per key
(deleteKernel, include/CuckooFilter.cuh:1157-1183).  Both are checked with the
delete-parity checker (tests/delete_parity.py): every delete that shares no
(bucket, fingerprint) with another has the sequential oracle's result, every
bucket outside the shared pairs ends byte-identical to the oracle's, each
group of interacting deletes ends as some order of them does, and the counter
falls by the successes.  Cases: present keys, absent keys (every delete
fails, so every key goes through the failure table), a half-and-half mix,
repeated keys, no output buffer, the partition's layouts (two-level 32-bit,
64-bit first level, LSD passes, one slice), non-power-of-two bucket counts,
4-slot buckets with failed inserts, and ragged batch sizes.
"""
import numpy as np
import pytest

from delete_parity import check_delete_parity

pytestmark = pytest.mark.gpu

CASES = [
    (dict(eviction="bfs"), 1 << 20),
    (dict(eviction="dfs", word_bytes=4), 1 << 20),
    (dict(eviction="bfs", bits_per_tag=8, word_bytes=4), 1 << 20),
    (dict(eviction="bfs", bits_per_tag=32, bucket_size=8), 1 << 20),
    (dict(eviction="bfs", bucket_size=4), 1 << 20),                # failed inserts at 95%
    (dict(eviction="bfs"), 1 << 23),                                # two-level, 32-bit values
    (dict(eviction="bfs", bits_per_tag=32, bucket_size=8), 1 << 23),  # two-level, 64-bit values
    (dict(eviction="bfs", bucket_size=4, word_bytes=8), 1 << 27),   # 64-bit first level, 32-bit second
    (dict(eviction="bfs", policy="addsub"), 12_500_000),            # non-power-of-two bucket count
    (dict(eviction="bfs"), 1 << 12),                                # one slice
    (dict(eviction="bfs", key_bytes=4), 1 << 20),
]


def _fill(cf, O, kw, cap, seed):
    c = cf.make_config(**kw)
    oc = O.make_config(**c.as_dict())
    f = cf.CuckooFilter(cap, config=c)
    n = int(f.capacity() * 0.95)
    keys = O.gen_keys(seed, 2 * n, c.key_bytes)
    if c.key_bytes == 4:
        keys = np.unique(keys)[: 2 * n]
        np.random.default_rng(seed).shuffle(keys)
        n = keys.size // 2
    dk = cf.keys_to_device(keys)
    occ = f.insertMany(dk[:n])
    return c, oc, f, keys, dk, n, occ


@pytest.mark.parametrize("path", ["binned", "direct"])
@pytest.mark.parametrize("which", ["present", "absent", "mixed"])
@pytest.mark.parametrize("kw,cap", CASES, ids=lambda x: str(x))
def test_delete_parity(cf, oracle_mod, kw, cap, which, path, monkeypatch):
    import torch

    if cap >= 1 << 23 and (which, path) != ("mixed", "binned"):
        pytest.skip("large tables: the mixed batch down the binned path covers them")
    monkeypatch.setenv("CUCKOO_DELETE_PATH", path)
    O = oracle_mod
    c, oc, f, keys, dk, n, occ = _fill(cf, O, kw, cap, 300 + len(which))
    pre = f.exportTable()
    if which == "present":
        sel = np.arange(0, n // 2)
    elif which == "absent":
        sel = np.arange(n, n + n // 2)
    else:
        sel = np.concatenate([np.arange(0, n // 4), np.arange(n, n + n // 4)])
        np.random.default_rng(1).shuffle(sel)
    d_idx = torch.from_numpy(sel).to(dk.device)
    dsel = dk[d_idx].contiguous()
    out = torch.zeros(sel.size, dtype=torch.bool, device=dk.device)
    rem = f.deleteMany(dsel, out=out)
    st = check_delete_parity(O, oc, f.capacity(), pre, occ, keys[sel], out.cpu().numpy(), f.exportTable(), rem)
    print(f"{kw} {cap} {which} {path}: {st}", flush=True)
    if which == "present" and occ == n:
        assert st["successes"] == sel.size
    f.close()


@pytest.mark.parametrize("path", ["binned", "direct"])
def test_delete_repeated_keys_and_no_output(cf, oracle_mod, path, monkeypatch):
    """Every key deleted twice in one batch (once more than it was inserted),
    and the same batch without an output buffer."""
    import torch

    monkeypatch.setenv("CUCKOO_DELETE_PATH", path)
    O = oracle_mod
    c, oc, f, keys, dk, n, occ = _fill(cf, O, dict(eviction="bfs"), 1 << 20, 17)
    g = cf.CuckooFilter(1 << 20, config=c)
    g.copyFrom(f)
    pre = f.exportTable()
    twice = torch.cat([dk[: n // 4], dk[: n // 4]])
    out = torch.zeros(twice.numel(), dtype=torch.bool, device=dk.device)
    rem = f.deleteMany(twice, out=out)
    h = np.concatenate([keys[: n // 4], keys[: n // 4]])
    st = check_delete_parity(O, oc, f.capacity(), pre, occ, h, out.cpu().numpy(), f.exportTable(), rem)
    assert st["successes"] >= n // 4
    assert g.deleteMany(twice) == rem
    assert g.countOccupiedSlots() == f.countOccupiedSlots()
    f.close()
    g.close()


@pytest.mark.parametrize("m", [1, 2, 63, 4097, 100_003, 524_289])
def test_binned_delete_ragged_sizes(cf, oracle_mod, m, monkeypatch):
    import torch

    monkeypatch.setenv("CUCKOO_DELETE_PATH", "binned")
    O = oracle_mod
    c, oc, f, keys, dk, n, occ = _fill(cf, O, dict(eviction="bfs"), 1 << 20, 23)
    pre = f.exportTable()
    out = torch.zeros(m, dtype=torch.bool, device=dk.device)
    rem = f.deleteMany(dk[:m], out=out)
    check_delete_parity(O, oc, f.capacity(), pre, occ, keys[:m], out.cpu().numpy(), f.exportTable(), rem)
    assert out.all().item() and rem == occ - m
    f.close()
