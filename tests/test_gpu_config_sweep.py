"""Every configuration the C ABI builds, under both eviction policies, filled
to alpha = 0.98: the parameter sweep's points as parity cases.

Per configuration (key width x policy x tag bits x bucket size x word size,
138 of them, each with BFS and DFS), a 2^20-slot filter is offered
0.98 x capacity distinct keys through insertMany, then checked with
size-independent properties:
  * the counter never exceeds the keys offered, and the non-empty slot count
    equals it up to BFS strays;
  * keys that went missing are at most the failed inserts (an exhausted
    eviction chain drops one victim per failure); under Offset + BFS the
    reference's clean-up (include/CuckooFilter.cuh:1011-1012) may also remove
    another key's equal tag after a lost swap, which is allowed up to the
    stray count (each lost swap leaves a stray or, far more rarely, removes
    an equal tag);
  * the false-positive count of 2^18 absent keys is within a binomial band of
    1 - (1 - 1/(2^f - 1))^(2*b*alpha) (the filter's own occupancy as alpha),
    for tags of 8 and 16 bits (32-bit tags give ~0 expected);
  * when every insert succeeded, deleting every key succeeds and empties the
    counter.
"""
import functools
import itertools

import numpy as np
import pytest

from conftest import assert_table_count

pytestmark = pytest.mark.gpu

CONFIGS = []
for kb, pol, bits, bucket, wb in itertools.product((4, 8), ("xor", "addsub", "offset"), (8, 16, 32),
                                                   (4, 8, 16, 32), (4, 8)):
    tb = {8: 1, 16: 2, 32: 4}[bits]
    if tb > wb or bucket % (wb // tb):
        continue
    CONFIGS.append(dict(key_bytes=kb, policy=pol, bits_per_tag=bits, bucket_size=bucket, word_bytes=wb))
assert len(CONFIGS) == 138

CAP = 1 << 20
ALPHA = 0.98
N_NEG = 1 << 18


@functools.lru_cache(maxsize=None)
def _key_sets(O, kb, n):
    """n distinct positives and N_NEG distinct negatives disjoint from them."""
    if kb == 8:
        k = O.gen_keys(11, n + N_NEG)  # one stream never repeats a key
        return k[:n], k[n:]
    k = np.unique(O.gen_keys(11, 2 * (n + N_NEG)).astype(np.uint32))
    np.random.default_rng(11).shuffle(k)
    assert k.size >= n + N_NEG
    return k[:n], k[n:n + N_NEG]


@pytest.mark.parametrize("eviction", ["bfs", "dfs"])
@pytest.mark.parametrize("kw", CONFIGS,
                         ids=lambda d: "k{key_bytes}-{policy}-t{bits_per_tag}-b{bucket_size}-w{word_bytes}".format(**d))
def test_config_at_98_percent(cf, oracle_mod, kw, eviction):
    import torch

    f = cf.CuckooFilter(CAP, config=cf.make_config(eviction=eviction, **kw))
    n = int(f.capacity() * ALPHA)
    pos, neg = _key_sets(oracle_mod, kw["key_bytes"], n)
    dpos, dneg = cf.keys_to_device(pos), cf.keys_to_device(neg)
    occ = f.insertMany(dpos)
    assert occ <= n
    table = assert_table_count(f, occ, n)
    missing = int((~f.containsMany(dpos)).sum().item())
    quirk = table - occ if (kw["policy"], eviction) == ("offset", "bfs") else 0
    assert missing <= n - occ + quirk, (missing, n - occ, quirk)
    if kw["bits_per_tag"] <= 16:
        alpha = occ / f.capacity()
        p = 1.0 - (1.0 - 1.0 / (2 ** kw["bits_per_tag"] - 1)) ** (2 * kw["bucket_size"] * alpha)
        fps = int(f.containsMany(dneg).sum().item())
        mu, sd = N_NEG * p, np.sqrt(N_NEG * p * (1 - p)) + 1
        assert abs(fps - mu) < 5 * sd + 0.03 * mu, (fps, mu)
    if occ == n and missing == 0:
        out = torch.zeros(n, dtype=torch.bool, device=dpos.device)
        assert f.deleteMany(dpos, out=out) == 0
        assert out.all().item()
    f.close()
