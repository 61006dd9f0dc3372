"""The reference's behavioural tests (tests/test_cuckoo_filter.cu) on the HIP
library through the C ABI, with the same fixture configuration:
CuckooConfig<uint32_t, 16, 500, 256, 16> (uint32 keys, DFS, uint64 words)."""
import numpy as np
import pytest

import reference_behaviour as RB

pytestmark = pytest.mark.gpu


class GpuAdapter:
    def __init__(self, cf):
        self.cf = cf
        self.cfg = cf.make_config(key_bytes=4, bits_per_tag=16, max_evictions=500, block_size=256, bucket_size=16)

    def new(self, capacity):
        return self.cf.CuckooFilter(capacity, config=self.cfg)

    def _k(self, keys):
        return self.cf.keys_to_device(np.ascontiguousarray(keys, dtype=np.uint32))

    def insert(self, f, keys):
        return f.insertMany(self._k(keys))

    def contains(self, f, keys):
        return f.containsMany(self._k(keys)).cpu().numpy()

    def delete(self, f, keys):
        import torch

        k = self._k(keys)
        out = torch.zeros(k.numel(), dtype=torch.uint8, device=k.device)
        occ = f.deleteMany(k, out=out)
        return occ, out.cpu().numpy().astype(bool)

    def load_factor(self, f):
        return f.loadFactor()

    def capacity(self, f):
        return f.capacity()

    def clear(self, f):
        f.clear()


@pytest.mark.parametrize("case", RB.ALL, ids=[c.__name__ for c in RB.ALL])
def test_reference_behaviour_on_gpu(cf, case):
    case(GpuAdapter(cf))


def test_sorted_insert_reference_behaviour(cf):
    """insertManySorted satisfies the same insert/lookup expectations."""
    a = GpuAdapter(cf)
    n = int((1 << 20) * RB.TARGET_LOAD_FACTOR)
    f = a.new(n)
    keys = RB.random_keys(n)
    inserted = f.insertManySorted(a._k(keys))
    assert inserted > n * 0.99
    assert a.contains(f, keys).sum() > inserted * 0.99
