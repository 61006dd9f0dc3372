"""The reference's behavioural tests (tests/test_cuckoo_filter.cu) on the CPU oracle."""
import numpy as np
import pytest

import reference_behaviour as RB


class OracleAdapter:
    def __init__(self, O):
        self.O = O
        self.cfg = O.make_config(key_bytes=4, bits_per_tag=16, max_evictions=500, block_size=256, bucket_size=16)

    def new(self, capacity):
        return self.O.Filter(self.cfg, capacity)

    def insert(self, f, keys):
        return f.insert_many(keys)[0]

    def contains(self, f, keys):
        return f.contains_many(keys)

    def delete(self, f, keys):
        return f.delete_many(keys)

    def load_factor(self, f):
        return float(np.float32(f.occupied()) / np.float32(f.capacity()))

    def capacity(self, f):
        return f.capacity()

    def clear(self, f):
        f.clear()


@pytest.mark.parametrize("case", RB.ALL, ids=[c.__name__ for c in RB.ALL])
def test_reference_behaviour_on_oracle(oracle_mod, case):
    case(OracleAdapter(oracle_mod))
