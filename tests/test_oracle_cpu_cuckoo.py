"""The single-threaded CPU cuckoo filter of oracle/cpu_cuckoo.h: the reported
CPU baseline that mirrors the reference's CPU comparison
(benchmark/cuckoo_cpu_vs_gpu.cu:52-155, cuckoofilter::CuckooFilter<uint64_t,16>).

Its upstream source is unavailable, so these tests pin the published design's
behaviour rather than bit-exact images: sizing, the partial-key alternate
index, no false negatives up to the victim cache, the false-positive rate
against the closed form 2*b*alpha/2^f with b = 4 and f = 16, and delete round
trips.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _expected_buckets(max_keys):
    nb = 1
    while nb < max(1, max_keys // 4):
        nb <<= 1
    if max_keys / nb / 4 > 0.96:
        nb <<= 1
    return nb


@pytest.mark.parametrize("cap", [1, 8, 1000, 4096, 1 << 20, 3 << 20])
def test_sizing(O, cap):
    f = O.CpuCuckoo(cap)
    assert f.num_buckets() == _expected_buckets(cap)
    assert f.size_bytes() == f.num_buckets() * 4 * 2


def test_alternate_index_is_an_involution(O):
    f = O.CpuCuckoo(1 << 16)
    nb = f.num_buckets()
    rng = np.random.default_rng(3)
    for i, t in zip(rng.integers(0, nb, 2000), rng.integers(1, 1 << 16, 2000)):
        a = f.alt_index(int(i), int(t))
        assert 0 <= a < nb
        assert f.alt_index(a, int(t)) == int(i)
    for k in O.gen_keys(9, 2000):
        i, t = f.index_tag(int(k))
        assert 0 <= i < nb and 1 <= t < (1 << 16)


def test_no_false_negatives_and_closed_form_fpr(O):
    cap = 1 << 20
    f = O.CpuCuckoo(cap)
    n = int(cap * 0.95)
    keys = O.gen_keys(42, n + (1 << 20))
    assert f.add_many(keys[:n]) == n
    assert f.items() == n and not f.victim_used()
    assert f.contain_many(keys[:n]) == n
    absent = keys[n:]
    fp = f.contain_many(absent)
    alpha = n / (f.num_buckets() * 4)
    p = 2 * 4 * alpha / 2**16
    mean, sd = p * absent.size, np.sqrt(p * (1 - p) * absent.size)
    assert abs(fp - mean) <= 4 * sd + 2, (fp, mean)


def test_fill_until_refused_keeps_every_accepted_key(O):
    f = O.CpuCuckoo(4096)
    keys = O.gen_keys(5, 3 * f.num_buckets() * 4)
    ok, accepted = f.add_many(keys, want_out=True)
    assert f.victim_used()  # the table filled: one tag waits in the victim cache
    assert not accepted.all()
    first_refused = int(np.argmin(accepted))
    assert accepted[:first_refused].all() and not accepted[first_refused:].any()
    assert ok == first_refused
    assert f.items() == ok - 1  # all but the victim sit in buckets
    assert f.contain_many(keys[:ok]) == ok
    assert ok / (f.num_buckets() * 4) > 0.9  # four-slot buckets fill past 90%


def test_delete_round_trip_reinserts_the_victim(O):
    f = O.CpuCuckoo(4096)
    keys = O.gen_keys(6, 3 * f.num_buckets() * 4)
    ok = f.add_many(keys)
    assert f.victim_used()
    gone, out = f.delete_many(keys[:100], want_out=True)
    assert gone == 100 and out.all()
    assert not f.victim_used()  # the first delete made room for it
    assert f.contain_many(keys[100:ok]) == ok - 100
    assert f.delete_many(keys[100:ok]) == ok - 100
    assert f.items() == 0 and not f.victim_used()
    assert f.contain_many(keys[:ok]) == 0  # an empty table matches nothing


def test_duplicates_take_separate_slots(O):
    f = O.CpuCuckoo(1 << 10)
    k = O.gen_keys(8, 1)
    assert f.add_many(np.repeat(k, 3)) == 3 and f.items() == 3
    assert f.delete_many(k) == 1 and f.contain_many(k) == 1
    assert f.delete_many(np.repeat(k, 2)) == 2
    assert f.contain_many(k) == 0 and f.delete_many(k) == 0


def test_same_keys_same_table(O):
    keys = O.gen_keys(11, 50_000)
    a, b = O.CpuCuckoo(1 << 16), O.CpuCuckoo(1 << 16)
    a.add_many(keys)
    b.add_many(keys)
    assert a.checksum() == b.checksum()
    c = O.CpuCuckoo(1 << 16)
    c.add_many(keys[::-1])
    assert c.contain_many(keys) == keys.size
