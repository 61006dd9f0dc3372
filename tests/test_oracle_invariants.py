"""Pins the oracle's bucket logic by brute force, a second slot-level model and
invariants the algorithm fixes (CuckooFilter.cuh:805-1116).

* SWAR (helpers.cuh:78-151): exhaustive over lane patterns -- the lowest set
  bit of getZeroMask marks the lowest zero lane, for every tag/word width.
* Sequential insert / delete with BFS and DFS eviction: table images equal a
  slot-list model written independently below, on tiny tables driven past
  capacity so both eviction paths and chain exhaustion run.
* Lookup: brute force over both candidate buckets of hand-built images (full
  buckets, duplicate tags, tag 1, i1 == i2).
* At the 2^20-slot workload: occupancy equals the non-empty slot count,
  stored tags are exactly the keys' fingerprints, every stored tag sits in one
  of its key's buckets, no false negatives, FPR within binomial bounds of
  2*b*alpha/2^f (src/main.cu:114-115).
"""
import itertools

import numpy as np
import pytest

from conftest import stray_slack


def _cfg(O, **kw):
    return O.make_config(**kw)


# ---------------------------------------------------------------------------
# SWAR
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("tag_bits,word_bits", [(8, 32), (8, 64), (16, 32), (16, 64), (32, 32), (32, 64)])
def test_swar_lowest_bit_is_lowest_matching_lane(oracle_mod, tag_bits, word_bits):
    L = oracle_mod.lib()
    lanes = word_bits // tag_bits
    lane_mask = (1 << tag_bits) - 1
    rng = np.random.default_rng(tag_bits * 100 + word_bits)
    for tag in (1, lane_mask, 1 << (tag_bits - 1), int(rng.integers(2, lane_mask))):
        other = (tag ^ 0x5A5A5A5A) & lane_mask or 2
        near = tag ^ 1 if tag ^ 1 else 3
        values = (0, tag, other, near)
        rep = L.cfo_swar_replicate(tag, tag_bits, word_bits)
        assert rep == sum(tag << (tag_bits * i) for i in range(lanes))
        for pattern in itertools.product(range(4), repeat=min(lanes, 6)):
            pat = list(pattern) + [int(rng.integers(0, 4)) for _ in range(lanes - len(pattern))]
            word = sum(values[p] << (tag_bits * i) for i, p in enumerate(pat))
            zm = L.cfo_swar_zero_mask(word, tag_bits, word_bits)
            zero_lanes = [i for i, p in enumerate(pat) if values[p] == 0]
            if zero_lanes:
                assert zm != 0 and ((zm & -zm).bit_length() - 1) // tag_bits == zero_lanes[0]
            else:
                assert zm == 0
            mm = L.cfo_swar_zero_mask(word ^ rep, tag_bits, word_bits)
            match_lanes = [i for i, p in enumerate(pat) if values[p] == tag]
            if match_lanes:
                assert mm != 0 and ((mm & -mm).bit_length() - 1) // tag_bits == match_lanes[0]
            else:
                assert mm == 0


# ---------------------------------------------------------------------------
# independent slot-list model of the sequential algorithm
# ---------------------------------------------------------------------------

class SlotModel:
    """Buckets as lists of slots; slot s lives in word s // tpw, lane s % tpw.

    Follows CuckooFilter.cuh:805-1080 step by step, written over slots rather
    than packed words, using the oracle only for hashing (pinned elsewhere).
    """

    def __init__(self, O, cfg, nb):
        self.O, self.cfg, self.nb = O, cfg, nb
        tb = {8: 1, 16: 2, 32: 4}[cfg.bits_per_tag]
        self.tpw = cfg.word_bytes // tb
        self.wc = cfg.bucket_size // self.tpw
        self.bs = cfg.bucket_size
        self.t = [[0] * self.bs for _ in range(nb)]

    def alt(self, b, fp):
        return self.O.alt_bucket(self.cfg, self.nb, b, fp)

    def _scan(self, b, tag, want):
        start = (tag & (self.bs - 1)) // self.tpw
        for i in range(self.wc):
            w = (start + i) % self.wc
            for lane in range(self.tpw):
                s = w * self.tpw + lane
                if self.t[b][s] == want:
                    return s
        return None

    def try_insert(self, b, tag):
        s = self._scan(b, tag, 0)
        if s is None:
            return False
        self.t[b][s] = tag
        return True

    def try_remove(self, b, tag):
        s = self._scan(b, tag, tag)
        if s is None:
            return False
        self.t[b][s] = 0
        return True

    def dfs(self, fp, b):
        for e in range(self.cfg.max_evictions):
            s = (fp + e * 7) & (self.bs - 1)
            fp, self.t[b][s] = self.t[b][s], fp
            b, fp = self.alt(b, fp)
            if self.try_insert(b, fp):
                return True
        return False

    def bfs_candidates(self, fp, b):
        for i in range(min(8, self.bs // 2)):
            s = (fp + i * 7) & (self.bs - 1)
            cand = self.t[b][s]
            if cand == 0:
                if self.try_insert(b, fp):
                    return True
                continue
            ab, afp = self.alt(b, cand)
            if self.try_insert(ab, afp):
                self.t[b][s] = fp
                return True
        return False

    def bfs(self, fp, b):
        return self.bfs_candidates(fp, b) or self.dfs(fp, b)

    def insert(self, i1, fp):
        i2, fp2 = self.alt(i1, fp)
        if self.try_insert(i1, fp) or self.try_insert(i2, fp2):
            return True
        first = (fp & 1) == 0
        f, b = (fp, i1) if first else (fp2, i2)
        return self.bfs(f, b) if self.cfg.eviction == 0 else self.dfs(f, b)

    def remove(self, i1, fp):
        i2, fp2 = self.alt(i1, fp)
        return self.try_remove(i1, fp) or self.try_remove(i2, fp2)

    def image_tags(self):
        return np.array(self.t, dtype=np.uint64)


TINY = [
    # (bits, bucket, word_bytes, policy, eviction, max_evictions, nb-capacity)
    (16, 4, 8, 0, 0, 30, 256),
    (16, 4, 8, 0, 1, 30, 256),
    (16, 8, 8, 0, 0, 500, 512),
    (8, 16, 4, 0, 1, 40, 1024),
    (32, 4, 8, 0, 0, 25, 256),
    (8, 32, 8, 0, 0, 500, 2048),
    (16, 16, 4, 1, 0, 50, 1000),
    (16, 16, 8, 2, 1, 50, 1000),
    (8, 8, 8, 2, 0, 30, 600),
    (32, 8, 4, 1, 1, 20, 500),
]


@pytest.mark.parametrize("bits,bucket,wb,pol,ev,maxev,cap", TINY)
def test_sequential_insert_delete_matches_slot_model(oracle_mod, bits, bucket, wb, pol, ev, maxev, cap):
    O = oracle_mod
    cfg = O.make_config(key_bytes=8, bits_per_tag=bits, bucket_size=bucket, word_bytes=wb, policy=pol,
                        eviction=ev, max_evictions=maxev)
    f = O.Filter(cfg, cap)
    nb = f.num_buckets
    model = SlotModel(O, cfg, nb)
    n = int(f.capacity() * 1.08)  # past capacity: evictions and exhausted chains
    keys = O.gen_keys(1000 + bits + bucket, n)
    i1, _, fp = O.candidates(cfg, nb, keys)
    occ, ok = f.insert_many(keys)
    want_ok = np.array([model.insert(int(a), int(b)) for a, b in zip(i1, fp)])
    np.testing.assert_array_equal(ok, want_ok)
    np.testing.assert_array_equal(O.table_tags(f.table_bytes(), cfg).astype(np.uint64), model.image_tags())
    assert occ == int(want_ok.sum()) == f.count_occupied_slots()
    assert (~ok).sum() > 0, "workload should exhaust some eviction chains"
    # delete every third key
    dk = keys[::3]
    occ2, dok = f.delete_many(dk)
    want_d = np.array([model.remove(int(a), int(b)) for a, b in zip(i1[::3], fp[::3])])
    np.testing.assert_array_equal(dok, want_d)
    np.testing.assert_array_equal(O.table_tags(f.table_bytes(), cfg).astype(np.uint64), model.image_tags())
    assert occ2 == occ - int(want_d.sum()) == f.count_occupied_slots()


@pytest.mark.parametrize("bits,bucket,wb,pol,ev,maxev,cap", TINY)
def test_phased_insert_matches_slot_model(oracle_mod, bits, bucket, wb, pol, ev, maxev, cap):
    """insert_many_phased: the same steps per key, scheduled by step over the
    batch -- all primary attempts, then secondaries, then BFS candidate loops,
    then DFS chains, each in key order."""
    O = oracle_mod
    cfg = O.make_config(key_bytes=8, bits_per_tag=bits, bucket_size=bucket, word_bytes=wb, policy=pol,
                        eviction=ev, max_evictions=maxev)
    f = O.Filter(cfg, cap)
    nb = f.num_buckets
    m = SlotModel(O, cfg, nb)
    n = int(f.capacity() * 1.08)
    keys = O.gen_keys(2000 + bits + bucket, n)
    i1, _, fp = O.candidates(cfg, nb, keys)
    occ, ok, _ = f.insert_many_phased(keys)
    done = [m.try_insert(int(a), int(b)) for a, b in zip(i1, fp)]
    start = []
    for j in range(n):
        b, t = int(i1[j]), int(fp[j])
        b2, t2 = m.alt(b, t)
        if not done[j]:
            done[j] = m.try_insert(b2, t2)
        start.append((t, b) if (t & 1) == 0 else (t2, b2))
    if ev == O.EVICT_BFS:
        for j in range(n):
            if not done[j]:
                done[j] = m.bfs_candidates(*start[j])
    for j in range(n):
        if not done[j]:
            done[j] = m.dfs(*start[j])
    np.testing.assert_array_equal(ok, np.array(done))
    np.testing.assert_array_equal(O.table_tags(f.table_bytes(), cfg).astype(np.uint64), m.image_tags())
    assert occ == int(ok.sum()) == f.count_occupied_slots()
    assert (~ok).sum() > 0


def test_sorted_insert_is_sequential_insert_in_packed_order(oracle_mod):
    O = oracle_mod
    cfg = O.make_config(eviction=O.EVICT_BFS)
    f = O.Filter(cfg, 4096)
    keys = O.gen_keys(5, 3900)
    f.insert_many_sorted(keys)
    order = np.argsort(O.packed_tags(cfg, f.num_buckets, keys), kind="stable")
    g = O.Filter(cfg, 4096)
    g.insert_many(keys[order])
    np.testing.assert_array_equal(f.table_bytes(), g.table_bytes())


def test_sorted_insert_truncates_packed_bucket_for_8bit_tags(oracle_mod):
    """PackedTagType is u32 for 8-bit tags (CuckooFilter.cuh:159): primary bucket
    index bits above 24 are lost, exactly as the reference's sorted path loses them."""
    O = oracle_mod
    cfg = O.make_config(bits_per_tag=8, word_bytes=8, bucket_size=8)
    nb = 1 << 26
    keys = O.gen_keys(9, 2000)
    i1, _, fp = O.candidates(cfg, nb, keys)
    packed = O.packed_tags(cfg, nb, keys)
    np.testing.assert_array_equal(packed >> np.uint64(8), i1 & np.uint64((1 << 24) - 1))
    np.testing.assert_array_equal(packed & np.uint64(0xFF), fp.astype(np.uint64))


# ---------------------------------------------------------------------------
# lookup brute force on hand-built images
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("bits,bucket,wb,pol", [(16, 16, 8, 0), (8, 8, 4, 0), (32, 4, 8, 0), (16, 4, 4, 2),
                                                (8, 32, 8, 1), (16, 32, 4, 0)])
def test_contains_brute_force_on_hand_built_images(oracle_mod, bits, bucket, wb, pol):
    O = oracle_mod
    cfg = O.make_config(bits_per_tag=bits, bucket_size=bucket, word_bytes=wb, policy=pol)
    f = O.Filter(cfg, 64 * bucket)
    nb = f.num_buckets
    rng = np.random.default_rng(bits + bucket + pol)
    keys = O.gen_keys(77, 4000)
    i1, i2, fp = O.candidates(cfg, nb, keys)
    fp2 = np.array([O.alt_bucket(cfg, nb, int(a), int(b))[1] for a, b in zip(i1, fp)], dtype=np.uint32)
    tags = np.zeros((nb, bucket), dtype=np.uint64)
    tmax = (1 << bits) - 1
    # random fill with a mix of: real fingerprints, tag value 1, duplicates, full buckets
    for b in range(nb):
        kind = b % 4
        for s in range(bucket):
            r = rng.random()
            if kind == 0 and r < 0.5:
                continue
            if r < 0.3:
                tags[b, s] = int(fp[rng.integers(0, fp.size)])
            elif r < 0.4:
                tags[b, s] = 1
            else:
                tags[b, s] = int(rng.integers(1, tmax + 1))
        if kind == 2:
            tags[b, :] = tags[b, 0] or 1  # bucket full of one duplicated tag
    # plant some keys' own tags in one of their two buckets (true positives)
    for j in rng.choice(keys.size, 300, replace=False):
        if rng.random() < 0.5:
            tags[i1[j], rng.integers(0, bucket)] = fp[j]
        else:
            tags[i2[j], rng.integers(0, bucket)] = fp2[j]
    dt = {8: np.uint8, 16: np.uint16, 32: np.uint32}[bits]
    image = tags.astype(dt).view(np.uint8).reshape(-1)
    f.load_table(image, int((tags != 0).sum()))
    got = f.contains_many(keys)
    want = np.array([(tags[a] == b).any() or (tags[c] == d).any() for a, b, c, d in zip(i1, fp, i2, fp2)])
    np.testing.assert_array_equal(got, want)
    assert want.any() and (~want).any()


def test_contains_when_both_candidates_coincide(oracle_mod):
    """nb = 1 forces i1 == i2 == 0 for every key."""
    O = oracle_mod
    cfg = O.make_config()
    f = O.Filter(cfg, 16)
    assert f.num_buckets == 1
    keys = O.gen_keys(3, 16)
    occ, ok = f.insert_many(keys)
    assert occ == 16 and ok.all()
    assert f.contains_many(keys).all()
    occ, ok = f.insert_many(O.gen_keys(4, 4))
    assert occ == 16 and not ok.any()


# ---------------------------------------------------------------------------
# workload-scale invariants (2^20 slots, alpha = 0.95)
# ---------------------------------------------------------------------------

def _stored(O, f, cfg):
    t = O.table_tags(f.table_bytes(), cfg)
    b, s = np.nonzero(t)
    return b.astype(np.uint64), t[b, s].astype(np.uint64)


@pytest.mark.parametrize("eviction", [0, 1])
def test_workload_invariants(oracle_mod, eviction):
    O = oracle_mod
    cfg = O.make_config(eviction=eviction)
    f = O.Filter(cfg, 1 << 20)
    n = int((1 << 20) * 0.95)
    keys = O.gen_keys(42, n + (1 << 19))
    pos, neg = keys[:n], keys[n:]
    occ, ok = f.insert_many(pos)
    assert ok.all() and occ == n == f.count_occupied_slots()
    nb = f.num_buckets
    i1, i2, fp = O.candidates(cfg, nb, pos)
    b, tag = _stored(O, f, cfg)
    # stored tags are exactly the inserted fingerprints (as multisets)
    np.testing.assert_array_equal(np.sort(tag), np.sort(fp.astype(np.uint64)))
    # every stored (bucket, tag) is one of some key's (i1, fp) / (i2, fp)
    pairs = np.union1d((i1 << np.uint64(32)) | fp, (i2 << np.uint64(32)) | fp)
    assert np.isin((b << np.uint64(32)) | tag, pairs).all()
    assert f.contains_many(pos).all()
    # FPR against the closed form 2*b*alpha/2^f, 99.9% binomial band
    fps = int(f.contains_many(neg).sum())
    p = 2 * 16 * 0.95 / 65536
    mu, sd = neg.size * p, np.sqrt(neg.size * p * (1 - p))
    assert abs(fps - mu) < 3.3 * sd, (fps, mu, sd)


def test_delete_half_then_lookup(oracle_mod):
    O = oracle_mod
    cfg = O.make_config(eviction=O.EVICT_BFS)
    f = O.Filter(cfg, 1 << 20)
    n = int((1 << 20) * 0.95)
    keys = O.gen_keys(42, n)
    f.insert_many(keys)
    occ, ok = f.delete_many(keys[: n // 2])
    assert ok.all() and occ == n - n // 2 == f.count_occupied_slots()
    assert f.contains_many(keys[n // 2:]).all()


def test_exhaustion_keeps_count_consistent(oracle_mod):
    """Past capacity DFS drops one victim per failed insert (CuckooFilter.cuh:917-954)."""
    O = oracle_mod
    cfg = O.make_config(max_evictions=50)
    f = O.Filter(cfg, 1 << 12)
    keys = O.gen_keys(8, int(f.capacity() * 1.2))
    occ, ok = f.insert_many(keys)
    fails = int((~ok).sum())
    assert fails > 0 and occ == keys.size - fails == f.count_occupied_slots() <= f.capacity()
    _, tag = _stored(O, f, cfg)
    _, _, fp = O.candidates(cfg, f.num_buckets, keys)
    # stored multiset is the keys' fingerprints minus exactly `fails` of them
    vals, cnt_keys = np.unique(fp.astype(np.uint64), return_counts=True)
    sv, cnt_st = np.unique(tag, return_counts=True)
    idx = np.searchsorted(vals, sv)
    assert (vals[idx] == sv).all() and (cnt_st <= cnt_keys[idx]).all()
    assert cnt_keys.sum() - cnt_st.sum() == fails


def test_multithreaded_cas_variant_is_consistent(oracle_mod):
    O = oracle_mod
    cfg = O.make_config(eviction=O.EVICT_BFS)
    f = O.Filter(cfg, 1 << 20)
    keys = O.gen_keys(42, int((1 << 20) * 0.95))
    occ, ok = f.insert_many(keys, threads=4)
    assert occ == int(ok.sum())
    # concurrent BFS may leave a few uncounted duplicate tags (see stray_slack)
    assert occ <= f.count_occupied_slots() <= occ + stray_slack(keys.size, "bfs")
    assert f.contains_many(keys[ok], threads=4).all()
