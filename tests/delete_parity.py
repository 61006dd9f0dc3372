"""Parity of a concurrent deleteMany with the sequential oracle on one table image.

A delete removes one copy of its key's fingerprint: the first matching slot of
its first bucket that holds one, else of its second
(include/CuckooFilter.cuh:1106-1116; slot scan :805-848).  A delete never
moves a tag, so two deletes can change each other's outcome only when they
remove the same fingerprint from a bucket both keys may use.  Hence, for a
batch of deletes run in any order -- one after another, or concurrently with
each removal a single compare-and-swap, as on the GPU and in the reference:

  * a key that shares no (bucket, fingerprint) pair with another deleted key
    ("a lone key") has one outcome in every order: its result must equal the
    sequential oracle's, and every bucket that no shared pair touches must end
    byte-identical to the oracle's;
  * the keys that do share pairs fall into small groups (a few thousand keys
    of 2^26 at 2^28 slots).  A group's results and final buckets must be those
    of SOME order of its deletes, run one after another on the oracle.  The
    orders are searched depth-first, cutting a branch as soon as a delete's
    result differs from the one observed.

Whatever a linearizable batch of deletes can produce passes; anything else
(a wrong result, a wrong slot cleared, a tag removed twice) fails.  The
checker itself is tested on the CPU against the oracle's own multi-threaded
CAS deletes, and with planted mistakes (tests/test_delete_parity_cpu.py).
"""
from __future__ import annotations

import numpy as np

TAG_BYTES = {8: 1, 16: 2, 32: 4}


def _member(values: np.ndarray, sorted_set: np.ndarray) -> np.ndarray:
    if sorted_set.size == 0:
        return np.zeros(values.shape, bool)
    at = np.searchsorted(sorted_set, values)
    at = np.minimum(at, sorted_set.size - 1)
    return sorted_set[at] == values


def check_delete_parity(O, oc, capacity: int, pre_image: np.ndarray, occ_before: int, keys: np.ndarray,
                        got_ok: np.ndarray, got_image: np.ndarray, got_occ: int) -> dict:
    """Asserts that (got_ok, got_image, got_occ) is an outcome of deleting
    `keys` from `pre_image` in some order.  Returns counts for the log."""
    assert oc.policy != O.POLICY_OFFSET, "the Offset policy stores a different tag in each bucket"
    got_ok = np.asarray(got_ok, bool)
    assert got_occ == occ_before - int(got_ok.sum()), (got_occ, occ_before, int(got_ok.sum()))

    of = O.Filter(oc, capacity)
    nb = of.num_buckets
    bb = oc.bucket_size * TAG_BYTES[oc.bits_per_tag]
    of.load_table(pre_image, occ_before)
    occ_o, ok_o = of.delete_many(keys)
    assert occ_o == occ_before - int(ok_o.sum())
    want_rows = of.table_view().reshape(nb, bb)
    got_rows = np.ascontiguousarray(got_image).reshape(nb, bb)
    pre_rows = np.ascontiguousarray(pre_image).reshape(nb, bb)

    i1, i2, fp = O.candidates(oc, nb, keys)
    fp64 = fp.astype(np.uint64)
    ca = (i1 << np.uint64(32)) | fp64
    cb = (i2 << np.uint64(32)) | fp64
    two = i2 != i1
    pairs = np.sort(np.concatenate([ca, cb[two]]))
    shared = np.unique(pairs[1:][pairs[1:] == pairs[:-1]])
    del pairs
    in_a = _member(ca, shared)
    in_b = _member(cb, shared) & two
    grouped = in_a | in_b
    lone = ~grouped

    bad = np.nonzero(got_ok[lone] != ok_o[lone])[0]
    assert bad.size == 0, f"{bad.size} lone keys differ from the oracle, first at key {np.nonzero(lone)[0][bad[0]]}"

    gidx = np.nonzero(grouped)[0]
    group_buckets = np.unique(np.concatenate([i1[gidx], i2[gidx]]))
    differ = np.nonzero((got_rows != want_rows).any(axis=1))[0]
    outside = differ[~_member(differ.astype(np.uint64), group_buckets)]
    assert outside.size == 0, f"{outside.size} buckets differ outside the shared groups, first {outside[0]}"

    # groups: keys joined by shared (bucket, fingerprint) pairs.  Every key of
    # a group has the group's fingerprint, and no delete outside the group
    # removes that fingerprint from the group's buckets (it would share a
    # pair), so there only the group's deletes clear slots holding it
    parent = {int(k): int(k) for k in gidx}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    first = {}
    for k in gidx:
        for c, use in ((int(ca[k]), in_a[k]), (int(cb[k]), in_b[k])):
            if not use:
                continue
            if c in first:
                ra, rb = find(first[c]), find(int(k))
                if ra != rb:
                    parent[ra] = rb
            else:
                first[c] = int(k)
    groups = {}
    for k in gidx:
        groups.setdefault(find(int(k)), []).append(int(k))

    tag_dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[TAG_BYTES[oc.bits_per_tag]]
    pre_tags = pre_rows.view(tag_dt)
    got_tags = got_rows.view(tag_dt)
    want_tags = want_rows.view(tag_dt)
    # slots of the group buckets that hold no group's fingerprint are cleared
    # only by lone keys: they must match the oracle
    mine = np.zeros((group_buckets.size, oc.bucket_size), bool)
    gb_pos = {int(b): j for j, b in enumerate(group_buckets)}
    for members in groups.values():
        x = int(fp[members[0]])
        for b in {int(i1[k]) for k in members} | {int(i2[k]) for k in members}:
            mine[gb_pos[b]] |= pre_tags[b] == x
    gbi = group_buckets.astype(np.int64)
    other = ~mine & (got_tags[gbi] != want_tags[gbi])
    assert not other.any(), "a slot no shared group may clear differs from the oracle"

    sim = O.Filter(oc, capacity)
    sim_rows = sim.table_view().reshape(nb, bb)
    sim_tags = sim_rows.view(tag_dt)
    max_size = 0
    for members in groups.values():
        max_size = max(max_size, len(members))
        x = int(fp[members[0]])
        assert all(int(fp[k]) == x for k in members)
        gb = np.array(sorted({int(i1[k]) for k in members} | {int(i2[k]) for k in members}), np.int64)
        own = pre_tags[gb] == x
        seen = got_tags[gb][own]
        want_ok = {k: bool(got_ok[k]) for k in members}

        def search(remaining):
            if not remaining:
                return bool((sim_tags[gb][own] == seen).all())
            for k in list(remaining):
                state = sim_rows[gb].copy()
                _, r = sim.delete_many(keys[k:k + 1])
                if bool(r[0]) == want_ok[k]:
                    remaining.remove(k)
                    if search(remaining):
                        return True
                    remaining.add(k)
                sim_rows[gb] = state
            return False

        sim_rows[gb] = pre_rows[gb]
        assert search(set(members)), f"group {members} has no order giving the observed results and slots"
    return {"deletes": int(keys.size), "lone": int(lone.sum()), "grouped_keys": int(gidx.size),
            "groups": len(groups), "largest_group": max_size, "successes": int(got_ok.sum())}
