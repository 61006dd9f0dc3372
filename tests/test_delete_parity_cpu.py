"""The delete-parity checker (tests/delete_parity.py) on the CPU.

It must accept every outcome some order of the deletes can give -- the
sequential oracle's, and the oracle's own lock-free multi-threaded CAS run,
whose order is the threads' -- and reject planted mistakes: a lone key's
result flipped, a slot cleared that no delete clears, a group's results
changed to ones no order gives, the wrong copy of a shared tag cleared.
"""
import numpy as np
import pytest

from delete_parity import TAG_BYTES, check_delete_parity


def _setup(O, bits, cap, load, frac, seed):
    oc = O.make_config(bits_per_tag=bits, eviction=O.EVICT_BFS, word_bytes=8 if bits > 8 else 4)
    of = O.Filter(oc, cap)
    n = int(of.capacity() * load)
    keys = O.gen_keys(seed, n)
    occ, ok = of.insert_many(keys)
    assert ok.all()
    pre = of.table_bytes()
    dels = keys[: int(n * frac)]
    return oc, of, pre, occ, dels


def _run(O, oc, cap, pre, occ, dels, threads):
    f = O.Filter(oc, cap)
    f.load_table(pre, occ)
    occ2, ok = f.delete_many(dels, threads=threads)
    return ok, f.table_bytes(), occ2


@pytest.mark.parametrize("bits,cap,frac", [(16, 1 << 22, 0.25), (8, 1 << 16, 0.5)])
@pytest.mark.parametrize("threads", [1, 8])
def test_checker_accepts_sequential_and_concurrent_deletes(oracle_mod, bits, cap, frac, threads):
    O = oracle_mod
    oc, _, pre, occ, dels = _setup(O, bits, cap, 0.95, frac, 5)
    ok, img, occ2 = _run(O, oc, cap, pre, occ, dels, threads)
    st = check_delete_parity(O, oc, cap, pre, occ, dels, ok, img, occ2)
    assert st["groups"] > 0 and st["lone"] > 0, st


def _grouped_case(O):
    """8-bit tags in a small table: many deletes share a (bucket, tag) pair."""
    cap = 1 << 14
    oc, _, pre, occ, dels = _setup(O, 8, cap, 0.95, 0.5, 9)
    ok, img, occ2 = _run(O, oc, cap, pre, occ, dels, 1)
    return oc, cap, pre, occ, dels, ok, img, occ2


def _grouping(O, oc, cap, dels):
    nb = O.num_buckets(oc, cap)
    i1, i2, fp = O.candidates(oc, nb, dels)
    ca = (i1 << np.uint64(32)) | fp
    cb = (i2 << np.uint64(32)) | fp
    u, c = np.unique(np.concatenate([ca, cb[i1 != i2]]), return_counts=True)
    shared = u[c > 1]
    grouped = np.isin(ca, shared) | (np.isin(cb, shared) & (i1 != i2))
    return nb, i1, i2, fp, grouped


def test_checker_rejects_a_flipped_lone_result(oracle_mod):
    O = oracle_mod
    oc, cap, pre, occ, dels, ok, img, occ2 = _grouped_case(O)
    _, _, _, _, grouped = _grouping(O, oc, cap, dels)
    k = int(np.nonzero(~grouped)[0][0])
    bad = ok.copy()
    bad[k] = not bad[k]
    with pytest.raises(AssertionError, match="lone keys differ"):
        check_delete_parity(O, oc, cap, pre, occ, dels, bad, img, occ - int(bad.sum()))


def test_checker_rejects_a_stray_cleared_slot(oracle_mod):
    O = oracle_mod
    oc, cap, pre, occ, dels, ok, img, occ2 = _grouped_case(O)
    bb = oc.bucket_size * TAG_BYTES[oc.bits_per_tag]
    nb, i1, i2, _, grouped = _grouping(O, oc, cap, dels)
    near_groups = set(i1[grouped].tolist()) | set(i2[grouped].tolist())
    rows = img.reshape(nb, bb).copy()
    b = next(b for b in range(nb) if b not in near_groups and rows[b].any())
    rows[b][np.nonzero(rows[b])[0][0]] = 0
    with pytest.raises(AssertionError, match="buckets differ outside"):
        check_delete_parity(O, oc, cap, pre, occ, dels, ok, rows.reshape(-1), occ2)


def test_checker_rejects_impossible_group_outcomes(oracle_mod):
    O = oracle_mod
    oc, cap, pre, occ, dels, ok, img, occ2 = _grouped_case(O)
    bb = oc.bucket_size * TAG_BYTES[oc.bits_per_tag]
    nb, i1, i2, fp, grouped = _grouping(O, oc, cap, dels)
    grouped = np.nonzero(grouped)[0]
    assert grouped.size > 1
    # (a) a grouped key's result flipped, the image left as it was
    k = int(grouped[0])
    bad = ok.copy()
    bad[k] = not bad[k]
    with pytest.raises(AssertionError, match="no order"):
        check_delete_parity(O, oc, cap, pre, occ, dels, bad, img, occ - int(bad.sum()))
    # (b) the observed results kept, but an extra copy of a shared tag cleared
    rows = img.reshape(nb, bb).copy()
    tag_dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[TAG_BYTES[oc.bits_per_tag]]
    for k in grouped:
        b = int(i1[k])
        tags = rows[b].view(tag_dt)
        hit = np.nonzero(tags == fp[k])[0]
        if hit.size:
            tags[hit[0]] = 0
            break
    else:
        pytest.skip("no surviving copy of a shared tag in this image")
    with pytest.raises(AssertionError, match="no order"):
        check_delete_parity(O, oc, cap, pre, occ, dels, ok, rows.reshape(-1), occ2)
