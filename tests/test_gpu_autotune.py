"""The autotuner's measurements on the GPU (small shapes).

Each measured point must agree with what the configuration fixes: no failed
insert at 90% load with bucket sizes of 8 and up, a measured FPR within a
binomial band of the closed form at the measured load, and a primary-bucket
share between one half and one (BFS and DFS both settle most tags in their
primary bucket below full load).  choose() must return one of the measured
points that meets the budget.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_measure_and_choose(cf):
    import autotune

    points = [dict(key_bytes=8, policy="xor", bits_per_tag=8, bucket_size=16, word_bytes=8, eviction="bfs",
                   block_size=256),
              dict(key_bytes=8, policy="xor", bits_per_tag=16, bucket_size=8, word_bytes=8, eviction="dfs",
                   block_size=128),
              dict(key_bytes=8, policy="xor", bits_per_tag=16, bucket_size=16, word_bytes=4, eviction="bfs",
                   block_size=512)]
    res = [autotune.measure(kw, 1 << 18, 0.9) for kw in points]
    for r in res:
        assert r["failed"] == 0 and r["exhausted"] == 0
        assert r["insert_mkeys_s"] > 0 and r["lookup_mkeys_s"] > 0 and r["delete_mkeys_s"] > 0
        m = r["keys"] - r["keys"] // 2
        p = r["fpr_closed_form"]
        assert abs(r["fpr"] * m - p * m) < 5 * np.sqrt(m * p * (1 - p)) + 3, (r["fpr"], p)
        assert 0.5 < r["primary_share"] <= 1.0
        assert r["lookup_bucket_gbs"] > 0
    budget = autotune.fpr_closed_form(16, 16, 0.9)
    best = autotune.choose(res, budget)
    assert best is not None and best["bits_per_tag"] == 16  # the 8-bit point exceeds the budget
    assert best in res


def test_batched_trial_and_choice(cf):
    import autotune

    kw = dict(key_bytes=8, policy="xor", bits_per_tag=16, bucket_size=16, word_bytes=8, eviction="bfs",
              block_size=256)
    r = autotune.batched_trial(kw, n_filters=8, n_batches=64, seed=3, cap_log2=(12, 16))
    assert r["occupancies_match_plain"] and r["batched_mkeys_s"] > 0 and r["plain_mkeys_s"] > 0
    cands = [kw, dict(kw, word_bytes=4, eviction="dfs"), dict(kw, bits_per_tag=8)]
    budget = autotune.fpr_closed_form(16, 16, 0.95)
    best, results, confirmed = autotune.choose_batched(budget, n_filters=8, n_batches=64, cands=cands,
                                                       shortlist=1, confirm=(8, 128), repeats=2)
    assert len(results) == 2  # the 8-bit candidate exceeds the budget and is not run
    assert best in cands[:2]
    assert len(confirmed) == 1 and confirmed[0][0] == best
    # with a baseline and an unreachable margin, the baseline stays
    best, _, confirmed = autotune.choose_batched(budget, n_filters=8, n_batches=64, cands=cands, baseline=cands[1],
                                                 shortlist=2, confirm=(8, 128), repeats=1, margin=10.0)
    assert best == cands[1] and len(confirmed) == 2
