"""Host logic of the autotuner (no GPU): the FPR closed form, the candidate
set, the op-mix score and the admissibility rules of choose()."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def at():
    import autotune

    return autotune


def test_fpr_closed_form_limits_and_small_p(at):
    assert at.fpr_closed_form(16, 16, 0.0) == 0.0
    # small p: 1 - (1 - q)^k ~= k q with k = 2 b alpha, q = 1 / (2^f - 1)
    for bits, bucket, load in [(16, 16, 0.95), (32, 4, 0.5), (16, 4, 0.98)]:
        k, q = 2 * bucket * load, 1.0 / (2 ** bits - 1)
        assert at.fpr_closed_form(bits, bucket, load) == pytest.approx(k * q, rel=1e-3)
    # the reference demo's printed estimate for 16-bit tags, bucket 16, alpha 0.95 (src/main.cu:114-115)
    assert at.fpr_closed_form(16, 16, 0.95) == pytest.approx(2 * 16 * 0.95 / 65536, rel=1e-3)
    # exact form against a direct product for a large p
    assert at.fpr_closed_form(8, 32, 1.0) == pytest.approx(1 - (254 / 255) ** 64, rel=1e-12)


def test_candidates_are_the_built_configurations(at):
    c = at.candidates(blocks=(256,), evictions=("bfs",))
    assert len(c) == 23  # bits x bucket x word combinations the static_asserts allow
    full = at.candidates()
    assert len(full) == 23 * 2 * 4
    assert all(k["key_bytes"] == 8 and k["policy"] == "xor" for k in full)
    assert len(at.candidates(key_bytes=4, policy="offset", blocks=(128,), evictions=("dfs",))) == 23


def _r(**kw):
    base = dict(keys=1000, failed=0, fpr=1e-4, fpr_closed_form=1e-4, insert_mkeys_s=100.0, lookup_mkeys_s=100.0,
                delete_mkeys_s=100.0, bits_per_tag=16, bucket_size=16)
    base.update(kw)
    return base


def test_score_is_the_op_mix_rate(at):
    r = _r(insert_mkeys_s=10.0, lookup_mkeys_s=40.0, delete_mkeys_s=20.0)
    assert at.score(r, (1.0, 0.0, 0.0)) == pytest.approx(10.0)
    assert at.score(r, (0.5, 0.5, 0.0)) == pytest.approx(1 / (0.05 + 0.0125))


def test_choose_respects_failures_and_fpr_budget(at):
    fast_failing = _r(insert_mkeys_s=1000.0, failed=3)
    fast_fpr_too_high = _r(insert_mkeys_s=900.0, fpr=0.05, fpr_closed_form=0.05)
    closed_form_too_high = _r(insert_mkeys_s=800.0, fpr=1e-4, fpr_closed_form=0.02)
    good_slow = _r(insert_mkeys_s=50.0)
    good_fast = _r(insert_mkeys_s=120.0)
    got = at.choose([fast_failing, fast_fpr_too_high, closed_form_too_high, good_slow, good_fast], fpr_budget=1e-3)
    assert got is good_fast
    assert at.choose([fast_failing], fpr_budget=1e-3) is None


def test_choose_allows_sampling_noise_on_measured_fpr(at):
    # measured 1.05e-3 on 500 negatives is within 3 sigma of a 1e-3 budget
    noisy = _r(keys=1000, fpr=1.05e-3, fpr_closed_form=9e-4)
    assert at.choose([noisy], fpr_budget=1e-3) is noisy
    # on 10^8 negatives the same excess is significant
    tight = _r(keys=2 * 10 ** 8, fpr=1.05e-3, fpr_closed_form=9e-4)
    assert at.choose([tight], fpr_budget=1e-3) is None
