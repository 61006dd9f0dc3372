"""Host-side workload generators used by bench.py and the GPU tests (run on
the CPU with torch's CPU tensors).

* gen_keys_torch reproduces the oracle's key stream (cfo_gen_keys) bit for
  bit, for 8- and 4-byte keys and any start offset.
* zipf_ranks: ranks stay in range, the hottest item's share matches the
  closed form 1 - 2^-(s-1) of the continuous tail it inverts, and for a
  power-of-two item count the odd multiplier scatters distinct ranks to
  distinct items.
* dna_sequence: only A/C/G/T, and the requested fraction of the sequence is
  copied from earlier stretches, which shows up as duplicated k-mers.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def wl():
    import workloads

    return workloads


@pytest.mark.parametrize("seed,n,start", [(42, 1000, 0), (7, 513, 12345), (0, 64, (1 << 40) + 3)])
def test_gen_keys_torch_matches_oracle(wl, oracle_mod, seed, n, start):
    got = wl.gen_keys_torch(seed, n, start=start, device="cpu").numpy().view(np.uint64)
    want = oracle_mod.gen_keys(seed, start + n)[start:] if start < 1 << 20 else None
    if want is None:  # far offsets: compare against the closed form directly
        C, S = 0xD1B54A32D192ED03, 0x9E3779B97F4A7C15
        M = (1 << 64) - 1

        def mix(z):
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
            return z ^ (z >> 31)

        want = np.array([mix((seed * C + (start + j) * S) & M) for j in range(n)], dtype=np.uint64)
    np.testing.assert_array_equal(got, want)


def test_gen_keys_torch_32bit_is_low_half(wl):
    k64 = wl.gen_keys_torch(3, 300, device="cpu").numpy().view(np.uint64)
    k32 = wl.gen_keys_torch(3, 300, device="cpu", key_bytes=4).numpy().view(np.uint32)
    np.testing.assert_array_equal(k32, (k64 & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def test_zipf_ranks_head_share_and_range(wl):
    n_items, m, s = 1 << 16, 1 << 20, 1.1
    r = wl.zipf_ranks(m, n_items, s=s, seed=1, device="cpu").numpy()
    assert r.min() >= 0 and r.max() < n_items
    _, counts = np.unique(r, return_counts=True)
    head = counts.max() / m
    want = 1 - 2 ** -(s - 1)  # P(rank 1) = P(u^(-1/(s-1)) < 2)
    assert abs(head - want) < 4 * np.sqrt(want * (1 - want) / m) + 1e-3, (head, want)


def test_zipf_scatter_is_a_bijection_for_powers_of_two(wl):
    import torch

    n_items = 1 << 12
    ranks = torch.arange(n_items, dtype=torch.int64)
    items = torch.remainder(ranks * wl._s64(wl.C_STEP), n_items).numpy()
    assert np.unique(items).size == n_items


def test_dna_sequence_alphabet_and_duplication(wl):
    L, k = 1 << 18, 31
    seq = wl.dna_sequence(L, dup_fraction=0.1, regions=64, seed=9, device="cpu").numpy()
    assert seq.size == L and set(np.unique(seq).tolist()) <= {65, 67, 71, 84}
    codes = np.searchsorted(np.array([65, 67, 71, 84]), seq).astype(np.uint64)
    kmers = np.zeros(L - k + 1, dtype=np.uint64)
    for j in range(k):
        kmers = (kmers << np.uint64(2)) | codes[j:L - k + 1 + j]
    dup = 1 - np.unique(kmers).size / kmers.size
    # ~10% of the bases sit in copied stretches; random 31-mers almost never repeat
    assert 0.07 < dup < 0.11, dup
    plain = wl.dna_sequence(L, dup_fraction=0.0, seed=9, device="cpu").numpy()
    codes = np.searchsorted(np.array([65, 67, 71, 84]), plain).astype(np.uint64)
    kmers = np.zeros(L - k + 1, dtype=np.uint64)
    for j in range(k):
        kmers = (kmers << np.uint64(2)) | codes[j:L - k + 1 + j]
    assert np.unique(kmers).size == kmers.size
