"""Device k-mer encoding and table files on the GPU.

* cf_encode_kmers against a numpy restatement of the reference converter's
  2-bit code (src/kmc_to_binary.cpp:17-46: A=0 C=1 G=2 T=3, first base in the
  highest bits) and of canonical form (min of k-mer and reverse complement),
  with hand-checked values: "ACGT" -> 0b00011011, canonical("TTTT") = "AAAA".
* cf_save_table / cf_load_table round trip: same configuration, bucket count,
  image, counters and lookups.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3,
        ord("a"): 0, ord("c"): 1, ord("g"): 2, ord("t"): 3}


def ref_kmers(seq: bytes, k: int, canonical: bool):
    out = []
    for i in range(len(seq) - k + 1):
        win = seq[i:i + k]
        if any(c not in CODE for c in win):
            out.append((1 << 64) - 1)
            continue
        fwd = 0
        for c in win:
            fwd = (fwd << 2) | CODE[c]
        rev = 0
        for c in reversed(win):
            rev = (rev << 2) | (3 - CODE[c])
        out.append(min(fwd, rev) if canonical else fwd)
    return np.array(out, dtype=np.uint64)


def _dev_seq(seq: bytes):
    import torch

    return torch.frombuffer(bytearray(seq), dtype=torch.uint8).cuda()


def test_encode_known_values(cf):
    assert cf.to_numpy_u(cf.encode_kmers(_dev_seq(b"ACGT"), 4, canonical=False), np.uint64)[0] == 0b00011011
    assert cf.to_numpy_u(cf.encode_kmers(_dev_seq(b"TTTT"), 4, canonical=True), np.uint64)[0] == 0
    assert cf.to_numpy_u(cf.encode_kmers(_dev_seq(b"ACGN"), 4), np.uint64)[0] == (1 << 64) - 1


@pytest.mark.parametrize("k", [1, 5, 21, 31, 32])
@pytest.mark.parametrize("canonical", [False, True])
def test_encode_matches_reference_code(cf, k, canonical):
    rng = np.random.default_rng(k)
    seq = bytes(rng.choice(list(b"ACGTacgt"), size=3000).tolist())
    seq = seq[:1000] + b"N" + seq[1001:]  # one invalid base
    got = cf.to_numpy_u(cf.encode_kmers(_dev_seq(seq), k, canonical=canonical), np.uint64)
    np.testing.assert_array_equal(got, ref_kmers(seq, k, canonical))


def test_table_file_round_trip(cf, oracle_mod, tmp_path):
    c = cf.make_config(bits_per_tag=16, eviction="bfs")
    f = cf.CuckooFilter(1 << 18, config=c)
    keys = oracle_mod.gen_keys(8, int(f.capacity() * 0.97))
    dk = cf.keys_to_device(keys)
    occ = f.insertMany(dk)
    p = str(tmp_path / "f.tbl")
    f.saveTable(p)
    g = cf.CuckooFilter.loadTable(p)
    assert g.getNumBuckets() == f.getNumBuckets() and g.occupiedSlots() == occ
    assert g.evictionCount() == f.evictionCount()
    assert g.config.as_dict() == c.as_dict()
    np.testing.assert_array_equal(g.exportTable(), f.exportTable())
    q = cf.keys_to_device(oracle_mod.gen_keys(9, 100000))
    assert bool((g.containsMany(q) == f.containsMany(q)).all().item())
    assert bool(g.containsMany(dk).all().item())
