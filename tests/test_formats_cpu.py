"""Host-side file formats of the C ABI (no GPU): the reference's k-mer key
file (uint64 N + N uint64, src/kmc_to_binary.cpp:1-7, read back as
benchmark/kmer_benchmark.cu:26-47 does) and the table-file header layout."""
import ctypes
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def cfm():
    import cuckoo_amd

    return cuckoo_amd


def test_kmer_file_round_trip_and_layout(cfm, tmp_path):
    keys = np.array([0, 1, 27, (1 << 62) - 1, 0xDEADBEEFCAFEBABE], dtype=np.uint64)
    p = str(tmp_path / "k.bin")
    cfm.write_kmer_file(p, keys)
    raw = open(p, "rb").read()
    assert len(raw) == 8 + 8 * keys.size
    assert int.from_bytes(raw[:8], "little") == keys.size          # uint64 N first
    np.testing.assert_array_equal(np.frombuffer(raw[8:], "<u8"), keys)
    np.testing.assert_array_equal(cfm.read_kmer_file(p), keys)


def test_kmer_file_errors(cfm, tmp_path):
    p = str(tmp_path / "short.bin")
    with open(p, "wb") as f:
        f.write((5).to_bytes(8, "little") + b"\\0" * 16)  # claims 5 keys, holds 2
    with pytest.raises(cfm.CuckooError):
        cfm.read_kmer_file(p)
    with pytest.raises(cfm.CuckooError):
        cfm.read_kmer_file(str(tmp_path / "missing.bin"))


def test_table_file_rejects_foreign_files(cfm, tmp_path):
    p = str(tmp_path / "bogus.tbl")
    with open(p, "wb") as f:
        f.write(b"NOTATABL" + bytes(200))
    h = ctypes.c_void_p()
    rc = cfm.lib.cf_load_table(os.fsencode(p), ctypes.byref(h))
    assert rc == cfm.CF_ERR_INVALID and not h.value
