"""Host-only checks of the product library (no GPU needed, no kernel launched).

* libcuckoo_amd.so loads and exports every function include/cuckoo_filter.h declares.
* Configuration validation mirrors the reference's static_asserts.
* calculateNumBuckets and the host half of getCandidateBuckets (the shared
  __host__ __device__ header code the kernels also compile) match the golden
  python-xxhash vectors.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import golden_cases


@pytest.fixture(scope="module")
def cfm():
    import cuckoo_amd

    return cuckoo_amd


def test_library_exports_every_header_symbol(cfm):
    syms = cfm.header_symbols()
    assert len(syms) >= 25
    raw = ctypes.CDLL(cfm.LIB_PATH)
    missing = [s for s in syms if not hasattr(raw, s)]
    assert not missing, missing


def test_defaults_are_the_reference_config_defaults(cfm):
    c = cfm.make_config(key_bytes=8, bits_per_tag=16)
    assert c.as_dict() == dict(key_bytes=8, bits_per_tag=16, max_evictions=500, block_size=256, bucket_size=16,
                               policy=0, eviction=1, word_bytes=8)  # DFS is the default (CuckooFilter.cuh:51)


@pytest.mark.parametrize("bad", [
    dict(bits_per_tag=12),
    dict(bucket_size=12),
    dict(word_bytes=2),
    dict(bits_per_tag=8, word_bytes=8, bucket_size=4),   # bucketSize % tagsPerWord != 0
    dict(bits_per_tag=32, word_bytes=4, bucket_size=3),
    dict(block_size=0),
    dict(block_size=2048),
    dict(policy=7),
])
def test_invalid_configs_rejected(cfm, bad):
    c = cfm.make_config()
    for k, v in bad.items():
        setattr(c, k, v)
    with pytest.raises(cfm.CuckooError) as ei:
        cfm.config_supported(c)
    assert ei.value.code == cfm.CF_ERR_INVALID


def test_every_sweep_config_is_built(cfm):
    n = 0
    for kb in (4, 8):
        for pol in ("xor", "addsub", "offset"):
            for bits in (8, 16, 32):
                for bucket in (4, 8, 16, 32):
                    for wb in (4, 8):
                        tb = {8: 1, 16: 2, 32: 4}[bits]
                        if tb > wb or bucket % (wb // tb):
                            continue
                        c = cfm.make_config(key_bytes=kb, bits_per_tag=bits, bucket_size=bucket, word_bytes=wb,
                                            policy=pol)
                        assert cfm.config_supported(c)
                        n += 1
    assert n == 138


@pytest.mark.parametrize("capacity", [1, 15, 16, 17, 1000, 10000, 1 << 20, int((1 << 20) * 0.95), (1 << 28) + 1])
def test_num_buckets_matches_oracle(cfm, oracle_mod, capacity):
    for pol in (0, 1, 2):
        for bucket in (4, 16, 32):
            c = cfm.make_config(bucket_size=bucket, policy=pol)
            oc = oracle_mod.make_config(bucket_size=bucket, policy=pol)
            assert cfm.calculate_num_buckets(c, capacity) == oracle_mod.num_buckets(oc, capacity)


def test_host_candidates_match_golden(cfm, golden):
    for name in golden_cases(golden):
        pol, kb, bits, nb = (int(x) for x in golden[name + "/meta"])
        c = cfm.make_config(key_bytes=kb, bits_per_tag=bits, policy=pol,
                            word_bytes=8, bucket_size=16)
        i1, i2, fp = cfm.candidate_buckets_host(c, nb, golden[name + "/keys"])
        np.testing.assert_array_equal(i1, golden[name + "/i1"], err_msg=name)
        np.testing.assert_array_equal(i2, golden[name + "/i2"], err_msg=name)
        np.testing.assert_array_equal(fp, golden[name + "/fp"], err_msg=name)


def test_struct_layouts_match_header(cfm, tmp_path):
    """The Python records mirror the C structs field by field (gcc on the header)."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    structs = {
        "cf_request": cfm.REQUEST_DTYPE,
        "cf_request_result": cfm.RESULT_DTYPE,
        "cf_server_options": cfm.SERVER_OPTIONS_DTYPE,
        "cf_server_stats": cfm.SERVER_STATS_DTYPE,
    }
    ctypes_structs = {"cf_config": cfm.cf_config, "cf_batch_op": cfm.cf_batch_op}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cuckoo_filter.h"', "int main(void) {"]
    for name, dt in structs.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for fld in dt.names:
            lines.append(f'printf("{name} {fld} %zu\\n", offsetof({name}, {fld}));')
    for name, st in ctypes_structs.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for fld, _ in st._fields_:
            lines.append(f'printf("{name} {fld} %zu\\n", offsetof({name}, {fld}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        name, fld, val = line.split()
        got[(name, fld)] = int(val)
    for name, dt in structs.items():
        assert got[(name, "size")] == dt.itemsize, name
        for fld in dt.names:
            assert got[(name, fld)] == dt.fields[fld][1], (name, fld)
    for name, st in ctypes_structs.items():
        assert got[(name, "size")] == ctypes.sizeof(st), name
        for fld, _ in st._fields_:
            assert got[(name, fld)] == getattr(st, fld).offset, (name, fld)


@pytest.mark.parametrize("field,value", [("window_slots", 1), ("window_slots", 9), ("max_window_requests", 0),
                                         ("max_window_keys", 0), ("ordering", 2)])
def test_server_options_rejected_before_any_device_call(cfm, field, value):
    o = np.zeros(1, cfm.SERVER_OPTIONS_DTYPE)
    cfm.lib.cf_server_options_init(o.ctypes.data)
    assert int(o["window_slots"][0]) == 2 and int(o["max_window_levels"][0]) == 0
    assert int(o["ordering"][0]) == cfm.CF_ORDER_ARRIVAL
    o[field] = value
    h = ctypes.c_void_p()
    rc = cfm.lib.cf_server_create(o.ctypes.data, ctypes.byref(h))
    assert rc == cfm.CF_ERR_INVALID and not h.value
