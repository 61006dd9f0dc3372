"""Measured memory roofline for the cuckoo-filter kernels on one GPU.

Runs the kernels of libcuckoo_roofline.so with HIP events:
  gather32 / gather64 : uniformly random aligned 32-/64-byte block reads,
                        8-byte key in and 1-byte result out per key, over a
                        table of the filter's size (512 MiB) and 4x that;
  gather128           : random aligned 128-byte reads (one full L2 line);
  gather32nt          : gather32 with non-temporal loads;
  gather32x2 dep      : two dependent 32-byte reads per key (lookup miss shape);
  cas8                : one random 64-bit CAS per key (device scope);
  cas8 xcd            : the same, each workgroup confined to the eighth of the
                        table its XCD owns, at workgroup and at device scope;
  gather32 region     : gather32 with the keys grouped by 1 MiB table region,
                        routed to the XCD that owns the region and in order
                        (reported as G requests/s beside random gather32);
  store1 random       : one byte stored at a random place of a 64 MiB array;
  copy                : streaming copy (HBM bandwidth reference).
Each figure is algorithmic bytes (the bytes named above) per second.

    python cuckoo-filter_amd/roofline.py            # prints one JSON object
    python cuckoo-filter_amd/roofline.py --copy-only
        # only the 1 GiB copy: the known-size kernel that scales the
        # FETCH_SIZE / WRITE_SIZE counters to bytes
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
RLIB_PATH = os.path.join(_HERE, "libcuckoo_roofline.so")


def _lib():
    if not os.path.exists(RLIB_PATH):
        raise ImportError(f"{RLIB_PATH} is not built; run python cuckoo-filter_amd/build.py")
    L = C.CDLL(RLIB_PATH)
    P, sz, i = C.c_void_p, C.c_size_t, C.c_int
    L.rl_gather.argtypes = [P, sz, i, P, sz, P, i, i, i, i, P]
    L.rl_cas.argtypes = [P, sz, P, sz, i, P]
    L.rl_cas_xcd.argtypes = [P, sz, C.c_uint, i, P, sz, i, P]
    L.rl_store_bytes.argtypes = [P, sz, P, sz, i, P]
    L.rl_copy.argtypes = [P, P, sz, i, P]
    L.rl_gather_region.argtypes = [P, sz, sz, P, sz, P, i, i, P]
    L.rl_gather_coop.argtypes = [P, sz, i, i, P, sz, P, i, P]
    for f in (L.rl_gather, L.rl_cas, L.rl_cas_xcd, L.rl_store_bytes, L.rl_copy, L.rl_gather_region,
              L.rl_gather_coop):
        f.restype = i
    return L


def _time(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps / 1e3


def measure(table_bytes=512 << 20, n=1 << 26, reps=5, block=256, big_factor=4) -> dict:
    import torch

    sys.path.insert(0, _HERE)
    import workloads

    L = _lib()
    dev = torch.device("cuda:0")
    s = C.c_void_p(int(torch.cuda.current_stream().cuda_stream))
    keys = workloads.gen_keys_torch(7, n, device=dev)
    out = torch.empty(n, dtype=torch.uint8, device=dev)
    res = {"keys": n, "block": block}

    def gather(tab, tb, gb, g, dep, nt=0):
        def run():
            rc = L.rl_gather(C.c_void_p(tab.data_ptr()), tb, gb, C.c_void_p(keys.data_ptr()), n,
                             C.c_void_p(out.data_ptr()), g, dep, nt, block, s)
            assert rc == 0, rc
        t = _time(torch, run, reps)
        return (8 + 1 + g * gb) * n / t / 1e9

    for factor in (1, big_factor):
        tb = table_bytes * factor
        tab = torch.randint(0, 2**31 - 1, (tb // 4,), dtype=torch.int32, device=dev)
        tag = f"{tb >> 20}MiB"
        res[f"gather32_{tag}_gbs"] = round(gather(tab, tb, 32, 1, 0), 1)
        res[f"gather64_{tag}_gbs"] = round(gather(tab, tb, 64, 1, 0), 1)
        res[f"gather128_{tag}_gbs"] = round(gather(tab, tb, 128, 1, 0), 1)
        res[f"gather32nt_{tag}_gbs"] = round(gather(tab, tb, 32, 1, 0, 1), 1)
        res[f"gather32x2dep_{tag}_gbs"] = round(gather(tab, tb, 32, 2, 1), 1)
        if factor == 1:
            res[f"gather32_{tag}_greq"] = round(res[f"gather32_{tag}_gbs"] / (8 + 1 + 32), 1)
            props = torch.cuda.get_device_properties(dev)
            grid = 8 * props.multi_processor_count
            for routed in (1, 0):
                def region():
                    rc = L.rl_gather_region(C.c_void_p(tab.data_ptr()), tb, 1 << 20, C.c_void_p(keys.data_ptr()),
                                            n, C.c_void_p(out.data_ptr()), routed, grid, s)
                    assert rc == 0, rc
                t = _time(torch, region, reps)
                res[f"gather32_region1MiB_{'xcd' if routed else 'inorder'}_{tag}_greq"] = round(n / t / 1e9, 1)

        def cas():
            rc = L.rl_cas(C.c_void_p(tab.data_ptr()), tb, C.c_void_p(keys.data_ptr()), n, block, s)
            assert rc == 0, rc
        t = _time(torch, cas, reps)
        res[f"cas8_{tag}_gbs"] = round((8 + 8) * n / t / 1e9, 1)
        res[f"cas8_{tag}_gops"] = round(n / t / 1e9, 2)
        for wg in (1, 0):
            def cas_xcd():
                rc = L.rl_cas_xcd(C.c_void_p(tab.data_ptr()), tb, 8, wg, C.c_void_p(keys.data_ptr()), n, block, s)
                assert rc == 0, rc
            t = _time(torch, cas_xcd, reps)
            res[f"cas8_xcd_{'wg' if wg else 'agent'}_{tag}_gops"] = round(n / t / 1e9, 2)
        del tab
        torch.cuda.empty_cache()

    # address translation: past 2 GiB the random-gather rate falls with the
    # load instructions each lane issues per request; the same blocks read by
    # several lanes per key (one 16- or 32-byte part each) keep the rate
    def coop(tab, tb, gb, lanes):
        def run():
            rc = L.rl_gather_coop(C.c_void_p(tab.data_ptr()), tb, gb, lanes, C.c_void_p(keys.data_ptr()), n,
                                  C.c_void_p(out.data_ptr()), block, s)
            assert rc == 0, rc
        return n / _time(torch, run, reps) / 1e9

    for tb in (table_bytes, 8 * table_bytes):
        tab = torch.randint(0, 2**31 - 1, (tb // 4,), dtype=torch.int32, device=dev)
        tag = f"{tb >> 20}MiB"
        for gb, lanes in ((32, 1), (64, 1), (64, 2), (64, 4), (128, 1), (128, 2), (128, 4)):
            res[f"gather{gb}_lanes{lanes}_{tag}_greq"] = round(coop(tab, tb, gb, lanes), 1)
        del tab
        torch.cuda.empty_cache()

    sink = torch.empty(64 << 20, dtype=torch.uint8, device=dev)

    def store():
        rc = L.rl_store_bytes(C.c_void_p(sink.data_ptr()), sink.numel(), C.c_void_p(keys.data_ptr()), n, block, s)
        assert rc == 0, rc
    res["store1_random_64MiB_gops"] = round(n / _time(torch, store, reps) / 1e9, 2)
    del sink
    res["copy_gbs"] = measure_copy(reps=reps, block=block)
    return res


def measure_copy(copy_bytes=1 << 30, reps=5, block=256) -> float:
    """GB/s (bytes read + bytes written) of a streaming copy of copy_bytes."""
    import torch

    L = _lib()
    dev = torch.device("cuda:0")
    s = C.c_void_p(int(torch.cuda.current_stream().cuda_stream))
    src = torch.empty(copy_bytes // 4, dtype=torch.int32, device=dev)
    dst = torch.empty_like(src)

    def copy():
        rc = L.rl_copy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), copy_bytes, block, s)
        assert rc == 0, rc
    t = _time(torch, copy, reps)
    return round(2 * copy_bytes / t / 1e9, 1)


if __name__ == "__main__":
    if "--copy-only" in sys.argv[1:]:
        print(json.dumps({"copy_bytes": 1 << 30, "copy_gbs": measure_copy()}), flush=True)
    else:
        print(json.dumps(measure()), flush=True)
