"""ctypes wrapper around the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker.  The oracle restates the
reference algorithm (tiltedtoast/cuckoo-filter include/CuckooFilter.cuh,
include/bucket_policies.cuh, include/hashutil.cuh) in plain C; see
cuckoo_oracle.h for what pins each function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

POLICY_XOR, POLICY_ADDSUB, POLICY_OFFSET = 0, 1, 2
EVICT_BFS, EVICT_DFS = 0, 1


class Config(C.Structure):
    """Same field order as cf_config (include/cuckoo_filter.h)."""

    _fields_ = [
        ("key_bytes", C.c_uint32),
        ("bits_per_tag", C.c_uint32),
        ("max_evictions", C.c_uint32),
        ("block_size", C.c_uint32),
        ("bucket_size", C.c_uint32),
        ("policy", C.c_uint32),
        ("eviction", C.c_uint32),
        ("word_bytes", C.c_uint32),
    ]


def make_config(key_bytes=8, bits_per_tag=16, max_evictions=500, block_size=256,
                bucket_size=16, policy=POLICY_XOR, eviction=EVICT_DFS, word_bytes=8) -> Config:
    """Defaults mirror CuckooConfig's (CuckooFilter.cuh:44-52): DFS, u64 words."""
    return Config(key_bytes, bits_per_tag, max_evictions, block_size, bucket_size, policy,
                  eviction, word_bytes)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        u64, u32, sz = C.c_uint64, C.c_uint32, C.c_size_t
        cfgp = C.POINTER(Config)
        sig = {
            "cfo_xxh64": (u64, [P, sz, u64]),
            "cfo_next_pow2": (u64, [u64]),
            "cfo_config_check": (C.c_int, [cfgp]),
            "cfo_num_buckets": (u64, [cfgp, u64]),
            "cfo_candidates": (None, [cfgp, u64, P, sz, P, P, P]),
            "cfo_alt_bucket": (u64, [cfgp, u64, u64, u32, C.POINTER(u32)]),
            "cfo_packed_tags": (None, [cfgp, u64, P, sz, P]),
            "cfo_create": (P, [cfgp, u64]),
            "cfo_destroy": (None, [P]),
            "cfo_clear": (None, [P]),
            "cfo_num_buckets_of": (u64, [P]),
            "cfo_size_bytes": (u64, [P]),
            "cfo_occupied": (u64, [P]),
            "cfo_evictions": (u64, [P]),
            "cfo_count_occupied_slots": (u64, [P]),
            "cfo_table": (P, [P]),
            "cfo_set_occupied": (None, [P, u64]),
            "cfo_insert_many": (u64, [P, P, sz, P, P, C.c_int]),
            "cfo_insert_many_sorted": (u64, [P, P, sz, P]),
            "cfo_insert_many_phased": (u64, [P, P, sz, P, P]),
            "cfo_contains_many": (None, [P, P, sz, P, C.c_int]),
            "cfo_delete_many": (u64, [P, P, sz, P, C.c_int]),
            "cfo_bucket_contains": (C.c_int, [P, u64, u32]),
            "cfo_probe_counts": (None, [P, P, sz, P]),
            "cfo_gen_keys": (None, [u64, u64, sz, u32, P]),
            "cfo_stats": (None, [P, C.POINTER(u64), C.POINTER(u64)]),
            "cfo_reset_stats": (None, [P]),
            "cfo_swar_zero_mask": (u64, [u64, u32, u32]),
            "cfo_swar_replicate": (u64, [u32, u32, u32]),
            # the single-threaded CPU filter of cpu_cuckoo.h (baseline only)
            "ccf_create": (P, [u64]),
            "ccf_destroy": (None, [P]),
            "ccf_num_buckets": (u64, [P]),
            "ccf_size_bytes": (u64, [P]),
            "ccf_items": (u64, [P]),
            "ccf_victim_used": (C.c_int, [P]),
            "ccf_alt_index": (u64, [P, u64, u32]),
            "ccf_index_tag": (None, [P, u64, C.POINTER(u64), C.POINTER(u32)]),
            "ccf_checksum": (u64, [P]),
            "ccf_add_many": (u64, [P, P, sz, P]),
            "ccf_contain_many": (u64, [P, P, sz, P]),
            "ccf_delete_many": (u64, [P, P, sz, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def key_dtype(cfg: Config):
    return np.uint64 if cfg.key_bytes == 8 else np.uint32


def xxh64(data: bytes, seed: int = 0) -> int:
    buf = C.create_string_buffer(data, len(data))
    return int(lib().cfo_xxh64(buf, len(data), seed))


def num_buckets(cfg: Config, capacity: int) -> int:
    return int(lib().cfo_num_buckets(C.byref(cfg), capacity))


def gen_keys(seed: int, n: int, key_bytes: int = 8, start: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64 if key_bytes == 8 else np.uint32)
    lib().cfo_gen_keys(seed, start, n, key_bytes, _ptr(out))
    return out


def candidates(cfg: Config, nb: int, keys: np.ndarray):
    keys = np.ascontiguousarray(keys, dtype=key_dtype(cfg))
    n = keys.size
    i1 = np.empty(n, np.uint64)
    i2 = np.empty(n, np.uint64)
    fp = np.empty(n, np.uint32)
    lib().cfo_candidates(C.byref(cfg), nb, _ptr(keys), n, _ptr(i1), _ptr(i2), _ptr(fp))
    return i1, i2, fp


def alt_bucket(cfg: Config, nb: int, bucket: int, fp: int):
    nf = C.c_uint32(0)
    b = lib().cfo_alt_bucket(C.byref(cfg), nb, bucket, fp, C.byref(nf))
    return int(b), int(nf.value)


def packed_tags(cfg: Config, nb: int, keys: np.ndarray) -> np.ndarray:
    keys = np.ascontiguousarray(keys, dtype=key_dtype(cfg))
    out = np.empty(keys.size, np.uint64)
    lib().cfo_packed_tags(C.byref(cfg), nb, _ptr(keys), keys.size, _ptr(out))
    return out


class Filter:
    """Sequential oracle filter with the reference's batch API shape."""

    def __init__(self, cfg: Config, capacity: int):
        self.cfg = cfg
        self._h = lib().cfo_create(C.byref(cfg), capacity)
        if not self._h:
            raise ValueError("invalid oracle configuration or capacity")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().cfo_destroy(h)
            self._h = None

    @property
    def num_buckets(self) -> int:
        return int(lib().cfo_num_buckets_of(self._h))

    @property
    def size_bytes(self) -> int:
        return int(lib().cfo_size_bytes(self._h))

    def capacity(self) -> int:
        return self.num_buckets * self.cfg.bucket_size

    def occupied(self) -> int:
        return int(lib().cfo_occupied(self._h))

    def evictions(self) -> int:
        return int(lib().cfo_evictions(self._h))

    def count_occupied_slots(self) -> int:
        return int(lib().cfo_count_occupied_slots(self._h))

    def clear(self):
        lib().cfo_clear(self._h)

    def _k(self, keys):
        return np.ascontiguousarray(keys, dtype=key_dtype(self.cfg))

    def insert_many(self, keys, threads=1, want_chain=False):
        keys = self._k(keys)
        ok = np.zeros(keys.size, np.uint8)
        chain = np.zeros(keys.size, np.uint32) if want_chain else None
        occ = lib().cfo_insert_many(self._h, _ptr(keys), keys.size, _ptr(ok), _ptr(chain), threads)
        return (int(occ), ok.astype(bool), chain) if want_chain else (int(occ), ok.astype(bool))

    def insert_many_sorted(self, keys, want_chain=False):
        keys = self._k(keys)
        chain = np.zeros(keys.size, np.uint32) if want_chain else None
        occ = lib().cfo_insert_many_sorted(self._h, _ptr(keys), keys.size, _ptr(chain))
        return (int(occ), chain) if want_chain else int(occ)

    def insert_many_phased(self, keys):
        """The insert scheduled by step over the batch (all primary attempts,
        then secondaries, then BFS loops, then DFS), sequentially; returns
        (occupied, ok, chain)."""
        keys = self._k(keys)
        ok = np.zeros(keys.size, np.uint8)
        chain = np.zeros(keys.size, np.uint32)
        occ = lib().cfo_insert_many_phased(self._h, _ptr(keys), keys.size, _ptr(ok), _ptr(chain))
        return int(occ), ok.astype(bool), chain

    def contains_many(self, keys, threads=1) -> np.ndarray:
        keys = self._k(keys)
        out = np.zeros(keys.size, np.uint8)
        lib().cfo_contains_many(self._h, _ptr(keys), keys.size, _ptr(out), threads)
        return out.astype(bool)

    def delete_many(self, keys, threads=1):
        keys = self._k(keys)
        ok = np.zeros(keys.size, np.uint8)
        occ = lib().cfo_delete_many(self._h, _ptr(keys), keys.size, _ptr(ok), threads)
        return int(occ), ok.astype(bool)

    def probe_counts(self, keys) -> np.ndarray:
        keys = self._k(keys)
        out = np.zeros(keys.size, np.uint8)
        lib().cfo_probe_counts(self._h, _ptr(keys), keys.size, _ptr(out))
        return out

    def stats(self):
        """(bucket reads, word writes) of the insert/delete calls since reset_stats()."""
        r, w = C.c_uint64(), C.c_uint64()
        lib().cfo_stats(self._h, C.byref(r), C.byref(w))
        return int(r.value), int(w.value)

    def reset_stats(self):
        lib().cfo_reset_stats(self._h)

    def bucket_contains(self, b: int, tag: int) -> bool:
        return bool(lib().cfo_bucket_contains(self._h, b, tag))

    def table_bytes(self) -> np.ndarray:
        """Copy of the raw table image (little-endian words)."""
        return self.table_view().copy()

    def table_view(self) -> np.ndarray:
        """The raw table image in place (writes change the filter's table;
        valid while the filter lives)."""
        p = lib().cfo_table(self._h)
        n = self.size_bytes
        return np.ctypeslib.as_array((C.c_uint8 * n).from_address(p))

    def load_table(self, image: np.ndarray, occupied: int):
        img = np.ascontiguousarray(image, dtype=np.uint8)
        assert img.size == self.size_bytes
        C.memmove(lib().cfo_table(self._h), img.ctypes.data, img.size)
        lib().cfo_set_occupied(self._h, occupied)


class CpuCuckoo:
    """The single-threaded CPU cuckoo filter of cpu_cuckoo.h: a restatement of
    the published design behind cuckoofilter::CuckooFilter<uint64_t, 16>, the
    CPU filter the reference compares against
    (benchmark/cuckoo_cpu_vs_gpu.cu:52-155).  A reported baseline only."""

    def __init__(self, max_keys: int):
        self._h = lib().ccf_create(max_keys)
        if not self._h:
            raise MemoryError("ccf_create failed")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().ccf_destroy(h)
            self._h = None

    def num_buckets(self) -> int:
        return int(lib().ccf_num_buckets(self._h))

    def size_bytes(self) -> int:
        return int(lib().ccf_size_bytes(self._h))

    def items(self) -> int:
        return int(lib().ccf_items(self._h))

    def victim_used(self) -> bool:
        return bool(lib().ccf_victim_used(self._h))

    def alt_index(self, index: int, tag: int) -> int:
        return int(lib().ccf_alt_index(self._h, index, tag))

    def index_tag(self, key: int):
        i, t = C.c_uint64(), C.c_uint32()
        lib().ccf_index_tag(self._h, key, C.byref(i), C.byref(t))
        return int(i.value), int(t.value)

    def checksum(self) -> int:
        return int(lib().ccf_checksum(self._h))

    def _run(self, fn, keys, want_out):
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.zeros(k.size, dtype=np.uint8) if want_out else None
        ok = int(fn(self._h, _ptr(k), k.size, _ptr(out)))
        return (ok, out.astype(bool)) if want_out else ok

    def add_many(self, keys, want_out=False):
        return self._run(lib().ccf_add_many, keys, want_out)

    def contain_many(self, keys, want_out=False):
        return self._run(lib().ccf_contain_many, keys, want_out)

    def delete_many(self, keys, want_out=False):
        return self._run(lib().ccf_delete_many, keys, want_out)


def table_tags(image: np.ndarray, cfg: Config) -> np.ndarray:
    """Decode a raw table image into a (num_buckets, bucket_size) array of tags
    (slot s <-> word s / tagsPerWord, lane s % tagsPerWord; CuckooFilter.cuh:226-249)."""
    tag_dt = {8: np.uint8, 16: np.uint16, 32: np.uint32}[cfg.bits_per_tag]
    # little-endian packing of equal-width lanes inside LE words is the same as
    # a flat LE array of tags
    tags = np.frombuffer(np.ascontiguousarray(image).tobytes(), dtype=tag_dt)
    return tags.reshape(-1, cfg.bucket_size)
