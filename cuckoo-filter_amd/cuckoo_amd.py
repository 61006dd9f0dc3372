"""Python host layer over libcuckoo_amd.so (the C ABI in include/cuckoo_filter.h).

Mirrors the reference's CuckooFilter<Config> interface
(tiltedtoast/cuckoo-filter include/CuckooFilter.cuh:139-785): the same
method names, argument meaning and return values (insert / delete return the
total occupied slots after the call).  Device buffers are torch tensors on a
ROCm device; torch is only the allocator and stream source here -- every
filter operation runs in the HIP kernels of libcuckoo_amd.so.  If the library
is missing this module raises at import time: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcuckoo_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "cuckoo_filter.h")

CF_OK, CF_ERR_INVALID, CF_ERR_UNSUPPORTED, CF_ERR_HIP, CF_ERR_ALLOC = 0, -1, -2, -3, -4
POLICIES = {"xor": 0, "addsub": 1, "offset": 2}
CF_NOALLOC_MAX_KEYS = 1 << 21  # include/cuckoo_filter.h
EVICTIONS = {"bfs": 0, "dfs": 1}


class CuckooError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cuckoo filter error {code}: {msg}")
        self.code = code


OP_INSERT, OP_LOOKUP, OP_DELETE = 0, 1, 2


class cf_batch_op(C.Structure):
    _fields_ = [
        ("filter", C.c_void_p),
        ("op", C.c_uint32),
        ("reserved", C.c_uint32),
        ("d_keys", C.c_void_p),
        ("n", C.c_size_t),
        ("d_output", C.c_void_p),
    ]


CF_REQUEST_HOST_MEMORY = 1
CF_ORDER_ARRIVAL, CF_ORDER_PHASED = 0, 1

# cf_request / cf_request_result / cf_server_options / cf_server_stats as numpy
# records (same layout as the C structs), so request lists need no per-item
# ctypes objects.
REQUEST_DTYPE = np.dtype([("filter", "u8"), ("op", "u4"), ("flags", "u4"), ("keys", "u8"), ("n", "u8"),
                          ("output", "u8")])
RESULT_DTYPE = np.dtype([("status", "i4"), ("window_levels", "u4"), ("window", "u8"), ("occupied", "u8"),
                         ("arrival_ns", "u8"), ("start_ns", "u8"), ("done_ns", "u8")])
SERVER_OPTIONS_DTYPE = np.dtype([("device", "i4"), ("max_window_requests", "u4"), ("max_window_keys", "u8"),
                                 ("staging_bytes", "u8"), ("window_slots", "u4"), ("max_window_levels", "u4"),
                                 ("ordering", "u4"), ("form_late", "u4")])
SERVER_STATS_DTYPE = np.dtype([(k, "u8") for k in ("windows", "requests", "keys", "levels", "staging_bytes",
                                                    "max_queue")])
assert REQUEST_DTYPE.itemsize == 40 and RESULT_DTYPE.itemsize == 48
assert SERVER_OPTIONS_DTYPE.itemsize == 40 and SERVER_STATS_DTYPE.itemsize == 48


class cf_config(C.Structure):
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

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def _load():
    # One HIP runtime per process: torch ships its own libamdhip64, and the
    # library resolves against whichever copy is already loaded.  Loading the
    # library first would bring in a second runtime, after which torch finds
    # no device; so torch (when installed) is imported first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run python cuckoo-filter_amd/build.py "
                          "(or __graft_entry__.build()) first")
    L = C.CDLL(LIB_PATH)
    P, sz, i32 = C.c_void_p, C.c_size_t, C.c_int
    cfgp = C.POINTER(cf_config)
    u64p, u32p = C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)
    sig = {
        "cf_last_error": (C.c_char_p, []),
        "cf_config_init": (i32, [cfgp, C.c_uint32, C.c_uint32]),
        "cf_config_supported": (i32, [cfgp]),
        "cf_calculate_num_buckets": (sz, [cfgp, sz]),
        "cf_candidate_buckets_host": (i32, [cfgp, sz, P, sz, P, P, P]),
        "cf_create": (i32, [cfgp, sz, C.POINTER(P)]),
        "cf_destroy": (i32, [P]),
        "cf_insert_many": (i32, [P, P, sz, P, C.POINTER(sz)]),
        "cf_insert_many_noalloc": (i32, [P, P, sz, P, C.POINTER(sz)]),
        "cf_insert_many_sorted": (i32, [P, P, sz, P, C.POINTER(sz)]),
        "cf_contains_many": (i32, [P, P, sz, P, P]),
        "cf_delete_many": (i32, [P, P, sz, P, P, C.POINTER(sz)]),
        "cf_clear": (i32, [P]),
        "cf_occupied_slots": (i32, [P, C.POINTER(sz)]),
        "cf_load_factor": (i32, [P, C.POINTER(C.c_float)]),
        "cf_capacity": (sz, [P]),
        "cf_num_buckets": (sz, [P]),
        "cf_size_in_bytes": (sz, [P]),
        "cf_count_occupied_slots": (i32, [P, C.POINTER(sz)]),
        "cf_eviction_count": (i32, [P, C.POINTER(sz)]),
        "cf_reset_eviction_count": (i32, [P]),
        "cf_set_eviction_histogram": (i32, [P, i32]),
        "cf_eviction_histogram": (i32, [P, P, sz]),
        "cf_export_table": (i32, [P, P, sz]),
        "cf_import_table": (i32, [P, P, sz, sz]),
        "cf_compute_candidates": (i32, [P, P, sz, P, P, P, P]),
        "cf_compute_packed_tags": (i32, [P, P, sz, P, P]),
        "cf_sort_packed_tags": (i32, [P, P, sz, P]),
        "cf_insert_packed_sorted": (i32, [P, P, sz, P, C.POINTER(sz)]),
        "cf_copy_filter": (i32, [P, P, P]),
        "cf_batch_create": (i32, [C.POINTER(P)]),
        "cf_batch_destroy": (i32, [P]),
        "cf_batch_execute": (i32, [P, P, sz, P, P]),
        "cf_run_plain_calls": (i32, [P, sz, P, P]),
        "cf_save_table": (i32, [P, C.c_char_p]),
        "cf_load_table": (i32, [C.c_char_p, C.POINTER(P)]),
        "cf_kmer_file_write": (i32, [C.c_char_p, P, sz]),
        "cf_kmer_file_count": (i32, [C.c_char_p, C.POINTER(sz)]),
        "cf_kmer_file_read": (i32, [C.c_char_p, P, sz]),
        "cf_encode_kmers": (i32, [P, sz, C.c_uint32, i32, P, P]),
        "cf_server_options_init": (i32, [P]),
        "cf_server_create": (i32, [P, C.POINTER(P)]),
        "cf_server_destroy": (i32, [P]),
        "cf_server_submit": (i32, [P, P, C.POINTER(C.c_uint64)]),
        "cf_server_wait": (i32, [P, C.c_uint64, P]),
        "cf_server_poll": (i32, [P, C.c_uint64, P]),
        "cf_server_drain": (i32, [P]),
        "cf_server_stats_get": (i32, [P, P]),
        "cf_server_stats_reset": (i32, [P]),
        "cf_server_now_ns": (C.c_uint64, []),
        "cf_server_replay": (i32, [P, P, P, sz, P]),
    }
    del u64p, u32p
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


lib = _load()


def header_symbols() -> list[str]:
    """Every function the public header declares."""
    text = open(HEADER_PATH).read()
    decl = r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z_0-9]*\s*\**\s+(cf_[a-z_0-9]+)\s*\("
    return sorted(set(re.findall(decl, text, flags=re.M)))


def _check(rc: int):
    if rc != CF_OK:
        raise CuckooError(rc, lib.cf_last_error().decode(errors="replace"))
    return rc


def make_config(key_bytes=8, bits_per_tag=16, max_evictions=500, block_size=256, bucket_size=16,
                policy="xor", eviction="dfs", word_bytes=8) -> cf_config:
    """Run-time CuckooConfig; defaults are the reference's (CuckooFilter.cuh:44-52)."""
    c = cf_config()
    _check(lib.cf_config_init(C.byref(c), key_bytes, bits_per_tag))
    c.max_evictions = max_evictions
    c.block_size = block_size
    c.bucket_size = bucket_size
    c.policy = POLICIES[policy] if isinstance(policy, str) else int(policy)
    c.eviction = EVICTIONS[eviction] if isinstance(eviction, str) else int(eviction)
    c.word_bytes = word_bytes
    return c


def config_supported(cfg: cf_config) -> bool:
    r = lib.cf_config_supported(C.byref(cfg))
    if r < 0:
        raise CuckooError(r, lib.cf_last_error().decode(errors="replace"))
    return r == 1


def calculate_num_buckets(cfg: cf_config, capacity: int) -> int:
    return int(lib.cf_calculate_num_buckets(C.byref(cfg), capacity))


def candidate_buckets_host(cfg: cf_config, num_buckets: int, keys: np.ndarray):
    """Host getCandidateBuckets (the reference's __host__ static)."""
    dt = np.uint64 if cfg.key_bytes == 8 else np.uint32
    k = np.ascontiguousarray(keys, dtype=dt)
    i1 = np.empty(k.size, np.uint64)
    i2 = np.empty(k.size, np.uint64)
    fp = np.empty(k.size, np.uint32)
    _check(lib.cf_candidate_buckets_host(C.byref(cfg), num_buckets, k.ctypes.data, k.size,
                                         i1.ctypes.data, i2.ctypes.data, fp.ctypes.data))
    return i1, i2, fp


def _torch():
    import torch  # deferred: only device operations need it

    return torch


def _stream_ptr(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(int(stream.cuda_stream))


class CuckooFilter:
    """CuckooFilter<Config> over device tensors (CuckooFilter.cuh:139-785)."""

    def __init__(self, capacity: int, config: cf_config | None = None, **cfg_kwargs):
        self.config = config if config is not None else make_config(**cfg_kwargs)
        h = C.c_void_p()
        _check(lib.cf_create(C.byref(self.config), capacity, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib.cf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------
    def _keys(self, keys):
        torch = _torch()
        if not isinstance(keys, torch.Tensor) or not keys.is_cuda:
            raise TypeError("keys must be a device tensor")
        if keys.element_size() != self.config.key_bytes or keys.dtype.is_floating_point:
            raise TypeError(f"keys must be {8 * self.config.key_bytes}-bit integers, got {keys.dtype}")
        if not keys.is_contiguous():
            raise ValueError("keys must be contiguous")
        return keys

    def _out(self, out, n):
        torch = _torch()
        if out is None:
            return None
        if out.dtype not in (torch.bool, torch.uint8) or out.numel() < n or not out.is_cuda:
            raise TypeError("output must be a device bool/uint8 tensor with one entry per key")
        return out

    def insertMany(self, keys, stream=None) -> int:
        k = self._keys(keys)
        occ = C.c_size_t()
        _check(lib.cf_insert_many(self._h, C.c_void_p(k.data_ptr()), k.numel(), _stream_ptr(stream),
                                  C.byref(occ)))
        return occ.value

    def insertManyNoAlloc(self, keys, stream=None) -> int:
        """insertMany without any device allocation (cf_insert_many_noalloc):
        scratch comes from the library's fixed per-device array."""
        k = self._keys(keys)
        occ = C.c_size_t()
        _check(lib.cf_insert_many_noalloc(self._h, C.c_void_p(k.data_ptr()), k.numel(), _stream_ptr(stream),
                                          C.byref(occ)))
        return occ.value

    def insertManySorted(self, keys, stream=None) -> int:
        k = self._keys(keys)
        occ = C.c_size_t()
        _check(lib.cf_insert_many_sorted(self._h, C.c_void_p(k.data_ptr()), k.numel(), _stream_ptr(stream),
                                         C.byref(occ)))
        return occ.value

    def containsMany(self, keys, out=None, stream=None):
        torch = _torch()
        k = self._keys(keys)
        if out is None:
            out = torch.empty(k.numel(), dtype=torch.bool, device=k.device)
        self._out(out, k.numel())
        _check(lib.cf_contains_many(self._h, C.c_void_p(k.data_ptr()), k.numel(), C.c_void_p(out.data_ptr()),
                                    _stream_ptr(stream)))
        return out

    def deleteMany(self, keys, out=None, stream=None) -> int:
        k = self._keys(keys)
        self._out(out, k.numel())
        occ = C.c_size_t()
        optr = C.c_void_p(out.data_ptr()) if out is not None else None
        _check(lib.cf_delete_many(self._h, C.c_void_p(k.data_ptr()), k.numel(), optr, _stream_ptr(stream),
                                  C.byref(occ)))
        return occ.value

    def clear(self):
        _check(lib.cf_clear(self._h))

    def loadFactor(self) -> float:
        v = C.c_float()
        _check(lib.cf_load_factor(self._h, C.byref(v)))
        return v.value

    def occupiedSlots(self) -> int:
        v = C.c_size_t()
        _check(lib.cf_occupied_slots(self._h, C.byref(v)))
        return v.value

    def capacity(self) -> int:
        return int(lib.cf_capacity(self._h))

    def getNumBuckets(self) -> int:
        return int(lib.cf_num_buckets(self._h))

    @property
    def numBuckets(self) -> int:
        return self.getNumBuckets()

    def sizeInBytes(self) -> int:
        return int(lib.cf_size_in_bytes(self._h))

    def countOccupiedSlots(self) -> int:
        v = C.c_size_t()
        _check(lib.cf_count_occupied_slots(self._h, C.byref(v)))
        return v.value

    def evictionCount(self) -> int:
        v = C.c_size_t()
        _check(lib.cf_eviction_count(self._h, C.byref(v)))
        return v.value

    def resetEvictionCount(self):
        _check(lib.cf_reset_eviction_count(self._h))

    # ------------------------------------------------------------------
    # additions: histogram, table images, device hashing
    def setEvictionHistogram(self, enable: bool = True):
        _check(lib.cf_set_eviction_histogram(self._h, 1 if enable else 0))

    def evictionHistogram(self) -> np.ndarray:
        bins = int(self.config.max_evictions) + 2
        out = np.zeros(bins, np.uint64)
        _check(lib.cf_eviction_histogram(self._h, out.ctypes.data, bins))
        return out

    def exportTable(self) -> np.ndarray:
        img = np.empty(self.sizeInBytes(), np.uint8)
        _check(lib.cf_export_table(self._h, img.ctypes.data, img.size))
        return img

    def importTable(self, image: np.ndarray, occupied: int):
        img = np.ascontiguousarray(image, dtype=np.uint8)
        _check(lib.cf_import_table(self._h, img.ctypes.data, img.size, occupied))

    def computeCandidates(self, keys, stream=None):
        torch = _torch()
        k = self._keys(keys)
        n = k.numel()
        i1 = torch.empty(n, dtype=torch.int64, device=k.device)
        i2 = torch.empty(n, dtype=torch.int64, device=k.device)
        fp = torch.empty(n, dtype=torch.int32, device=k.device)
        _check(lib.cf_compute_candidates(self._h, C.c_void_p(k.data_ptr()), n, C.c_void_p(i1.data_ptr()),
                                         C.c_void_p(i2.data_ptr()), C.c_void_p(fp.data_ptr()),
                                         _stream_ptr(stream)))
        return i1, i2, fp

    def copyFrom(self, src: "CuckooFilter", stream=None):
        """Device copy of src's table and counters into this filter."""
        _check(lib.cf_copy_filter(self._h, src._h, _stream_ptr(stream)))

    def saveTable(self, path: str):
        """Header + table image file (cf_save_table)."""
        _check(lib.cf_save_table(self._h, os.fsencode(path)))

    @classmethod
    def loadTable(cls, path: str) -> "CuckooFilter":
        """A new filter (current device) from a cf_save_table file."""
        h = C.c_void_p()
        _check(lib.cf_load_table(os.fsencode(path), C.byref(h)))
        f = cls.__new__(cls)
        f._h = h
        cfg = cf_config()
        # the header's configuration: re-read it from the file
        with open(path, "rb") as fh:
            hdr = fh.read(80)
        C.memmove(C.addressof(cfg), hdr[16:48], 32)
        f.config = cfg
        return f

    def computePackedTags(self, keys, stream=None):
        torch = _torch()
        k = self._keys(keys)
        dt = torch.int32 if self.config.bits_per_tag <= 8 else torch.int64
        out = torch.empty(k.numel(), dtype=dt, device=k.device)
        _check(lib.cf_compute_packed_tags(self._h, C.c_void_p(k.data_ptr()), k.numel(),
                                          C.c_void_p(out.data_ptr()), _stream_ptr(stream)))
        return out

    def sortPackedTags(self, packed, stream=None):
        """In-place ascending sort of packed tags (the cub::DeviceRadixSort
        step of insertManySorted, CuckooFilter.cuh:485-498)."""
        self._packed(packed)
        _check(lib.cf_sort_packed_tags(self._h, C.c_void_p(packed.data_ptr()), packed.numel(),
                                       _stream_ptr(stream)))
        return packed

    def insertPackedSorted(self, packed, stream=None) -> int:
        """insertKernelSorted over packed tags (CuckooFilter.cuh:1209-1266);
        returns the total occupancy afterwards."""
        self._packed(packed)
        occ = C.c_size_t()
        _check(lib.cf_insert_packed_sorted(self._h, C.c_void_p(packed.data_ptr()), packed.numel(),
                                           _stream_ptr(stream), C.byref(occ)))
        return occ.value

    def _packed(self, packed):
        torch = _torch()
        want = 4 if self.config.bits_per_tag <= 8 else 8
        if (not isinstance(packed, torch.Tensor) or not packed.is_cuda or packed.element_size() != want
                or packed.dtype.is_floating_point or not packed.is_contiguous()):
            raise TypeError(f"packed tags must be a contiguous device tensor of {8 * want}-bit integers")


class Batch:
    """Batched many-filter executor (cf_batch_execute): a window of insert /
    lookup / delete batches over many filters of one configuration, run with
    one graph launch and the results of running them one by one."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib.cf_batch_create(C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib.cf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def make_ops(ops):
        """ops: sequence of (filter, op, keys tensor, output tensor or None)."""
        arr = (cf_batch_op * len(ops))()
        for i, (f, op, keys, out) in enumerate(ops):
            arr[i].filter = f._h
            arr[i].op = op
            arr[i].d_keys = keys.data_ptr() if keys.numel() else None
            arr[i].n = keys.numel()
            arr[i].d_output = out.data_ptr() if out is not None else None
        return arr

    def execute(self, ops, stream=None, want_occupied=True):
        arr = ops if isinstance(ops, C.Array) else self.make_ops(ops)
        occ = np.zeros(len(arr), np.uint64) if want_occupied else None
        _check(lib.cf_batch_execute(self._h, arr, len(arr), _stream_ptr(stream),
                                    occ.ctypes.data if occ is not None else None))
        return occ


def run_plain_calls(ops, stream=None, want_occupied=True):
    """ops (a Batch.make_ops array, or (filter, op, keys, out) tuples) one at a
    time through the plain calls, from the library's C loop."""
    arr = ops if isinstance(ops, C.Array) else Batch.make_ops(ops)
    occ = np.zeros(len(arr), np.uint64) if want_occupied else None
    _check(lib.cf_run_plain_calls(arr, len(arr), _stream_ptr(stream), occ.ctypes.data if occ is not None else None))
    return occ


def _ptr(buf) -> int:
    """Address of a torch tensor or numpy array (0 for None / empty)."""
    if buf is None:
        return 0
    if isinstance(buf, np.ndarray):
        return buf.ctypes.data if buf.size else 0
    return buf.data_ptr() if buf.numel() else 0


class Server:
    """Serving loop (cf_server_*): requests against resident filters, queued in
    arrival order and run in continuously formed batched windows.

    A request is (filter, op, keys, out, host): keys / out are device tensors,
    or numpy arrays with host=True (staged through the server's device arena).
    Results equal running the requests one by one in arrival order; with
    ordering="phased" a window's requests are concurrent and run per filter as
    its inserts (together), then its lookups and deletes (together).
    window_slots (2..8, default 2) is how many windows are formed ahead of the
    running one; 3 lets host-memory staging overlap the uploads.  With
    form_late a window is formed only once the one before it has reached its
    last level, so that requests arriving meanwhile join it."""

    def __init__(self, device=None, max_window_requests=None, max_window_keys=None, staging_bytes=None,
                 window_slots=None, max_window_levels=None, ordering="arrival", form_late=False):
        o = np.zeros(1, SERVER_OPTIONS_DTYPE)
        _check(lib.cf_server_options_init(o.ctypes.data))
        if device is not None:
            o["device"] = device
        if max_window_requests is not None:
            o["max_window_requests"] = max_window_requests
        if max_window_keys is not None:
            o["max_window_keys"] = max_window_keys
        if staging_bytes is not None:
            o["staging_bytes"] = staging_bytes
        if window_slots is not None:
            o["window_slots"] = window_slots
        if max_window_levels is not None:
            o["max_window_levels"] = max_window_levels
        o["ordering"] = {"arrival": CF_ORDER_ARRIVAL, "phased": CF_ORDER_PHASED}[ordering]
        o["form_late"] = int(bool(form_late))
        h = C.c_void_p()
        _check(lib.cf_server_create(o.ctypes.data, C.byref(h)))
        self._h = h
        self._alive = {}  # ticket -> buffers kept alive until the request completes

    def close(self):
        if getattr(self, "_h", None):
            _check(lib.cf_server_destroy(self._h))
            self._h = None
            self._alive.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def make_requests(reqs):
        """reqs: sequence of (filter, op, keys, out, host) -> REQUEST_DTYPE array."""
        arr = np.zeros(len(reqs), REQUEST_DTYPE)
        for i, (f, op, keys, out, host) in enumerate(reqs):
            n = keys.size if isinstance(keys, np.ndarray) else keys.numel()
            arr[i] = (f._h.value, op, CF_REQUEST_HOST_MEMORY if host else 0, _ptr(keys), n, _ptr(out))
        return arr

    def submit(self, f, op, keys, out=None, host=False) -> int:
        r = self.make_requests([(f, op, keys, out, host)])
        t = C.c_uint64()
        _check(lib.cf_server_submit(self._h, r.ctypes.data, C.byref(t)))
        self._alive[t.value] = (keys, out)
        return t.value

    def _result(self, res):
        return {k: int(res[k][0]) for k in RESULT_DTYPE.names}

    def wait(self, ticket: int) -> dict:
        res = np.zeros(1, RESULT_DTYPE)
        _check(lib.cf_server_wait(self._h, ticket, res.ctypes.data))
        self._alive.pop(ticket, None)
        return self._result(res)

    def poll(self, ticket: int):
        res = np.zeros(1, RESULT_DTYPE)
        rc = lib.cf_server_poll(self._h, ticket, res.ctypes.data)
        if rc < 0:
            _check(rc)
        if rc == 0:
            return None
        self._alive.pop(ticket, None)
        return self._result(res)

    def drain(self):
        _check(lib.cf_server_drain(self._h))

    def stats(self) -> dict:
        st = np.zeros(1, SERVER_STATS_DTYPE)
        _check(lib.cf_server_stats_get(self._h, st.ctypes.data))
        return {k: int(st[k][0]) for k in SERVER_STATS_DTYPE.names}

    def reset_stats(self):
        _check(lib.cf_server_stats_reset(self._h))

    def replay(self, reqs: np.ndarray, arrival_s: np.ndarray) -> np.ndarray:
        """Open-loop replay of a REQUEST_DTYPE array at arrival_s (seconds from
        the start, non-decreasing); returns a RESULT_DTYPE array."""
        reqs = np.ascontiguousarray(reqs, dtype=REQUEST_DTYPE)
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        if arr.shape != (reqs.size,):
            raise ValueError("arrival_s must have one entry per request")
        res = np.zeros(reqs.size, RESULT_DTYPE)
        _check(lib.cf_server_replay(self._h, reqs.ctypes.data, arr.ctypes.data, reqs.size, res.ctypes.data))
        return res


def server_now_ns() -> int:
    return int(lib.cf_server_now_ns())


def write_kmer_file(path: str, keys: np.ndarray):
    """k-mer key file: uint64 N then N uint64 (src/kmc_to_binary.cpp:1-7 format)."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    _check(lib.cf_kmer_file_write(os.fsencode(path), k.ctypes.data, k.size))


def read_kmer_file(path: str) -> np.ndarray:
    n = C.c_size_t()
    _check(lib.cf_kmer_file_count(os.fsencode(path), C.byref(n)))
    out = np.empty(n.value, np.uint64)
    _check(lib.cf_kmer_file_read(os.fsencode(path), out.ctypes.data, out.size))
    return out


def encode_kmers(seq, k: int, canonical: bool = True, stream=None):
    """2-bit keys of every k-mer of an ASCII DNA sequence (device uint8 tensor)."""
    torch = _torch()
    n = seq.numel() - k + 1
    out = torch.empty(max(n, 0), dtype=torch.int64, device=seq.device)
    _check(lib.cf_encode_kmers(C.c_void_p(seq.data_ptr()), seq.numel(), k, 1 if canonical else 0,
                               C.c_void_p(out.data_ptr()), _stream_ptr(stream)))
    return out


def keys_to_device(keys: np.ndarray, device="cuda"):
    """Host uint64/uint32 array -> device tensor holding the same bit patterns."""
    torch = _torch()
    if keys.dtype == np.uint64:
        t = torch.from_numpy(keys.view(np.int64))
    elif keys.dtype == np.uint32:
        t = torch.from_numpy(keys.view(np.int32))
    else:
        raise TypeError(keys.dtype)
    return t.to(device)


def to_numpy_u(t, dtype) -> np.ndarray:
    """Device integer tensor -> host numpy array reinterpreted as unsigned dtype."""
    return t.cpu().numpy().view(dtype)
