"""Generate the golden hashing fixtures (tests/golden/candidates.npz).

The vectors come from python-xxhash 3.x (an implementation of XXH64 that is
independent of both this repository's HIP code and its C oracle) combined with
the closed-form bucket policies of the reference:

  XorAltBucketPolicy     bucket_policies.cuh:45-74
  AddSubAltBucketPolicy  bucket_policies.cuh:126-166
  OffsetAltBucketPolicy  bucket_policies.cuh:248-285
  packed tags            CuckooFilter.cuh:159, 1185-1207

Keys are the repository's seeded stream key_j = mix64(seed*C1 + j*C2)
(splitmix64 finaliser), computed here with Python integers.

Run:  python tests/golden/gen_golden.py      (needs the `xxhash` module)
"""
from __future__ import annotations

import os

import numpy as np
import xxhash

M64 = (1 << 64) - 1
HERE = os.path.dirname(os.path.abspath(__file__))


def mix64(z: int) -> int:
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def gen_keys(seed: int, n: int, key_bytes: int) -> list[int]:
    out = []
    for j in range(n):
        k = mix64((seed * 0xD1B54A32D192ED03 + j * 0x9E3779B97F4A7C15) & M64)
        out.append(k if key_bytes == 8 else k & 0xFFFFFFFF)
    return out


def h64(value: int, nbytes: int) -> int:
    return xxhash.xxh64_intdigest(value.to_bytes(nbytes, "little"), seed=0)


def tag_bytes(bits: int) -> int:
    return 1 if bits <= 8 else (2 if bits <= 16 else 4)


def xor_triple(key, key_bytes, bits, nb):
    h = h64(key, key_bytes)
    masked = (h >> 32) & ((1 << bits) - 1)
    fp = masked if masked else 1
    i1 = (h & 0xFFFFFFFF) & (nb - 1)
    i2 = i1 ^ (h64(fp, tag_bytes(bits)) & (nb - 1))
    return i1, i2, fp


def addsub_triple(key, key_bytes, bits, nb):
    h = h64(key, key_bytes)
    masked = (h >> 32) & ((1 << bits) - 1)
    fp = masked if masked else 1
    bpb = nb // 2
    i1 = (h & 0xFFFFFFFF) % bpb
    fh = h64(fp, tag_bytes(bits)) % bpb
    i2 = ((i1 + fh) % bpb) + bpb
    return i1, i2, fp


def offset_triple(key, key_bytes, bits, nb):
    h = h64(key, key_bytes)
    masked = (h >> 32) & ((1 << (bits - 1)) - 1)
    pure = masked if masked else 1
    i1 = (h & 0xFFFFFFFF) % nb
    off = h64(pure, tag_bytes(bits)) % nb
    off = off if off else 1
    return i1, (i1 + off) % nb, pure


def packed(i1, fp, bits):
    if bits <= 8:
        return (((i1 & 0xFFFFFFFF) << bits) & 0xFFFFFFFF) | fp
    return ((i1 << bits) & M64) | fp


CASES = [
    # (name, policy, key_bytes, bits, nb)
    *[(f"xor_k{kb}_b{b}_nb{nb}", "xor", kb, b, nb)
      for kb in (8, 4) for b in (8, 16, 32) for nb in (1 << 16, 1 << 24)],
    ("xor_k8_b16_nb1", "xor", 8, 16, 1),
    ("xor_k8_b8_nb64M", "xor", 8, 8, 1 << 26),
    *[(f"addsub_k8_b{b}_nb{nb}", "addsub", 8, b, nb) for b in (8, 16, 32) for nb in (1000, 65538)],
    *[(f"offset_k8_b{b}_nb{nb}", "offset", 8, b, nb) for b in (8, 16, 32) for nb in (1000, 65537)],
    ("offset_k4_b16_nb12345", "offset", 4, 16, 12345),
]

N_KEYS = 512
SEED = 42


def main():
    fn = {"xor": xor_triple, "addsub": addsub_triple, "offset": offset_triple}
    arrays = {}
    for name, pol, kb, bits, nb in CASES:
        keys = gen_keys(SEED, N_KEYS, kb)
        # edge keys: 0, 1, all-ones
        keys[:3] = [0, 1, (1 << (8 * kb)) - 1]
        trip = [fn[pol](k, kb, bits, nb) for k in keys]
        arrays[name + "/keys"] = np.array(keys, dtype=np.uint64 if kb == 8 else np.uint32)
        arrays[name + "/i1"] = np.array([t[0] for t in trip], dtype=np.uint64)
        arrays[name + "/i2"] = np.array([t[1] for t in trip], dtype=np.uint64)
        arrays[name + "/fp"] = np.array([t[2] for t in trip], dtype=np.uint32)
        arrays[name + "/packed"] = np.array([packed(t[0], t[2], bits) for t in trip],
                                            dtype=np.uint64)
        arrays[name + "/meta"] = np.array([{"xor": 0, "addsub": 1, "offset": 2}[pol], kb, bits, nb],
                                          dtype=np.uint64)
    # raw XXH64 vectors over lengths 0..80 (every tail path of hashutil.cuh:70-161)
    blob = bytes((i * 131 + 7) & 0xFF for i in range(80))
    arrays["xxh64/blob"] = np.frombuffer(blob, dtype=np.uint8).copy()
    arrays["xxh64/seed0"] = np.array([xxhash.xxh64_intdigest(blob[:n], seed=0) for n in range(81)],
                                     dtype=np.uint64)
    arrays["xxh64/seed_prime"] = np.array(
        [xxhash.xxh64_intdigest(blob[:n], seed=2654435761) for n in range(81)], dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "candidates.npz"), **arrays)
    print("wrote", len(CASES), "cases")


if __name__ == "__main__":
    main()
