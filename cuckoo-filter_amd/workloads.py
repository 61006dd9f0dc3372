"""Synthetic key streams generated directly in device memory.

key_j = mix64(seed * 0xD1B54A32D192ED03 + j * 0x9E3779B97F4A7C15) (mod 2^64),
with mix64 the splitmix64 finaliser.  The map j -> key_j is a bijection for a
fixed seed, so one stream never repeats a key; the first n keys of a stream
are "positives" and the next m keys are guaranteed-disjoint "negatives".  The
same stream is computed on the host by the oracle (cfo_gen_keys) and by
tests/golden/gen_golden.py.

Torch int64 arithmetic wraps modulo 2^64, so the 64-bit products are exact;
right shifts are made logical by masking.
"""
from __future__ import annotations

M64 = (1 << 64) - 1
C_SEED = 0xD1B54A32D192ED03
C_STEP = 0x9E3779B97F4A7C15
MIX1 = 0xBF58476D1CE4E5B9
MIX2 = 0x94D049BB133111EB


def _s64(v: int) -> int:
    v &= M64
    return v - (1 << 64) if v >= (1 << 63) else v


def _lsr(z, s: int):
    return (z >> s) & ((1 << (64 - s)) - 1)


def gen_keys_torch(seed: int, n: int, start: int = 0, device="cuda", key_bytes: int = 8, chunk: int = 1 << 26):
    """Keys start..start+n-1 of stream `seed` as an int64 (or int32) tensor."""
    import torch

    out = torch.empty(n, dtype=torch.int64 if key_bytes == 8 else torch.int32, device=device)
    base = _s64(seed * C_SEED + start * C_STEP)
    step = _s64(C_STEP)
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        j = torch.arange(lo, hi, dtype=torch.int64, device=device)
        z = j * step + base
        z = (z ^ _lsr(z, 30)) * _s64(MIX1)
        z = (z ^ _lsr(z, 27)) * _s64(MIX2)
        z = z ^ _lsr(z, 31)
        if key_bytes == 8:
            out[lo:hi] = z
        else:
            out[lo:hi] = _low32(z)
        del j, z
    return out


def _low32(z):
    import torch

    lo = z & 0xFFFFFFFF
    lo = torch.where(lo >= (1 << 31), lo - (1 << 32), lo)
    return lo.to(torch.int32)
