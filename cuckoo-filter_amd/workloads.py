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


def dna_sequence(length: int, dup_fraction: float = 0.1, regions: int = 1024, seed: int = 42, device="cuda"):
    """ASCII ACGT sequence of `length` uniform random bases in which
    `dup_fraction` of the bases are copies of earlier stretches: `regions`
    stretches of equal length, each copied from a random earlier position to a
    random later one (the repeats a k-mer set of a real genome holds)."""
    import torch

    g = torch.Generator(device=device).manual_seed(seed)
    codes = torch.randint(0, 4, (length,), dtype=torch.uint8, device=device, generator=g)
    seq = torch.tensor([65, 67, 71, 84], dtype=torch.uint8, device=device)[codes.long()]
    del codes
    span = int(length * dup_fraction) // max(regions, 1)
    if span > 0:
        gen = torch.Generator().manual_seed(seed + 1)
        for _ in range(regions):
            src = int(torch.randint(0, length - 2 * span, (1,), generator=gen))
            dst = int(torch.randint(src + span, length - span + 1, (1,), generator=gen))
            seq[dst:dst + span] = seq[src:src + span].clone()
    return seq


def zipf_ranks(m: int, n_items: int, s: float = 1.1, seed: int = 42, device="cuda"):
    """m item indices in [0, n_items) with Zipf(s) popularity: rank r >= 1 is
    drawn by inverting the continuous power-law tail P(R > r) = r^-(s-1)
    (r = u^(-1/(s-1)), u uniform in (0, 1]), folded into [0, n_items), and
    scattered over the items by an odd multiplier mod 2^64, so the hottest
    items are not neighbours in memory."""
    import torch

    g = torch.Generator(device=device).manual_seed(seed)
    u = torch.rand(m, dtype=torch.float64, device=device, generator=g).clamp_(min=1e-300)
    r = torch.pow(u, -1.0 / (s - 1.0)).clamp_(max=2.0 ** 62)
    rank = (r.to(torch.int64) - 1) % n_items
    return torch.remainder(rank * _s64(C_STEP), n_items)
