"""The measurement kernels behind the measured ceilings (csrc/roofline.hip):
each must read the blocks its description names and write every output, or
the ceilings it reports mean nothing.  The table marks every even 32-byte
block (first word 0x5A5A5A5A, the rest zero), so a kernel's output byte says
which block it read; a torch restatement of the address computation gives
the expected bytes.  Ragged key counts check that the work-queue kernel
covers every position, routed by XCD and in order."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAGIC = 0x5A5A5A5A


def _mix(torch, z):
    # z ^= z >> 31; z *= 0x9E3779B97F4A7C15; z ^= z >> 29 (wrapping u64, logical shifts)
    z = z ^ ((z >> 31) & ((1 << 33) - 1))
    z = z * torch.tensor(0x9E3779B97F4A7C15 - (1 << 64), dtype=torch.int64, device=z.device)
    return z ^ ((z >> 29) & ((1 << 35) - 1))


@pytest.fixture(scope="module")
def rl():
    import roofline

    return roofline._lib()


def _marked_table(torch, blocks):
    t = torch.zeros(blocks * 8, dtype=torch.int32, device="cuda")
    t.view(blocks, 8)[::2, 0] = MAGIC
    return t


def _stream(torch):
    return C.c_void_p(int(torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("n", [1, 1000, (1 << 20) + 17])
def test_gather_reads_the_hashed_block(rl, device, n):
    import torch
    import workloads

    blocks = 1 << 18
    tab = _marked_table(torch, blocks)
    keys = workloads.gen_keys_torch(3, n)
    out = torch.full((n,), 0xFF, dtype=torch.uint8, device=device)
    rc = rl.rl_gather(C.c_void_p(tab.data_ptr()), blocks * 32, 32, C.c_void_p(keys.data_ptr()), n,
                      C.c_void_p(out.data_ptr()), 1, 0, 0, 256, _stream(torch))
    assert rc == 0
    torch.cuda.synchronize()
    blk = _mix(torch, keys) & (blocks - 1)
    torch.testing.assert_close(out, (blk % 2 == 0).to(torch.uint8), rtol=0, atol=0)


@pytest.mark.parametrize("routed", [1, 0])
@pytest.mark.parametrize("n", [1, 4097, (1 << 21) + 5])
def test_region_gather_covers_every_position(rl, device, routed, n):
    import torch
    import workloads

    region_bytes = 1 << 16
    blocks = 1 << 18  # 8 MiB: 128 regions of 2048 blocks
    regions, per = blocks * 32 // region_bytes, region_bytes // 32
    tab = _marked_table(torch, blocks)
    keys = workloads.gen_keys_torch(4, n)
    out = torch.full((n,), 0xFF, dtype=torch.uint8, device=device)
    grid = 8 * torch.cuda.get_device_properties(device).multi_processor_count
    rc = rl.rl_gather_region(C.c_void_p(tab.data_ptr()), blocks * 32, region_bytes, C.c_void_p(keys.data_ptr()),
                             n, C.c_void_p(out.data_ptr()), routed, grid, _stream(torch))
    assert rc == 0
    torch.cuda.synchronize()
    i = torch.arange(n, dtype=torch.int64, device=device)
    blk = (i * regions // n) * per + (_mix(torch, keys) & (per - 1))
    torch.testing.assert_close(out, (blk % 2 == 0).to(torch.uint8), rtol=0, atol=0)


def test_region_gather_rejects_bad_shapes(rl, device):
    import torch

    t = torch.zeros(1 << 16, dtype=torch.int32, device=device)
    k = torch.zeros(16, dtype=torch.int64, device=device)
    o = torch.zeros(16, dtype=torch.uint8, device=device)
    args = (C.c_void_p(k.data_ptr()), 16, C.c_void_p(o.data_ptr()), 1, 64, _stream(torch))
    assert rl.rl_gather_region(C.c_void_p(t.data_ptr()), 1 << 18, 3 << 10, *args) == -1  # region not 2^k
    assert rl.rl_gather_region(C.c_void_p(t.data_ptr()), 1 << 18, 1 << 16, *args) == -1  # 4 regions, not 8k
    assert rl.rl_gather_region(C.c_void_p(t.data_ptr()), 1 << 18, 1 << 10, *args) == 0
    torch.cuda.synchronize()


def test_copy_and_cas(rl, device):
    import torch
    import workloads

    src = torch.randint(-(2**31), 2**31 - 1, (1 << 20,), dtype=torch.int32, device=device)
    dst = torch.zeros_like(src)
    assert rl.rl_copy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), 4 << 20, 256, _stream(torch)) == 0
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
    words = torch.zeros(1 << 16, dtype=torch.int64, device=device)
    keys = workloads.gen_keys_torch(5, 100_000)
    rc = rl.rl_cas(C.c_void_p(words.data_ptr()), words.numel() * 8, C.c_void_p(keys.data_ptr()), keys.numel(), 256,
                   _stream(torch))
    assert rc == 0
    torch.cuda.synchronize()
    hit = torch.unique(_mix(torch, keys) & (words.numel() - 1))
    assert int((words != 0).sum()) == hit.numel()  # each word CAS-ed from zero exactly once
    assert bool((words[hit] & 1 == 1).all())
