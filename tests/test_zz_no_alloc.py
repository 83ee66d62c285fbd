"""kmst_solve_device_noalloc (the fixed-scratch variant of kmst_solve_device,
no device allocation) through the C ABI against kmst_solve_device on the same
inputs: identical MSTs on consecutive calls (the second smaller, so stale
scratch contents from the first must not matter), at the size limit, with
positional keys, on a caller's stream; sizes past the limit are refused."""
import numpy as np
import pytest

import kamsta_amd as k

pytestmark = pytest.mark.gpu


def gnm(torch, ctx, n, m, max_w, seed):
    arrs = [torch.empty(2 * m, dtype=torch.int32, device="cuda") for _ in range(3)]
    ctx.gen_gnm_device(n, m, max_w, seed, *arrs)
    return arrs


def outs(torch, n):
    return [torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda") for _ in range(3)]


def both(torch, ctx, arrs, n, algorithm, stream=None):
    N = arrs[0].numel()
    a = outs(torch, n)
    ca = ctx.solve_device(*arrs, N, n, *a, n, algorithm=algorithm)
    torch.cuda.synchronize()
    b = outs(torch, n)
    cb = k.solve_device_noalloc(*arrs, N, n, *b, n, algorithm=algorithm, stream=stream)
    torch.cuda.synchronize()
    assert cb == ca
    for x, y in zip(a, b):
        assert torch.equal(x[:ca[0]], y[:ca[0]])
    return ca


@pytest.mark.parametrize("algorithm", [k.BORUVKA, k.FILTER_BORUVKA])
def test_noalloc_equals_solve_device_twice_then_smaller(gpu, ctx, algorithm):
    torch = gpu
    big = gnm(torch, ctx, 1 << 17, 1 << 19, 254, 41)      # 2^20 records: the limit
    small = gnm(torch, ctx, 3000, 20000, 254, 42)
    r1 = both(torch, ctx, big, 1 << 17, algorithm)
    r2 = both(torch, ctx, small, 3000, algorithm)
    assert r1[0] > r2[0] > 0


def test_noalloc_positional_keys_and_a_callers_stream(gpu, ctx):
    torch = gpu
    s = torch.cuda.Stream()
    for n, m, seed in ((1 << 19, 1 << 19, 43), (5000, 40000, 44)):
        arrs = gnm(torch, ctx, n, m, (1 << 31) - 1, seed)  # 31-bit weights: positional keys
        with torch.cuda.stream(s):
            both(torch, ctx, arrs, n, k.FILTER_BORUVKA, stream=s.cuda_stream)


def test_noalloc_allocates_no_device_memory(gpu, ctx):
    torch = gpu
    arrs = gnm(torch, ctx, 1 << 16, 1 << 18, 254, 45)
    o = outs(torch, 1 << 16)
    k.solve_device_noalloc(*arrs, arrs[0].numel(), 1 << 16, *o, 1 << 16)  # set-up done
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for algorithm in (k.BORUVKA, k.FILTER_BORUVKA):
        k.solve_device_noalloc(*arrs, arrs[0].numel(), 1 << 16, *o, 1 << 16, algorithm=algorithm)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert abs(free0 - free1) < (1 << 20), (free0, free1)


def test_noalloc_refuses_sizes_past_the_limit(gpu, ctx):
    torch = gpu
    arrs = gnm(torch, ctx, 1 << 16, (1 << 19) + 1, 254, 46)  # 2^20 + 2 records
    o = outs(torch, 1 << 16)
    with pytest.raises(k.KmstError) as e:
        k.solve_device_noalloc(*arrs, arrs[0].numel(), 1 << 16, *o, 1 << 16)
    assert e.value.status == k.KMST_EINVAL and "limit" in str(e.value)
    small = gnm(torch, ctx, 1000, 4000, 254, 47)
    o = outs(torch, (1 << 19) + 1)
    with pytest.raises(k.KmstError) as e:  # n past 2^19
        k.solve_device_noalloc(*small, small[0].numel(), (1 << 19) + 1, *o, (1 << 19) + 1)
    assert e.value.status == k.KMST_EINVAL
    # and the entry point still works afterwards
    both(torch, ctx, small, 1000, k.BORUVKA)
