"""Device generators (kmst_gen_rgg2d_device, kmst_gen_rhg_device) against the
host generators: identical arrays, record for record.  The host generators
are pinned to their definitions by tests/test_generators.py; the full-size
comparison is in tests/test_gpu_full_configs.py."""
import ctypes

import numpy as np
import pytest

import kamsta_amd as k

pytestmark = pytest.mark.gpu


def host_of(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("n,avg,seed", [(1 << 16, 16.0, 3), (1 << 16, 4.0, 21), (50000, 30.0, 8),
                                         (40, 30.0, 5), (2, 8.0, 1)])
def test_rgg2d_device_equals_host(gpu, ctx, n, avg, seed):
    mw = (1 << 31) - 1
    hs, hd, hw = k.gen_rgg2d_host(n, avg, mw, seed)
    ds, dd, dw = ctx.gen_rgg2d_device(n, avg, mw, seed)
    assert np.array_equal(host_of(ds), hs)
    assert np.array_equal(host_of(dd), hd)
    assert np.array_equal(host_of(dw), hw)


@pytest.mark.parametrize("n,avg,gamma,seed", [(1 << 16, 32.0, 2.8, 4), (1 << 16, 8.0, 2.2, 9),
                                               (30000, 16.0, 3.5, 17), (1000, 32.0, 2.8, 2)])
def test_rhg_device_equals_host(gpu, ctx, n, avg, gamma, seed):
    hs, hd, hw = k.gen_rhg_host(n, avg, gamma, 254, seed)
    ds, dd, dw = ctx.gen_rhg_device(n, avg, gamma, 254, seed)
    assert len(ds) == len(hs)
    assert np.array_equal(host_of(ds), hs)
    assert np.array_equal(host_of(dd), hd)
    assert np.array_equal(host_of(dw), hw)


def test_device_generator_capacity_rule(gpu, ctx):
    L = k.load()
    m = ctypes.c_uint64(0)
    st = L.kmst_gen_rgg2d_device(ctx._h, 5000, 16.0, 254, 3, None, None, None, 0, ctypes.byref(m))
    assert st == k.KMST_EINVAL
    hs, _, _ = k.gen_rgg2d_host(5000, 16.0, 254, 3)
    assert m.value == len(hs)
    st = L.kmst_gen_rhg_device(ctx._h, 5000, 32.0, 2.8, 254, 4, None, None, None, 0,
                               ctypes.byref(m))
    assert st == k.KMST_EINVAL
    hs, _, _ = k.gen_rhg_host(5000, 32.0, 2.8, 254, 4)
    assert m.value == len(hs)
    # bad arguments are refused before any device work
    assert L.kmst_gen_rhg_device(ctx._h, 5000, 32.0, 1.5, 254, 4, None, None, None, 0,
                                 ctypes.byref(m)) == k.KMST_EINVAL


@pytest.mark.parametrize("gen", ["rgg2d", "rhg"])
def test_device_generated_graphs_solve_like_the_oracle(gpu, ctx, gen):
    import oracle
    torch = gpu
    if gen == "rgg2d":
        s, d, w = ctx.gen_rgg2d_device(1 << 15, 16.0, (1 << 31) - 1, 31)
    else:
        s, d, w = ctx.gen_rhg_device(1 << 15, 32.0, 2.8, 254, 32)
    n = 1 << 15
    outs = [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3)]
    cnt, tw = ctx.solve_device(s, d, w, s.numel(), n, *outs, n, algorithm=k.FILTER_BORUVKA)
    torch.cuda.synchronize()
    exp, ew = oracle.kruskal(oracle.wedges_from_soa(host_of(s), host_of(d), host_of(w)))
    assert tw == ew and cnt == len(exp)
    assert np.array_equal(host_of(outs[0][:cnt]), exp["src"].astype(np.uint32))
    assert np.array_equal(host_of(outs[1][:cnt]), exp["dst"].astype(np.uint32))
