"""Tunings and the autotuner (kmst_context_set_tuning / kmst_autotune): every
launch shape returns the oracle's MST; the autotuner checks its candidates
against each other and keeps the fastest on the context."""
import numpy as np
import pytest

import kamsta_amd as k
import oracle

pytestmark = pytest.mark.gpu


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


@pytest.fixture(scope="module")
def gnm(gpu, ctx):
    """GNM n=2^16, m=2^20: 2^21 records, so sorts of both tile widths run."""
    torch = gpu
    n, m = 1 << 16, 1 << 20
    s = torch.empty(2 * m, dtype=torch.int32, device="cuda")
    d, w = torch.empty_like(s), torch.empty_like(s)
    ctx.gen_gnm_device(n, m, 254, 91, s, d, w)
    rec = k.to_wedges(host(s, np.uint32), host(d, np.uint32), host(w, np.uint32))
    exp, exp_w = oracle.kruskal(rec)
    return n, m, s, d, w, exp, exp_w


def solve_and_check(torch, ctx, gnm, algo, threshold=4):
    n, m, s, d, w, exp, exp_w = gnm
    os_, od, ow = (torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3))
    cnt, tw = ctx.solve_device(s, d, w, 2 * m, n, os_, od, ow, n, algorithm=algo,
                               config=k.AlgorithmConfig(filter_threshold=threshold))
    assert tw == exp_w and cnt == len(exp)
    assert np.array_equal(host(os_, np.uint32)[:cnt], exp["src"])
    assert np.array_equal(host(od, np.uint32)[:cnt], exp["dst"])
    assert np.array_equal(host(ow, np.uint32)[:cnt], exp["weight"])


@pytest.mark.parametrize("threads", [0, 256, 512])
@pytest.mark.parametrize("balanced", [0, 1])
def test_every_tuning_gives_the_oracle_mst(gpu, ctx, gnm, threads, balanced):
    entry = ctx.get_tuning()
    try:
        # wide_min 2^12: the by-size rule switches inside this solve's sorts
        ctx.set_tuning(k.Tuning(threads, balanced, 1 << 12))
        assert ctx.get_tuning() == k.Tuning(threads, balanced, 1 << 12)
        for algo in (k.BORUVKA, k.FILTER_BORUVKA):
            solve_and_check(gpu, ctx, gnm, algo)
    finally:
        ctx.set_tuning(entry)


def test_invalid_tunings_are_rejected(gpu, ctx):
    for bad in (k.Tuning(128, 1, 1 << 20), k.Tuning(0, 2, 1 << 20), k.Tuning(0, 1, 0)):
        with pytest.raises(k.KmstError) as err:
            ctx.set_tuning(bad)
        assert err.value.status == k.KMST_EINVAL


def test_autotune_picks_a_checked_tuning(gpu, ctx, gnm):
    n, m, s, d, w, exp, exp_w = gnm
    entry = ctx.get_tuning()
    try:
        res = ctx.autotune(s, d, w, 2 * m, n, algorithm=k.FILTER_BORUVKA, repeats=1)
        # default + 3 other thresholds + 2 other sort widths + the other digit plan
        # + 3 other persistent-round limits
        assert res["candidates"] == 10
        assert res["filter_threshold"] in (2, 3, 4, 6)
        assert res["tuning"].sort_threads in (0, 256, 512)
        assert res["best_ms"] > 0
        assert ctx.get_tuning() == res["tuning"]
        solve_and_check(gpu, ctx, gnm, k.FILTER_BORUVKA, res["filter_threshold"])
    finally:
        ctx.set_tuning(entry)
