"""The persistent round kernel (kmst_persist.hip): solves that run some or all
of their Boruvka rounds inside one cooperative launch return the oracle's MST
bit for bit, for both algorithms, packed and positional keys, and graphs
whose trees are deep (a path with increasing weights), wide (a star) or
tied (equal weights, parallel records); the solve statistics mark those
rounds as persistent."""
import dataclasses

import numpy as np
import pytest

import graphs
import kamsta_amd as k
import oracle

pytestmark = pytest.mark.gpu

ALL = 1 << 22   # every round of these graphs fits
OFF = 0
MAX_DEGREE = 32  # rounds with more records per vertex id run as separate launches
SMALL = 1 << 14  # ... unless they are this small


def qualifies(m, n, pm, deduped=False):
    """The library's rule for running the remaining rounds persistently (after
    a dedup sort at most n^2 of the m records are live)."""
    if deduped:
        m = min(m, n * n)
    return bool(pm) and m <= pm and n <= 2 * m + 1024 and (m <= MAX_DEGREE * n or m <= SMALL)


def deduped_before(rounds, i):
    """Round i starts on records a dedup sort marked (the previous round
    ended with more than 8 n^2 records)."""
    prev = rounds[i - 1] if i else None
    return prev is not None and prev["m_next"] > 8 * prev["n_next"] ** 2


def round_qualifies(rounds, i, pm):
    return qualifies(rounds[i]["m"], rounds[i]["n"], pm, deduped_before(rounds, i))


def live_bound(rounds, i):
    r = rounds[i]
    return min(r["m"], r["n"] ** 2) if deduped_before(rounds, i) else r["m"]


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def host(t):
    return t.cpu().numpy().view(np.uint32)


def solve(torch, ctx, rec, n, algo, persistent_max, threshold=4):
    entry = ctx.get_tuning()
    try:
        ctx.set_tuning(dataclasses.replace(entry, persistent_max=persistent_max))
        s, d, w = (dev(torch, rec[f]) for f in ("src", "dst", "weight"))
        cap = max(n, 1)
        os_, od, ow = (torch.empty(cap, dtype=torch.int32, device="cuda") for _ in range(3))
        cnt, tw = ctx.solve_device(s, d, w, len(rec), n, os_, od, ow, cap, algorithm=algo,
                                   config=k.AlgorithmConfig(filter_threshold=threshold))
        got = (host(os_)[:cnt], host(od)[:cnt], host(ow)[:cnt])
        return cnt, tw, got, ctx.stats()
    finally:
        ctx.set_tuning(entry)


def check(torch, ctx, rec, n, thresholds=(4,)):
    exp, exp_w = oracle.kruskal(rec)
    for algo in (k.BORUVKA, k.FILTER_BORUVKA):
        for thr in thresholds:
            for pm in (OFF, ALL):
                cnt, tw, got, st = solve(torch, ctx, rec, n, algo, pm, thr)
                assert tw == exp_w and cnt == len(exp), (algo, thr, pm, tw, exp_w)
                assert np.array_equal(got[0], exp["src"]), (algo, thr, pm)
                assert np.array_equal(got[1], exp["dst"]), (algo, thr, pm)
                assert np.array_equal(got[2], exp["weight"]), (algo, thr, pm)
                rounds = st["rounds"]
                flags = [r["persistent"] for r in rounds]
                if pm == OFF:
                    assert not any(flags)
                elif algo == k.BORUVKA and rounds:
                    # one round loop: persistent from the first qualifying round on
                    first = next((i for i in range(len(rounds))
                                  if round_qualifies(rounds, i, pm)), len(rounds))
                    assert flags == [i >= first for i in range(len(rounds))], flags
    return exp


def test_hand_graphs(gpu, ctx):
    cases = [
        [(i, i + 1, i + 1) for i in range(20000)],                         # path, one deep chain
        [(i, i + 1, 20000 - i) for i in range(20000)],                     # path, the other way
        [(0, i, 3) for i in range(1, 4000)],                               # star (hub)
        [(i, (i + 1) % 500, 9) for i in range(500)],                       # cycle, ties
        [(u, v, 1) for u in range(60) for v in range(u + 1, 60)],          # K60 equal weights
        [(0, 1, 5), (1, 2, 5), (0, 2, 5), (10, 11, 1), (200, 300, 7)],     # forest, isolated ids
    ]
    for und in cases:
        rec = graphs.records(und)
        n = int(rec["src"].max()) + 1
        check(gpu, ctx, rec, n, thresholds=(2,))


@pytest.mark.parametrize("seed", range(8))
def test_small_random_and_multigraphs(gpu, ctx, seed):
    und = graphs.random_undirected(300 + 37 * seed, 2000 + 500 * seed, 1 + 3 * seed, 900 + seed,
                                   allow_multi=seed % 2 == 1)
    rec = graphs.records(und)
    check(gpu, ctx, rec, int(rec["src"].max()) + 1, thresholds=(1, 4))


@pytest.mark.parametrize("n,m,mw,seed", [(1 << 14, 1 << 18, 254, 0x4B414D5354410001),
                                         (70000, 400000, (1 << 31) - 1, 31),   # positional keys
                                         (1 << 12, 1 << 19, 254, 32)])         # dense
def test_gnm(gpu, ctx, n, m, mw, seed):
    s, d, w = k.gen_gnm_host(n, m, mw, seed)
    check(gpu, ctx, k.to_wedges(s, d, w), n, thresholds=(3, 16))


def grid_graph(side, seed):
    """side x side grid, random weights in [1, 254]: contraction keeps it
    sparse (planar), so later rounds stay below the degree limit."""
    rng = np.random.default_rng(seed)
    ids = np.arange(side * side, dtype=np.uint32).reshape(side, side)
    u = np.concatenate([ids[:, :-1].ravel(), ids[:-1, :].ravel()])
    v = np.concatenate([ids[:, 1:].ravel(), ids[1:, :].ravel()])
    w = rng.integers(1, 255, size=len(u), dtype=np.uint32)
    s, d, ww = np.concatenate([u, v]), np.concatenate([v, u]), np.concatenate([w, w])
    order = np.lexsort((d, s))
    return k.to_wedges(s[order], d[order], ww[order])


def test_switch_inside_a_solve(gpu, ctx):
    """Large first rounds run as separate launches, the tail persistently:
    the parent array and the renumbered ids carry over."""
    torch = gpu
    side = 512
    rec = grid_graph(side, 33)
    n = side * side
    exp, exp_w = oracle.kruskal(rec)
    _, _, _, st0 = solve(torch, ctx, rec, n, k.BORUVKA, OFF)
    sizes = [live_bound(st0["rounds"], i) for i in range(1, len(st0["rounds"]))
             if round_qualifies(st0["rounds"], i, ALL)]
    assert sizes, st0["rounds"]
    pm = max(sizes)
    for algo in (k.BORUVKA, k.FILTER_BORUVKA):
        cnt, tw, got, st = solve(torch, ctx, rec, n, algo, pm, 3)
        assert tw == exp_w and cnt == len(exp)
        assert np.array_equal(got[0], exp["src"]) and np.array_equal(got[1], exp["dst"])
        assert np.array_equal(got[2], exp["weight"])
        flags = [r["persistent"] for r in st["rounds"]]
        assert True in flags, flags
        if algo == k.BORUVKA:
            assert flags[0] is False, flags


@pytest.mark.parametrize("algo", [k.BORUVKA, k.FILTER_BORUVKA])
def test_tail_after_dedup_rounds(gpu, ctx, algo):
    """Rounds with more than 8 n^2 records sort by (src, dst) and mark all
    but the lightest of each run of parallel records dead (key all ones); a
    persistent tail that starts on such records must drop the dead ones."""
    torch = gpu
    n, m = 1 << 10, 1 << 17
    s, d, w = k.gen_gnm_host(n, m, 254, 35)
    rec = k.to_wedges(s, d, w)
    exp, exp_w = oracle.kruskal(rec)
    # a filter threshold above the average degree (256) sends the filter
    # straight to its base case (rounds that also track the component of every
    # original vertex) on the whole, dense graph; base cases of a split graph
    # hold about threshold * n records and never reach 8 n^2
    thr = 1 << 12
    # thresholds at the sizes of this solve's own qualifying rounds: the tail
    # starts there, after dense rounds of separate launches
    _, _, _, st0 = solve(torch, ctx, rec, n, algo, OFF, thr)
    rounds = st0["rounds"]
    if algo == k.FILTER_BORUVKA:
        assert [f["base"] for f in st0["filter"]] == [True], st0["filter"]
    assert any(deduped_before(rounds, i) for i in range(len(rounds))), "no dedup round"
    sizes = sorted({live_bound(rounds, i) for i in range(1, len(rounds))
                    if round_qualifies(rounds, i, ALL) and deduped_before(rounds, i)})
    assert sizes, rounds
    for pm in sizes[-3:]:
        cnt, tw, got, st = solve(torch, ctx, rec, n, algo, pm, thr)
        assert tw == exp_w and cnt == len(exp), (pm, tw, exp_w)
        assert np.array_equal(got[0], exp["src"]) and np.array_equal(got[1], exp["dst"])
        assert np.array_equal(got[2], exp["weight"])
        flags = [r["persistent"] for r in st["rounds"]]
        assert False in flags and True in flags, (pm, flags)


def test_persistent_rounds_logged(gpu, ctx):
    torch = gpu
    s, d, w = k.gen_gnm_host(5000, 40000, 254, 34)
    rec = k.to_wedges(s, d, w)
    cnt, _, _, st = solve(torch, ctx, rec, 5000, k.BORUVKA, ALL)
    rounds = st["rounds"]
    assert rounds and all(r["persistent"] for r in rounds)
    assert rounds[0]["m"] == len(rec) and rounds[-1]["m_next"] == 0
    # every MST edge is added by exactly one round
    assert sum(r["n"] - r["n_next"] for r in rounds) == cnt
    for a, b in zip(rounds, rounds[1:]):
        assert a["m_next"] == b["m"] and b["n"] <= a["n_next"]


def test_tuning_field_validated(gpu, ctx):
    entry = ctx.get_tuning()
    with pytest.raises(k.KmstError) as err:
        ctx.set_tuning(dataclasses.replace(entry, persistent_max=1 << 32))
    assert err.value.status == k.KMST_EINVAL
    assert ctx.get_tuning() == entry


def test_round_run_samples(gpu, ctx):
    """Every round that leaves records samples 4096 of them for run starts
    (the sort-skip rule reads them): plain Boruvka on a GNM multigraph keeps
    each original vertex's records adjacent, so runs are long at first."""
    torch = gpu
    n = 1 << 12
    s, d, w = k.gen_gnm_host(n, 32 * n, 254, 36)
    rec = k.to_wedges(s, d, w)
    exp, exp_w = oracle.kruskal(rec)
    cnt, tw, _, st = solve(torch, ctx, rec, n, k.BORUVKA, OFF)
    assert tw == exp_w and cnt == len(exp)
    rounds = st["rounds"]
    assert rounds and not any(r["persistent"] for r in rounds)
    for r in rounds:
        if r["m_next"] == 0:
            assert r["run_starts"] == 0 and r["far_starts"] == 0
            continue
        assert 0 < r["run_starts"] <= 4096, r
        assert r["far_starts"] <= r["run_starts"], r
    # 64 records per vertex in the input: the first relabel keeps runs of
    # about that length, so far fewer than a quarter of the samples start one
    assert rounds[0]["run_starts"] < 1024, rounds[0]
