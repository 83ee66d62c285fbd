"""MST output order (K9): the keys are sorted by their group only -- the
smaller endpoint of a packed key, the twin position / 64 of a positional one --
and one pass orders every group while it decodes: groups of at most 64 keys by
a thread per key, longer ones (hubs) by a workgroup sort, and a group longer
than 8192 keys (a star) by a sort of the whole order field.  Every path returns
the oracle's MST in canonical (min, max) order, bit for bit, for both
algorithms; the solve statistics show which paths ran."""
import numpy as np
import pytest

import kamsta_amd as k
import oracle

pytestmark = pytest.mark.gpu

GROUP_SMALL = 64
GROUP_MAX = 8 * 1024


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def host(t):
    return t.cpu().numpy().view(np.uint32)


def symmetric(u, v, w):
    """Directed records in both directions, sorted by (src, dst)."""
    s = np.concatenate([u, v]).astype(np.uint64)
    d = np.concatenate([v, u]).astype(np.uint64)
    ww = np.concatenate([w, w]).astype(np.uint32)
    order = np.lexsort((ww, d, s))
    return k.to_wedges(s[order], d[order], ww[order])


def star(leaves, center_first, max_w, seed):
    """A star on leaves + 1 vertices; its MST is the star itself."""
    rng = np.random.default_rng(seed)
    n = leaves + 1
    center = 0 if center_first else n - 1
    others = np.delete(np.arange(n, dtype=np.uint64), center)
    u = np.full(leaves, center, dtype=np.uint64)
    w = rng.integers(1, max_w + 1, size=leaves, dtype=np.uint64)
    return symmetric(u, others, w), n


def run(torch, ctx, rec, n, algo):
    s, d, w = (dev(torch, rec[f]) for f in ("src", "dst", "weight"))
    cap = max(n, 1)
    os_, od, ow = (torch.empty(cap, dtype=torch.int32, device="cuda") for _ in range(3))
    cnt, tw = ctx.solve_device(s, d, w, len(rec), n, os_, od, ow, cap, algorithm=algo,
                               config=k.AlgorithmConfig(filter_threshold=4))
    return cnt, tw, (host(os_)[:cnt], host(od)[:cnt], host(ow)[:cnt]), ctx.stats()


def check(torch, ctx, rec, n):
    exp, exp_w = oracle.kruskal(rec)
    stats = []
    for algo in (k.BORUVKA, k.FILTER_BORUVKA):
        cnt, tw, got, st = run(torch, ctx, rec, n, algo)
        assert tw == exp_w and cnt == len(exp), (algo, tw, exp_w, cnt, len(exp))
        assert np.array_equal(got[0], exp["src"]), algo
        assert np.array_equal(got[1], exp["dst"]), algo
        assert np.array_equal(got[2], exp["weight"]), algo
        stats.append(st)
    return exp, stats


@pytest.mark.parametrize("leaves", [1, 2, 63, 64, 65, 1023, 1024, 1025, 4097, GROUP_MAX,
                                    GROUP_MAX + 1, 40000])
def test_star_center_first(gpu, ctx, leaves):
    """All MST edges share the smaller endpoint 0: one group of `leaves` keys."""
    rec, n = star(leaves, True, 254, 70 + leaves)
    _, stats = check(gpu, ctx, rec, n)
    for st in stats:
        assert st["packed_keys"]
        assert st["mst_long_groups"] == (1 if leaves > GROUP_SMALL else 0), st["mst_long_groups"]
        assert st["mst_resorted"] == (leaves > GROUP_MAX)


@pytest.mark.parametrize("leaves", [65, 5000])
def test_star_center_last(gpu, ctx, leaves):
    """The same star with the center as the largest id: groups of one key."""
    rec, n = star(leaves, False, 254, 80 + leaves)
    _, stats = check(gpu, ctx, rec, n)
    for st in stats:
        assert st["mst_long_groups"] == 0 and not st["mst_resorted"]


def test_hubs_in_a_random_graph(gpu, ctx):
    """Hubs of 70 to 3000 leaves among random edges: long groups next to short
    ones, at the group boundaries of the sorted keys."""
    rng = np.random.default_rng(90)
    n = 20000
    hubs = {3: 70, 10: 500, 11: 1500, 4000: 3000}
    us, vs = [], []
    taken = np.zeros(n, dtype=bool)
    taken[list(hubs)] = True
    for h, cnt in hubs.items():
        pool = np.flatnonzero(~taken & (np.arange(n) > h))
        leaves = rng.choice(pool, size=cnt, replace=False)
        taken[leaves] = True
        us.append(np.full(cnt, h))
        vs.append(leaves)
    # a sparse random background on all vertices, heavier than the hub edges
    bu = rng.integers(0, n, size=3 * n)
    bv = rng.integers(0, n, size=3 * n)
    keep = bu != bv
    us.append(np.minimum(bu, bv)[keep])
    vs.append(np.maximum(bu, bv)[keep])
    u = np.concatenate(us).astype(np.uint64)
    v = np.concatenate(vs).astype(np.uint64)
    pairs = np.unique(np.stack([np.minimum(u, v), np.maximum(u, v)], axis=1), axis=0)
    w = rng.integers(1, 255, size=len(pairs), dtype=np.uint64)
    hub_edge = np.isin(pairs[:, 0], list(hubs))
    w[hub_edge] = rng.integers(1, 20, size=int(hub_edge.sum()), dtype=np.uint64)
    rec = symmetric(pairs[:, 0], pairs[:, 1], w)
    exp, stats = check(gpu, ctx, rec, n)
    groups = np.bincount(exp["src"].astype(np.int64), minlength=n)
    long_groups = int((groups > GROUP_SMALL).sum())
    assert long_groups >= 3
    for st in stats:
        assert st["mst_long_groups"] == long_groups and not st["mst_resorted"]


def test_positional_star(gpu, ctx):
    """Weights of 31 bits on 2^17 vertices do not fit a packed key: groups are
    ranges of 64 twin positions, never longer than 64 keys."""
    rec, n = star(1 << 17, True, (1 << 31) - 1, 91)
    _, stats = check(gpu, ctx, rec, n)
    for st in stats:
        assert not st["packed_keys"]
        assert st["mst_long_groups"] == 0 and not st["mst_resorted"]


def test_positional_geometric(gpu, ctx):
    """31-bit weights on 2^17 vertices: positional keys (31 + 2 * 17 > 64 bits)."""
    s, d, w = k.gen_rgg2d_host(1 << 17, seed=92)
    rec = k.to_wedges(s, d, w)
    _, stats = check(gpu, ctx, rec, 1 << 17)
    for st in stats:
        assert not st["packed_keys"] and not st["mst_resorted"]


def test_hyperbolic_hubs(gpu, ctx):
    """A small hyperbolic graph: power-law hubs give groups past 64 keys."""
    n = 1 << 17
    s, d, w = k.gen_rhg_host(n, seed=93)
    rec = k.to_wedges(s, d, w)
    exp, stats = check(gpu, ctx, rec, n)
    groups = np.bincount(exp["src"].astype(np.int64), minlength=n)
    long_groups = int((groups > GROUP_SMALL).sum())
    assert long_groups > 0
    for st in stats:
        assert st["mst_long_groups"] == long_groups and not st["mst_resorted"]
