"""Pins the CPU oracle (oracle/kmst_oracle.cpp) -- CPU only.

Kruskal is checked against brute-force enumeration of spanning forests and
closed forms; the three restatements of the reference's algorithms
(local filter-Boruvka, hybrid rounds, filter recursion) are checked against
Kruskal, and the oracle against the committed golden fixtures.  A planted
mistake in any of these (dropped tie-break term, wrong root rule, missing
dedup key) changes an edge set that these tests compare exactly."""
import os

import numpy as np
import pytest

import graphs
import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def canon(rec):
    return graphs.canonical(rec)


@pytest.mark.parametrize("seed", range(300))
def test_kruskal_and_restatements_vs_brute_force(seed):
    n = 2 + seed % 6
    m = min(n * (n - 1) // 2, 1 + seed % 10)
    und = graphs.random_undirected(n, m, 1 + seed % 3, seed)
    rec = graphs.records(und)
    if len(rec) == 0:
        return
    bf = graphs.brute_force_msf(n, und)
    k, tk = oracle.kruskal(rec)
    assert canon(k) == bf
    assert tk == sum(w for _, _, w in bf)
    for lp in (0, 1):
        assert canon(oracle.boruvka(rec, lp)[0]) == bf
        for thr in (0, 1, 4):
            assert canon(oracle.filter_boruvka(rec, lp, thr)[0]) == bf


def test_multigraph_brute_force():
    rng = np.random.default_rng(5)
    for seed in range(60):
        n = 3 + seed % 4
        und = graphs.random_undirected(n, 6 + seed % 5, 3, 1000 + seed, allow_multi=True)
        rec = graphs.records(und)
        bf = graphs.brute_force_msf(n, und)
        k, _ = oracle.kruskal(rec)
        # canonical triples of the chosen copies equal the brute-force choice
        assert canon(k) == bf, seed
        for lp in (0, 1):
            assert canon(oracle.boruvka(rec, lp)[0]) == bf
            assert canon(oracle.filter_boruvka(rec, lp, 1)[0]) == bf
    del rng


def test_closed_forms():
    # path: every edge is in the MST
    path = [(i, i + 1, 1 + (7 * i) % 5) for i in range(50)]
    k, t = oracle.kruskal(graphs.records(path))
    assert canon(k) == sorted(path) and t == sum(w for *_, w in path)
    # star: every edge
    star = [(0, i, 3) for i in range(1, 40)]
    assert canon(oracle.kruskal(graphs.records(star))[0]) == star
    # cycle with all-equal weights: drop the largest (min, max) edge
    n = 30
    cyc = [(i, (i + 1) % n, 9) for i in range(n)]
    exp = sorted((min(u, v), max(u, v), w) for u, v, w in cyc)[:-1]
    assert canon(oracle.kruskal(graphs.records(cyc))[0]) == exp
    # complete graph with equal weights: the star at vertex 0
    kn = [(u, v, 1) for u in range(12) for v in range(u + 1, 12)]
    assert canon(oracle.kruskal(graphs.records(kn))[0]) == [(0, v, 1) for v in range(1, 12)]
    # cycle with distinct weights: drop the heaviest
    cyc2 = [(i, (i + 1) % n, 100 - i) for i in range(n)]
    exp2 = sorted((min(u, v), max(u, v), w) for u, v, w in cyc2 if w != 100)
    assert canon(oracle.kruskal(graphs.records(cyc2))[0]) == exp2


def test_forest_invariants_and_isolated_vertices():
    # two components plus isolated ids 100..199 (never mentioned) and 300
    und = [(0, 1, 5), (1, 2, 5), (0, 2, 5), (10, 11, 1), (11, 12, 2), (200, 300, 7)]
    k, t = oracle.kruskal(graphs.records(und))
    assert canon(k) == [(0, 1, 5), (0, 2, 5), (10, 11, 1), (11, 12, 2), (200, 300, 7)]
    for lp in (0, 1):
        assert canon(oracle.boruvka(graphs.records(und), lp)[0]) == canon(k)
        assert canon(oracle.filter_boruvka(graphs.records(und), lp, 0)[0]) == canon(k)


@pytest.mark.parametrize("n,m,max_w,seed", [
    (2000, 30000, 254, 1), (5000, 20000, 3, 2), (50000, 200000, 254, 3),
    (60000, 400000, (1 << 31) - 1, 4), (40000, 45000, 254, 5),
])
def test_restatements_equal_kruskal(n, m, max_w, seed):
    rec = graphs.records(graphs.random_undirected(n, m, max_w, seed))
    k, tk = oracle.kruskal(rec)
    ck = canon(k)
    for lp in (0, 1):
        b, tb = oracle.boruvka(rec, lp)
        assert tb == tk and canon(b) == ck
        for thr in (2, 4, 16):
            f, tf = oracle.filter_boruvka(rec, lp, thr)
            assert tf == tk and canon(f) == ck


def test_pair_records_errors():
    good = graphs.records([(0, 1, 3), (1, 2, 4)])
    uid, n = oracle.pair_records(good)
    assert n == 3
    # twin of each reverse record is the forward record with the same pair
    for i, e in enumerate(good):
        t = good[uid[i]]
        assert (min(e["src"], e["dst"]), max(e["src"], e["dst"]), e["weight"]) == \
            (t["src"], t["dst"], t["weight"]) and t["src"] < t["dst"]
    bad = good.copy()
    bad[[0, 1]] = bad[[1, 0]]
    with pytest.raises(oracle.OracleError) as e:
        oracle.kruskal(bad)
    assert e.value.status == 2
    asym = good.copy()
    asym["weight"][0] = 99
    with pytest.raises(oracle.OracleError) as e:
        oracle.kruskal(asym)
    assert e.value.status == 3
    loop = graphs.records([(0, 1, 3)])
    loop["dst"][0] = 0
    with pytest.raises(oracle.OracleError) as e:
        oracle.kruskal(loop)
    assert e.value.status in (1, 2)


def test_golden_fixtures():
    files = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))
    assert files, "no golden fixtures"
    for f in files:
        z = np.load(os.path.join(GOLDEN, f))
        rec = np.zeros(len(z["src"]), dtype=oracle.WEDGE_DTYPE)
        rec["src"], rec["dst"], rec["weight"] = z["src"], z["dst"], z["weight"]
        k, t = oracle.kruskal(rec)
        assert t == int(z["total_weight"]), f
        assert np.array_equal(k["src"], z["mst_src"]) and np.array_equal(k["dst"], z["mst_dst"]) \
            and np.array_equal(k["weight"], z["mst_weight"]), f
        for lp in (0, 1):
            b, tb = oracle.boruvka(rec, lp)
            assert tb == t and np.array_equal(b, k)
