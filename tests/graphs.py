"""Small-graph helpers for the tests: record construction, random graphs and
a brute-force minimum spanning forest used to pin the oracle."""
import itertools

import numpy as np

WEDGE_DTYPE = np.dtype([("src", "<u8"), ("dst", "<u8"), ("weight", "<u4"), ("_pad", "<u4")])


def records(undirected):
    """Symmetric directed records sorted by (src, dst) from (u, v, w) triples,
    the input contract of hybridMST::boruvka (every edge in both directions)."""
    rows = []
    for u, v, w in undirected:
        assert u != v
        rows.append((u, v, w))
        rows.append((v, u, w))
    rows.sort(key=lambda r: (r[0], r[1], r[2]))
    out = np.zeros(len(rows), dtype=WEDGE_DTYPE)
    if rows:
        arr = np.array(rows, dtype=np.uint64)
        out["src"], out["dst"], out["weight"] = arr[:, 0], arr[:, 1], arr[:, 2].astype(np.uint32)
    return out


def random_undirected(n, m, max_w, seed, allow_multi=False):
    """m distinct (or, with allow_multi, possibly repeated) pairs on n vertices."""
    rng = np.random.default_rng(seed)
    seen = set()
    out = []
    guard = 0
    while len(out) < m and guard < 50 * m + 100:
        guard += 1
        u, v = (int(x) for x in rng.integers(0, n, size=2))
        if u == v:
            continue
        a, b = min(u, v), max(u, v)
        if not allow_multi and (a, b) in seen:
            continue
        seen.add((a, b))
        out.append((a, b, int(rng.integers(1, max_w + 1))))
    return out


def canonical(edges):
    """Canonical (min, max, w) triples sorted by (min, max) of an edge array."""
    s = np.minimum(edges["src"], edges["dst"])
    d = np.maximum(edges["src"], edges["dst"])
    order = np.lexsort((edges["weight"], d, s))
    return [(int(s[i]), int(d[i]), int(edges["weight"][i])) for i in order]


def brute_force_msf(n, undirected):
    """Minimum spanning forest by enumeration of edge subsets (tiny inputs).

    Minimises sum(w) * 2^32 + sum(rank) over all spanning forests, where rank
    is the position of the edge in (min, max, w) order.  Since the rank term
    never exceeds 2^32, this picks the minimum-weight forest and, among those,
    the one that is minimal under the strict (w, min, max) order, independent
    of any greedy argument."""
    edges = sorted((min(u, v), max(u, v), w) for u, v, w in undirected)
    # components of the whole graph decide the forest size
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x

    for a, b, _ in edges:
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[ra] = rb
    size = n - len({find(v) for v in range(n)})
    best, best_set = None, None
    for subset in itertools.combinations(range(len(edges)), size):
        p = list(range(n))

        def f(x):
            while p[x] != x:
                x = p[x]
            return x

        ok = True
        for i in subset:
            a, b, _ = edges[i]
            ra, rb = f(a), f(b)
            if ra == rb:
                ok = False
                break
            p[ra] = rb
        if not ok:
            continue
        cost = sum(edges[i][2] for i in subset) * (1 << 32) + sum(subset)
        if best is None or cost < best:
            best, best_set = cost, subset
    return sorted(edges[i] for i in best_set)
