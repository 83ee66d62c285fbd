"""The geometric generators against their definitions, by brute force over
all pairs (CPU).

RGG2D is integer arithmetic, so the host generator must equal the brute force
exactly.  RHG decides each pair in double precision with the library's own
elementary functions; the brute force here uses numpy's, so the two may
disagree only on pairs whose two sides of the test lie within a relative
1e-9 of each other (none occur at these sizes).  Both checks also show that
the candidate windows (cells, angular bands) miss no edge.
"""
import math

import numpy as np
import pytest

import kamsta_amd as k

U64 = np.uint64


def mix64(x):
    x = x.astype(U64) + U64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> U64(30))) * U64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> U64(27))) * U64(0x94D049BB133111EB)
    return x ^ (x >> U64(31))


def unit(h):
    return (h >> U64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def geo_weights(s, d, seed, max_w):
    wseed = mix64(np.array([seed ^ 0x5745494748543A32], dtype=U64))[0]
    a, b = np.minimum(s, d).astype(U64), np.maximum(s, d).astype(U64)
    return (U64(1) + mix64(wseed ^ ((a << U64(32)) | b)) % U64(max_w)).astype(np.uint32)


def records_from_pairs(pairs, n):
    """Sorted symmetric records from an (i < j) boolean upper triangle."""
    i, j = np.nonzero(pairs)
    s = np.concatenate([i, j]).astype(np.uint32)
    d = np.concatenate([j, i]).astype(np.uint32)
    o = np.lexsort((d, s))
    return s[o], d[o]


def rgg_reference(n, avg, seed, max_w):
    r = math.sqrt(avg / (math.pi * n))
    rr = int(r * 4294967296.0)
    cells = min(max(1, int(1.0 / r)), 1 << 15)
    h = mix64(mix64(np.array([seed], dtype=U64))[0] + np.arange(n, dtype=U64))
    x, y = (h >> U64(32)).astype(np.uint64), (h & U64(0xFFFFFFFF)).astype(np.uint64)
    cell = ((y * U64(cells)) >> U64(32)) * U64(cells) + ((x * U64(cells)) >> U64(32))
    order = np.argsort(cell, kind="stable")
    x, y = x[order].astype(np.int64), y[order].astype(np.int64)
    dx = np.abs(x[:, None] - x[None, :])
    dy = np.abs(y[:, None] - y[None, :])
    near = np.triu((dx <= rr) & (dy <= rr), 1)
    adj = np.zeros_like(near)
    for a, b in zip(*np.nonzero(near)):  # exact integer test (sums reach 2^65)
        adj[a, b] = int(dx[a, b]) ** 2 + int(dy[a, b]) ** 2 <= rr * rr
    s, d = records_from_pairs(adj, n)
    return s, d, geo_weights(s, d, seed, max_w)


@pytest.mark.parametrize("n,avg,seed", [(3000, 16.0, 3), (2500, 6.0, 11), (40, 30.0, 5),
                                         (2, 8.0, 1)])
def test_rgg2d_host_is_the_definition(n, avg, seed):
    mw = (1 << 31) - 1
    s, d, w = k.gen_rgg2d_host(n, avg, mw, seed)
    es, ed, ew = rgg_reference(n, avg, seed, mw)
    assert np.array_equal(s, es) and np.array_equal(d, ed) and np.array_equal(w, ew)


def rhg_reference(n, avg, gamma, seed):
    alpha = (gamma - 1.0) / 2.0
    xi = alpha / (alpha - 0.5)
    R = 2.0 * math.log((2.0 / math.pi) * xi * xi * n / avg)
    coshR = math.cosh(R)
    cr = math.cosh(alpha * R) - 1.0
    sm = mix64(np.array([seed ^ 0x524847], dtype=U64))[0]
    i = np.arange(n, dtype=U64)
    th = 2.0 * math.pi * unit(mix64(sm + U64(2) * i))
    r = np.arccosh(1.0 + unit(mix64(sm + U64(2) * i + U64(1))) * cr) / alpha
    o = np.lexsort((r, th))
    th, r = th[o], r[o]
    sh = np.sinh(r)
    h = np.sin(0.5 * (th[:, None] - th[None, :]))
    lhs = 2.0 * h * h * sh[:, None] * sh[None, :]
    rhs = coshR - np.cosh(r[:, None] - r[None, :])
    return np.triu(lhs <= rhs, 1), np.triu(np.abs(lhs - rhs) <= 1e-9 * coshR, 1)


@pytest.mark.parametrize("n,avg,gamma,seed", [(2000, 32.0, 2.8, 4), (1500, 8.0, 2.2, 9),
                                               (1200, 16.0, 3.5, 17)])
def test_rhg_host_is_the_definition(n, avg, gamma, seed):
    s, d, w = k.gen_rhg_host(n, avg, gamma, 254, seed)
    adj, close = rhg_reference(n, avg, gamma, seed)
    got = np.zeros((n, n), dtype=bool)
    got[s, d] = True
    assert np.array_equal(got, got.T)
    got = np.triu(got, 1)
    differ = got != adj
    assert not np.any(differ & ~close), np.argwhere(differ & ~close)[:5]
    assert differ.sum() == 0  # no pair lies that close to the threshold here
    assert np.array_equal(w, geo_weights(s, d, seed, 254))
    assert np.bincount(s, minlength=n).max() > 4 * (len(s) / n)  # heavy-tailed degrees
