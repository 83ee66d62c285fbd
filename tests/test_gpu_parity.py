"""GPU parity: libkmst (through the C ABI) against the CPU oracle and numpy
references, bit-exact (integer work)."""
import time

import numpy as np
import pytest

import graphs
import kamsta_amd as k
import oracle

pytestmark = pytest.mark.gpu


def dev(torch, arr, dtype):
    a = np.ascontiguousarray(arr)
    if dtype == np.uint32:
        return torch.from_numpy(a.astype(np.uint32).view(np.int32)).cuda()
    return torch.from_numpy(a.astype(np.uint64).view(np.int64)).cuda()


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


def strip(rec):
    return rec[["src", "dst", "weight"]]


def check_equal(rec, threshold=4):
    exp, exp_w = oracle.kruskal(rec)
    for fn in (k.boruvka, k.filter_boruvka):
        got, got_w = fn(rec, k.AlgorithmConfig(filter_threshold=threshold))
        assert got_w == exp_w, (fn.__name__, got_w, exp_w)
        assert np.array_equal(strip(got), strip(exp)), fn.__name__
    return exp, exp_w


# ------------------------------------------------------------------ kernels
@pytest.mark.parametrize("case", ["uniform", "hubs", "single", "sparse_ids"])
def test_segmented_argmin_kernel(gpu, ctx, case):
    torch = gpu
    rng = np.random.default_rng(11)
    if case == "uniform":
        n, deg = 5000, rng.integers(1, 40, 5000)
    elif case == "hubs":
        n = 300
        deg = rng.integers(1, 5, n)
        deg[[3, 150, 299]] = [70000, 5000, 9000]  # cross many 2048-record tiles
    elif case == "single":
        n, deg = 1, np.array([100000])
    else:
        n = 20000
        deg = np.where(rng.random(n) < 0.3, rng.integers(1, 9, n), 0)
    src = np.repeat(np.arange(n, dtype=np.uint32), deg)
    m = len(src)
    dst = rng.integers(0, n, m).astype(np.uint32)
    val = rng.permutation(m).astype(np.uint64) * 7 + 3  # unique keys
    val[rng.random(m) < 0.05] = np.uint64(2**64 - 1)  # some dead records
    bv = dev(torch, np.zeros(n, np.uint64), np.uint64)
    bd = dev(torch, np.zeros(n, np.uint32), np.uint32)
    ctx.segmented_argmin(dev(torch, src, np.uint32), dev(torch, dst, np.uint32),
                         dev(torch, val, np.uint64), m, n, bv, bd)
    torch.cuda.synchronize()
    gv, gd = host(bv, np.uint64), host(bd, np.uint32)
    # numpy reference
    exp_v = np.full(n, np.uint64(2**64 - 1))
    exp_d = np.zeros(n, np.uint32)
    order = np.lexsort((val, src))
    first = np.ones(m, bool)
    first[1:] = src[order][1:] != src[order][:-1]
    idx = order[first]
    exp_v[src[idx]] = val[idx]
    exp_d[src[idx]] = dst[idx]
    assert np.array_equal(gv, exp_v)
    live = exp_v != np.uint64(2**64 - 1)
    assert np.array_equal(gd[live], exp_d[live])


@pytest.mark.parametrize("case", ["uniform", "hubs", "sparse", "grouped"])
def test_segmented_argmin_packed_kernel(gpu, ctx, case):
    """Packed keys made from the weights: tiles whose sources lie within
    4096 ids of their first stage the minima in LDS and unpack best_dst from
    the key; sparse ids and groups out of order take the direct stores."""
    torch = gpu
    rng = np.random.default_rng(12)
    if case == "uniform":  # 0-39 records per vertex, empty vertices among them
        n = 20000
        deg = rng.integers(0, 40, n)
        order = np.arange(n)
    elif case == "hubs":  # runs across many 4096-record tiles
        n = 300
        deg = rng.integers(1, 5, n)
        deg[[0, 150, 299]] = [70000, 5000, 9000]
        order = np.arange(n)
    elif case == "sparse":  # sources far apart: a tile spans more than 4096 ids
        n = 1 << 22
        deg = np.zeros(n, np.int64)
        ids = rng.choice(n, 30000, replace=False)
        deg[ids] = rng.integers(1, 9, len(ids))
        order = np.arange(n)
    else:  # grouped by source, groups in random order
        n = 20000
        deg = rng.integers(1, 12, n)
        order = rng.permutation(n)
    src = np.repeat(order.astype(np.uint32), deg[order])
    m = len(src)
    dst = rng.integers(0, n, m).astype(np.uint32)
    vbits = max(1, int(n - 1).bit_length())
    w = rng.integers(0, 1 << 19, m).astype(np.uint32)  # no key is all ones
    assert 2 * vbits + 20 <= 64
    lo = np.minimum(src, dst).astype(np.uint64)
    hi = np.maximum(src, dst).astype(np.uint64)
    key = (w.astype(np.uint64) << np.uint64(2 * vbits)) | (lo << np.uint64(vbits)) | hi
    bv = dev(torch, np.zeros(n, np.uint64), np.uint64)
    bd = dev(torch, np.zeros(n, np.uint32), np.uint32)
    ctx.segmented_argmin_packed(dev(torch, src, np.uint32), dev(torch, dst, np.uint32),
                                dev(torch, w, np.uint32), vbits, m, n, bv, bd)
    torch.cuda.synchronize()
    gv, gd = host(bv, np.uint64), host(bd, np.uint32)
    exp_v = np.full(n, np.uint64(2**64 - 1))
    exp_d = np.zeros(n, np.uint32)
    srt = np.lexsort((key, src))
    first = np.ones(m, bool)
    first[1:] = src[srt][1:] != src[srt][:-1]
    idx = srt[first]
    exp_v[src[idx]] = key[idx]
    exp_d[src[idx]] = dst[idx]
    assert np.array_equal(gv, exp_v)
    live = exp_v != np.uint64(2**64 - 1)
    assert np.array_equal(gd[live], exp_d[live])


def test_segmented_argmin_packed_rejects_bad_vbits(gpu, ctx):
    torch = gpu
    z32 = dev(torch, np.zeros(4, np.uint32), np.uint32)
    z64 = dev(torch, np.zeros(4, np.uint64), np.uint64)
    with pytest.raises(k.KmstError):
        ctx.segmented_argmin_packed(z32, z32, z32, 0, 4, 4, z64, z32)
    with pytest.raises(k.KmstError):
        ctx.segmented_argmin_packed(z32, z32, z32, 1, 4, 4, z64, z32)  # n > 2^vbits


@pytest.mark.parametrize("m,bits", [(1, 3), (5000, 4), (100003, 17), (300000, 24), (70000, 32)])
def test_radix_sort_kernel_stable(gpu, ctx, m, bits):
    torch = gpu
    rng = np.random.default_rng(m)
    hi = (1 << bits) if bits < 32 else (1 << 32)
    s = rng.integers(0, min(hi, 50), m).astype(np.uint32)  # many equal keys: stability matters
    d = rng.integers(0, hi, m).astype(np.uint32) if bits > 8 else rng.integers(0, hi, m).astype(np.uint32)
    v = np.arange(m, dtype=np.uint64)
    ts, td, tv = dev(torch, s, np.uint32), dev(torch, d, np.uint32), dev(torch, v, np.uint64)
    ctx.sort_pairs(ts, td, tv, m, bits)
    torch.cuda.synchronize()
    order = np.lexsort((v, d, s))
    assert np.array_equal(host(ts, np.uint32), s[order])
    assert np.array_equal(host(td, np.uint32), d[order])
    assert np.array_equal(host(tv, np.uint64), v[order])


def test_pointer_jump_kernel(gpu, ctx):
    torch = gpu
    rng = np.random.default_rng(3)
    n = 200000
    # random forest: parent[v] < v or v (roots), plus long chains
    parent = np.array([v if (v == 0 or rng.random() < 0.01) else rng.integers(max(0, v - 50), v)
                       for v in range(n)], dtype=np.uint32)
    chain = np.arange(50000, 60000, dtype=np.uint32)
    parent[chain[1:]] = chain[:-1]
    p = dev(torch, parent, np.uint32)
    ctx.pointer_jump(p, n)
    got = host(p, np.uint32)
    exp = parent.copy()
    for v in range(n):
        r = exp[v]
        while exp[r] != r:
            r = exp[r]
        exp[v] = r
    assert np.array_equal(got, exp)


# ------------------------------------------------------------------ end to end
@pytest.mark.parametrize("seed", range(40))
def test_small_random_graphs(gpu, seed):
    n = 2 + seed % 9
    m = min(n * (n - 1) // 2, 1 + seed % 12)
    und = graphs.random_undirected(n, m, 1 + seed % 3, seed)
    rec = graphs.records(und)
    if len(rec) == 0:
        return
    exp, _ = check_equal(rec, threshold=seed % 3)
    assert graphs.canonical(exp) == graphs.brute_force_msf(n, und)


def test_hand_graphs(gpu):
    cases = [
        [(i, i + 1, 1 + (7 * i) % 5) for i in range(300)],                 # path
        [(0, i, 3) for i in range(1, 400)],                                # star (hub)
        [(i, (i + 1) % 500, 9) for i in range(500)],                       # cycle, ties
        [(u, v, 1) for u in range(60) for v in range(u + 1, 60)],          # K60 equal weights
        [(0, 1, 5), (1, 2, 5), (0, 2, 5), (10, 11, 1), (200, 300, 7)],     # forest, isolated ids
    ]
    for und in cases:
        check_equal(graphs.records(und), threshold=2)


def test_multigraph(gpu):
    for seed in range(10):
        und = graphs.random_undirected(40, 300, 4, 500 + seed, allow_multi=True)
        check_equal(graphs.records(und), threshold=1)


def test_positional_multigraph(gpu):
    """Parallel copies under positional keys (31-bit weights, n > 2^16), many
    of equal weight: every copy of a pair, in both directions, keys the
    position of the pair's first record (equal-weight copies tie, as with
    packed keys), in the key pass (plain Boruvka) and in the keys that
    filter-Boruvka's class scatter and outside gather make for the records
    they write."""
    rng = np.random.default_rng(41)
    n, m = 70000, 300000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    u, v = u[keep], v[keep]
    w = (1 << 30) + rng.integers(0, 1 << 12, len(u))
    dup = rng.random(len(u)) < 0.3
    copies = rng.integers(1, 4, int(dup.sum()))
    du = np.repeat(u[dup], copies)
    dv = np.repeat(v[dup], copies)
    same = rng.random(len(du)) < 0.5
    dw = np.where(same, np.repeat(w[dup], copies), (1 << 30) + rng.integers(0, 1 << 12, len(du)))
    uu, vv, ww = np.concatenate([u, du]), np.concatenate([v, dv]), np.concatenate([w, dw])
    src = np.concatenate([uu, vv])
    dst = np.concatenate([vv, uu])
    wt = np.concatenate([ww, ww])
    order = np.lexsort((wt, dst, src))
    rec = np.zeros(len(src), dtype=graphs.WEDGE_DTYPE)
    rec["src"], rec["dst"], rec["weight"] = src[order], dst[order], wt[order]
    runs = (rec["src"][1:] == rec["src"][:-1]) & (rec["dst"][1:] == rec["dst"][:-1])
    assert runs.sum() > 50000  # many parallel records
    for thr in (2, 4):
        check_equal(rec, threshold=thr)


@pytest.mark.parametrize("n,m,mw,seed", [(1 << 14, 1 << 18, 254, 0x4B414D5354410001),
                                         (1 << 16, 1 << 20, 254, 5), (100000, 150000, 254, 6),
                                         (1 << 12, 1 << 20, 254, 7), (70000, 900000, (1 << 31) - 1, 8)])
def test_gnm_vs_oracle(gpu, n, m, mw, seed):
    s, d, w = k.gen_gnm_host(n, m, mw, seed)
    rec = k.to_wedges(s, d, w)
    for thr in (4, 16):
        check_equal(rec, threshold=thr)


def test_golden_fixtures_gpu(gpu):
    import os
    gdir = os.path.join(os.path.dirname(__file__), "golden")
    for f in sorted(os.listdir(gdir)):
        if not f.endswith(".npz"):
            continue
        z = np.load(os.path.join(gdir, f))
        rec = np.zeros(len(z["src"]), dtype=k.WEDGE_DTYPE)
        rec["src"], rec["dst"], rec["weight"] = z["src"], z["dst"], z["weight"]
        for fn in (k.boruvka, k.filter_boruvka):
            got, tw = fn(rec, k.AlgorithmConfig(filter_threshold=1))
            assert tw == int(z["total_weight"]), f
            assert np.array_equal(got["src"], z["mst_src"]) and np.array_equal(got["dst"], z["mst_dst"])
            assert np.array_equal(got["weight"], z["mst_weight"])


def test_contract_errors(gpu):
    good = graphs.records([(0, 1, 3), (1, 2, 4), (0, 2, 9)])
    bad = good.copy()
    bad[[0, 1]] = bad[[1, 0]]
    with pytest.raises(k.KmstError) as e:
        k.boruvka(bad)
    assert e.value.status == k.KMST_EUNSORTED
    asym = good.copy()
    asym["weight"][0] = 99
    with pytest.raises(k.KmstError) as e:
        k.filter_boruvka(asym)
    assert e.value.status == k.KMST_EASYMMETRIC
    loop = good.copy()
    loop["dst"][0] = 0
    with pytest.raises(k.KmstError) as e:
        k.boruvka(loop)
    assert e.value.status in (k.KMST_EINVAL, k.KMST_EUNSORTED)
    odd = good[:-1].copy()
    with pytest.raises(k.KmstError) as e:
        k.boruvka(odd)
    assert e.value.status == k.KMST_EASYMMETRIC
    # the library still works after errors
    check_equal(good)
    # empty input
    got, tw = k.boruvka(np.zeros(0, dtype=k.WEDGE_DTYPE))
    assert len(got) == 0 and tw == 0


def test_device_generator_matches_host(gpu, ctx):
    torch = gpu
    for n, m, mw, seed in [(1 << 14, 1 << 18, 254, 1), (256, 4000, 254, 2), (1 << 20, 1 << 22, 254, 3)]:
        s = torch.empty(2 * m, dtype=torch.int32, device="cuda")
        d = torch.empty_like(s)
        w = torch.empty_like(s)
        ctx.gen_gnm_device(n, m, mw, seed, s, d, w)
        hs, hd, hw = k.gen_gnm_host(n, m, mw, seed)
        assert np.array_equal(host(s, np.uint32), hs)
        assert np.array_equal(host(d, np.uint32), hd)
        assert np.array_equal(host(w, np.uint32), hw)


def test_device_solve_and_stats(gpu, ctx):
    torch = gpu
    n, m = 1 << 16, 1 << 20
    s = torch.empty(2 * m, dtype=torch.int32, device="cuda")
    d, w = torch.empty_like(s), torch.empty_like(s)
    ctx.gen_gnm_device(n, m, 254, 42, s, d, w)
    os_, od, ow = (torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3))
    rec = k.to_wedges(host(s, np.uint32), host(d, np.uint32), host(w, np.uint32))
    exp, exp_w = oracle.kruskal(rec)
    for algo in (k.BORUVKA, k.FILTER_BORUVKA):
        ctx.set_timing(True)  # totals from zero
        t0 = time.perf_counter()
        cnt, tw = ctx.solve_device(s, d, w, 2 * m, n, os_, od, ow, n, algorithm=algo)
        wall_ms = (time.perf_counter() - t0) * 1e3
        assert tw == exp_w and cnt == len(exp)
        assert np.array_equal(host(os_, np.uint32)[:cnt], exp["src"])
        assert np.array_equal(host(od, np.uint32)[:cnt], exp["dst"])
        assert np.array_equal(host(ow, np.uint32)[:cnt], exp["weight"])
        st = ctx.stats()
        assert st["mst_edges"] == cnt and st["rounds"]
        kern = st["kernels"]
        # filter-Boruvka may solve every piece of this size in the persistent
        # kernel
        assert sum(kern.get(kk, {}).get("launches", 0) for kk in ("argmin", "persistent_rounds")) >= 1
        # per-launch event pairs: every interval lies inside the solve
        assert all(v["ms"] >= 0 for v in kern.values())
        assert sum(v["ms"] for v in kern.values()) <= wall_ms
        if any(f["depth"] > 0 and f["heavy_kept"] > 0 for f in st.get("filter", [])):
            # heavy records kept below the top level: the relabel of
            # intermediate records (component bitmap, two passes) ran timed
            assert kern["relabel_scatter"]["launches"] >= 2
    ctx.set_timing(False)


def test_copy_readbacks_match_published(gpu, ctx, monkeypatch):
    """A context reads its counters back by publication into coherent pinned
    host words (the default) or, with KMST_READBACK_COPY=1 (and where the
    system has no coherent pinned memory), by copies and stream
    synchronisation; both give the oracle's MST, plain and filter."""
    torch = gpu
    monkeypatch.setenv("KMST_READBACK_COPY", "1")
    cp = k.Context(0)
    monkeypatch.delenv("KMST_READBACK_COPY")
    try:
        n, m = 1 << 15, 1 << 19
        s = torch.empty(2 * m, dtype=torch.int32, device="cuda")
        d, w = torch.empty_like(s), torch.empty_like(s)
        ctx.gen_gnm_device(n, m, 254, 7, s, d, w)
        rec = k.to_wedges(host(s, np.uint32), host(d, np.uint32), host(w, np.uint32))
        exp, exp_w = oracle.kruskal(rec)
        for c, mode in ((ctx, "publish"), (cp, "copy")):
            for algo in (k.BORUVKA, k.FILTER_BORUVKA):
                os_, od, ow = (torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3))
                cnt, tw = c.solve_device(s, d, w, 2 * m, n, os_, od, ow, n, algorithm=algo)
                assert c.stats()["readback"] == mode
                assert tw == exp_w and cnt == len(exp)
                assert np.array_equal(host(os_, np.uint32)[:cnt], exp["src"])
                assert np.array_equal(host(od, np.uint32)[:cnt], exp["dst"])
                assert np.array_equal(host(ow, np.uint32)[:cnt], exp["weight"])
    finally:
        cp.close()


def test_device_solve_misaligned_inputs(gpu, ctx):
    """Input arrays at a 4-byte offset (not 16-byte aligned) and an odd
    record count exercise the scalar tail of the twin search."""
    torch = gpu
    und = graphs.random_undirected(3001, 20001, 254, 77)
    rec = graphs.records(und)
    exp, exp_w = oracle.kruskal(rec)
    N = len(rec)
    n = int(max(rec["src"].max(), rec["dst"].max())) + 1
    bufs = []
    for col in ("src", "dst", "weight"):
        t = torch.zeros(N + 1, dtype=torch.int32, device="cuda")
        t[1:] = torch.from_numpy(rec[col].astype(np.uint32).view(np.int32)).cuda()
        bufs.append(t)
    views = [b[1:] for b in bufs]
    assert views[1].data_ptr() % 16 != 0
    outs = [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3)]
    for algo in (k.BORUVKA, k.FILTER_BORUVKA):
        cnt, tw = ctx.solve_device(views[0], views[1], views[2], N, n, outs[0], outs[1], outs[2], n,
                                   algorithm=algo, config=k.AlgorithmConfig(filter_threshold=1))
        assert tw == exp_w and cnt == len(exp)
        assert np.array_equal(host(outs[0], np.uint32)[:cnt], exp["src"])
        assert np.array_equal(host(outs[1], np.uint32)[:cnt], exp["dst"])


@pytest.mark.parametrize("n", [1 << 12, 1 << 16])
def test_rgg2d_positional_keys(gpu, n):
    """Config 3 shape at reduced n: weights up to 2^31-1 force positional keys."""
    s, d, w = k.gen_rgg2d_host(n, 16.0, (1 << 31) - 1, 3)
    check_equal(k.to_wedges(s, d, w), threshold=4)


@pytest.mark.parametrize("n", [1 << 13, 1 << 16])
def test_rhg_hubs(gpu, n):
    """Config 4 shape at reduced n: power-law degrees (hubs across many tiles)."""
    s, d, w = k.gen_rhg_host(n, 32.0, 2.8, 254, 4)
    check_equal(k.to_wedges(s, d, w), threshold=4)


@pytest.mark.parametrize("n,m,mw", [(4096, 1 << 16, 254),              # packed keys
                                    (70000, 400000, (1 << 31) - 1)])   # positional keys
def test_contract_errors_with_weight_classes(gpu, n, m, mw):
    """Inputs large enough for filter-Boruvka's weight classes: the input
    checks then run inside the class scatter (packed keys and positional keys,
    whose twin search there runs on unchecked records) instead of
    validate_kernel."""
    s, d, w = k.gen_gnm_host(n, m, mw, 13)
    good = k.to_wedges(s, d, w)
    exp, exp_w = oracle.kruskal(good)
    got, got_w = k.filter_boruvka(good)
    assert got_w == exp_w and np.array_equal(strip(got), strip(exp))
    i = len(good) // 2
    unsorted = good.copy()
    unsorted[[i, i + 1]] = unsorted[[i + 1, i]]
    asym = good.copy()
    asym["weight"][i] = asym["weight"][i] % mw + 1
    loop = good.copy()
    loop["dst"][i] = loop["src"][i]
    wide = good.copy()
    wide["dst"][-1] = n + 5
    for bad, ok in ((unsorted, (k.KMST_EUNSORTED,)), (asym, (k.KMST_EASYMMETRIC,)),
                    (loop, (k.KMST_EINVAL, k.KMST_EUNSORTED)),
                    (wide, (k.KMST_EINVAL, k.KMST_EUNSORTED, k.KMST_EASYMMETRIC))):
        with pytest.raises(k.KmstError) as e:
            k.filter_boruvka(bad)
        assert e.value.status in ok, (e.value.status, ok)
    # the context still works after the errors
    got, got_w = k.filter_boruvka(good)
    assert got_w == exp_w and np.array_equal(strip(got), strip(exp))


@pytest.mark.parametrize("n,m,mw,seed", [(1 << 12, 1 << 16, 254, 21), (70000, 300000, (1 << 31) - 1, 22),
                                         (2, 1, 254, 23)])
def test_compression_level_1(gpu, n, m, mw, seed):
    """compression=1: the host API sends the records difference-encoded and
    decodes them on the device; the MST is the same as without it."""
    s, d, w = k.gen_gnm_host(n, m, mw, seed)
    rec = k.to_wedges(s, d, w)
    exp, exp_w = oracle.kruskal(rec)
    for fn in (k.boruvka, k.filter_boruvka):
        got, got_w = fn(rec, k.AlgorithmConfig(compression=1))
        assert got_w == exp_w and np.array_equal(strip(got), strip(exp)), fn.__name__


@pytest.mark.parametrize("threshold", [2, 4])
def test_twin_relabels_on_a_random_graph(gpu, ctx, threshold):
    """A random graph past 2^20 vertices: the first round of the lightest
    class and every heavy class relabel only the src < dst half of their
    (symmetric) records and write both directions; the MSF equals the
    oracle's."""
    torch = gpu
    n, m = 1 << 21, 3 << 21
    arrs = [torch.empty(2 * m, dtype=torch.int32, device="cuda") for _ in range(3)]
    ctx.gen_gnm_device(n, m, 254, 91, *arrs)
    outs = [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3)]
    cnt, tw = ctx.solve_device(*arrs, 2 * m, n, *outs, n, algorithm=k.FILTER_BORUVKA,
                               config=k.AlgorithmConfig(filter_threshold=threshold))
    torch.cuda.synchronize()
    st = ctx.stats()
    assert st["rounds"][0]["twins"]
    rec = oracle.wedges_from_soa(*(host(a, np.uint32) for a in arrs))
    exp, exp_w = oracle.boruvka(rec, 1, 8)
    assert tw == exp_w and cnt == len(exp)
    assert np.array_equal(host(outs[0][:cnt], np.uint32), exp["src"].astype(np.uint32))
    assert np.array_equal(host(outs[1][:cnt], np.uint32), exp["dst"].astype(np.uint32))
    assert np.array_equal(host(outs[2][:cnt], np.uint32), exp["weight"])
