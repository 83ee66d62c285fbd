"""Batched serving: every request's result is identical alone and in the mix,
and equals the oracle's Kruskal."""
import numpy as np
import pytest

import graphs
import kamsta_amd as k
import oracle

pytestmark = pytest.mark.gpu


def requests(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(2 ** rng.uniform(4, 12))
        deg = rng.uniform(2, 12)
        m = int(min(n * deg / 2, n * (n - 1) // 2))
        s, d, w = k.gen_gnm_host(n, m, 254, 1000 * seed + i) if m else (np.zeros(0, np.uint32),) * 3
        out.append((s, d, w, n))
    return out


def solve_all(server, reqs):
    tickets = [server.submit(s, d, w, n) for s, d, w, n in reqs]
    return [server.wait(t) for t in tickets]


@pytest.mark.parametrize("algo", [k.BORUVKA, k.FILTER_BORUVKA])
def test_batched_equals_alone_and_oracle(gpu, algo):
    reqs = requests(60, 3)
    batched = k.Server(algorithm=algo, max_batch_requests=4096)
    alone = k.Server(algorithm=algo, max_batch_requests=1)
    rb = solve_all(batched, reqs)
    ra = solve_all(alone, reqs)
    nb, qb = batched.counters()
    batched.close()
    alone.close()
    assert qb == len(reqs) and nb < len(reqs), "requests were batched"
    assert max(r[0]["batch_requests"] for r in rb) > 1
    for (s, d, w, n), (resb, eb), (resa, ea) in zip(reqs, rb, ra):
        assert resb["status"] == 0 and resa["status"] == 0
        assert resb["total_weight"] == resa["total_weight"]
        for x, y in zip(eb, ea):
            assert np.array_equal(x, y)
        if len(s):
            exp, exp_w = oracle.kruskal(k.to_wedges(s, d, w))
            assert resb["total_weight"] == exp_w
            assert np.array_equal(eb[0], exp["src"]) and np.array_equal(eb[1], exp["dst"])
            assert np.array_equal(eb[2], exp["weight"])
        assert resb["done_ns"] >= resb["start_ns"] >= resb["admit_ns"] >= resb["submit_ns"]


def test_batched_wide_weights(gpu):
    """31-bit weights on a batch of more than 2^17 vertex ids: positional
    keys, made by the class scatter and the outside gather of a partitioned
    solve."""
    rng = np.random.default_rng(17)
    reqs = []
    for i in range(40):
        n = int(2 ** rng.uniform(11, 13.5))
        m = int(n * rng.uniform(4, 16) / 2)
        s, d, w = k.gen_gnm_host(n, m, (1 << 31) - 1, 7000 + i)
        reqs.append((s, d, w, n))
    assert sum(r[3] for r in reqs) > (1 << 17)
    batched = k.Server(algorithm=k.FILTER_BORUVKA, max_batch_requests=4096)
    alone = k.Server(algorithm=k.FILTER_BORUVKA, max_batch_requests=1)
    rb = solve_all(batched, reqs)
    ra = solve_all(alone, reqs)
    batched.close()
    alone.close()
    assert max(r[0]["batch_requests"] for r in rb) > 1
    for (s, d, w, n), (resb, eb), (resa, ea) in zip(reqs, rb, ra):
        assert resb["status"] == 0 and resa["status"] == 0
        exp, exp_w = oracle.kruskal(k.to_wedges(s, d, w))
        assert resb["total_weight"] == exp_w == resa["total_weight"]
        assert np.array_equal(eb[0], exp["src"]) and np.array_equal(eb[1], exp["dst"])
        assert np.array_equal(eb[2], exp["weight"])
        for x, y in zip(eb, ea):
            assert np.array_equal(x, y)


def test_batch_of_mixed_upload_forms(gpu):
    """A batch mixing requests of byte weights and of wide weights, with
    isolated vertices, an empty one, and one whose records are out of order
    (refused when it is solved alone; every other request is exact)."""
    rng = np.random.default_rng(23)
    reqs = []
    for i in range(24):
        n = int(2 ** rng.uniform(6, 12))
        m = int(n * rng.uniform(2, 8) / 2)
        mw = 254 if i % 3 else (1 << 20)
        s, d, w = k.gen_gnm_host(n, m, mw, 9000 + i)
        reqs.append((s, d, w, n + (7 if i % 5 == 0 else 0)))  # some with isolated ids
    reqs[10] = (np.zeros(0, np.uint32),) * 3 + (16,)
    bad = 17
    s, d, w, n = reqs[bad]
    s, d = s.copy(), d.copy()
    s[[3, len(s) - 3]] = s[[len(s) - 3, 3]]
    d[[3, len(d) - 3]] = d[[len(d) - 3, 3]]
    reqs[bad] = (s, d, w, n)
    srv = k.Server(algorithm=k.FILTER_BORUVKA, max_batch_requests=4096)
    res = solve_all(srv, reqs)
    srv.close()
    for i, ((s, d, w, n), (r, e)) in enumerate(zip(reqs, res)):
        if i == bad:
            assert r["status"] in (k.KMST_EUNSORTED, k.KMST_EASYMMETRIC), r["status"]
            continue
        assert r["status"] == 0, (i, r["status"])
        if len(s) == 0:
            assert r["mst_edges"] == 0
            continue
        exp, exp_w = oracle.kruskal(k.to_wedges(s, d, w))
        assert r["total_weight"] == exp_w
        assert np.array_equal(e[0], exp["src"]) and np.array_equal(e[1], exp["dst"])
        assert np.array_equal(e[2], exp["weight"])


def test_bad_request_does_not_poison_batch(gpu):
    reqs = requests(12, 5)
    s, d, w, n = reqs[4]
    bad = (s.copy(), d.copy(), w.copy(), n)
    bad[2][0] += 1  # asymmetric weight
    reqs[4] = bad
    srv = k.Server(max_batch_requests=4096)
    res = solve_all(srv, reqs)
    srv.close()
    for i, ((s, d, w, n), (r, e)) in enumerate(zip(reqs, res)):
        if i == 4:
            assert r["status"] == k.KMST_EASYMMETRIC
        else:
            assert r["status"] == 0
            if len(s):
                assert r["total_weight"] == oracle.kruskal(k.to_wedges(s, d, w))[1]


def test_admission_rejects_oversize(gpu):
    srv = k.Server(max_batch_records=1000)
    s, d, w = k.gen_gnm_host(200, 800, 254, 1)
    with pytest.raises(k.KmstError) as e:
        srv.submit(s, d, w, 200)
    assert e.value.status == k.KMST_ENOMEM
    s, d, w = k.gen_gnm_host(100, 300, 254, 2)
    r, edges = srv.wait(srv.submit(s, d, w, 100))
    assert r["status"] == 0 and r["total_weight"] == oracle.kruskal(k.to_wedges(s, d, w))[1]
    srv.close()


@pytest.mark.parametrize("where", ["device", "pinned", "mixed"])
def test_direct_delivery_to_device_writable_outputs(gpu, where):
    """Outputs in device memory or pinned host memory are written by the
    split kernel itself; pageable ones are copied.  Every kind, alone or mixed
    in one batch, gets the oracle's result."""
    torch = gpu
    reqs = requests(45, 7)
    kinds = {"device": ["device"], "pinned": ["pinned"], "mixed": ["device", "pinned", "host"]}[where]
    outs = []
    for i, (s, d, w, n) in enumerate(reqs):
        cap = max(n, 1)
        kind = kinds[i % len(kinds)]
        if kind == "device":
            outs.append([torch.zeros(cap, dtype=torch.int32, device="cuda") for _ in range(3)])
        elif kind == "pinned":
            outs.append([torch.zeros(cap, dtype=torch.int32).pin_memory() for _ in range(3)])
        else:
            outs.append([np.zeros(cap, np.uint32) for _ in range(3)])
    srv = k.Server(max_batch_requests=4096)
    tickets = [srv.submit(s, d, w, n, out=o) for (s, d, w, n), o in zip(reqs, outs)]
    res = [srv.wait(t) for t in tickets]
    nb, _ = srv.counters()
    srv.close()
    assert nb < len(reqs), "requests were batched"
    for (s, d, w, n), (r, e) in zip(reqs, res):
        assert r["status"] == 0
        got = [x.cpu().numpy().view(np.uint32) if hasattr(x, "cpu") else x for x in e]
        if len(s):
            exp, exp_w = oracle.kruskal(k.to_wedges(s, d, w))
            assert r["total_weight"] == exp_w
            assert np.array_equal(got[0], exp["src"]) and np.array_equal(got[1], exp["dst"])
            assert np.array_equal(got[2], exp["weight"])
        else:
            assert r["mst_edges"] == 0


def test_request_leaving_its_id_range_is_refused(gpu):
    """A request with a vertex id at or past its own n could, after the batch
    offsets, join the next request's id range while the batch stays sorted
    and symmetric.  The offset pass catches it: that request reports
    KMST_EINVAL (as it does alone) and the other is solved correctly."""
    a = (np.array([0, 1, 3, 12], np.uint32), np.array([1, 0, 12, 3], np.uint32),
         np.array([5, 5, 7, 7], np.uint32), 10)
    b = (np.array([3, 4], np.uint32), np.array([4, 3], np.uint32), np.array([9, 9], np.uint32), 5)
    srv = k.Server(max_batch_requests=4096)
    ta = srv.submit(*a)
    tb = srv.submit(*b)
    ra, _ = srv.wait(ta)
    rb, eb = srv.wait(tb)
    srv.close()
    assert ra["status"] == k.KMST_EINVAL
    assert rb["status"] == 0 and rb["total_weight"] == 9
    assert list(eb[0]) == [3] and list(eb[1]) == [4]


@pytest.mark.parametrize("max_batch", [4096, 1])
def test_device_resident_inputs_gathered_or_in_place(gpu, max_batch):
    """Requests whose arrays are in device memory are not copied: a batch's
    offset pass reads them in place (mixed here with host requests, which
    are copied), and a request alone is solved straight from its arrays."""
    torch = gpu
    reqs = requests(40, 11)
    srv = k.Server(max_batch_requests=max_batch)
    subs = []
    for i, (s, d, w, n) in enumerate(reqs):
        if i % 2 == 0:
            dev = [torch.from_numpy(np.ascontiguousarray(x, np.uint32).view(np.int32)).cuda()
                   for x in (s, d, w)]
            subs.append(srv.submit(*dev, n))
        else:
            subs.append(srv.submit(s, d, w, n))
    res = [srv.wait(t) for t in subs]
    srv.close()
    for (s, d, w, n), (r, e) in zip(reqs, res):
        assert r["status"] == 0
        if len(s):
            exp, exp_w = oracle.kruskal(k.to_wedges(s, d, w))
            assert r["total_weight"] == exp_w
            assert np.array_equal(e[0], exp["src"]) and np.array_equal(e[1], exp["dst"])
            assert np.array_equal(e[2], exp["weight"])
