"""C-ABI and host-side logic, CPU only (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import kamsta_amd as k

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "kmst.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kmst_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(k.LIB_PATH)
    declared = header_symbols()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(k.EXPORTED_SYMBOLS) == declared


def test_config_default_matches_reference():
    cfg = k._Config()
    k.load().kmst_config_default(ctypes.byref(cfg))
    # AlgorithmConfig defaults: NO_PREPROCESSING, NO_COMPRESSION, threshold 4
    assert (cfg.algorithm, cfg.local_preprocessing, cfg.compression, cfg.filter_threshold) == (0, 0, 0, 4)


def test_host_api_rejects_wide_ids_without_gpu():
    e = np.zeros(2, dtype=k.WEDGE_DTYPE)
    e["src"] = [0, 1 << 33]
    e["dst"] = [1 << 33, 0]
    e["weight"] = 1
    with pytest.raises(k.KmstError) as err:
        k.boruvka(e)
    assert err.value.status == k.KMST_EINVAL
    with pytest.raises(k.KmstError) as err:
        k.boruvka(e, k.AlgorithmConfig(compression=1))
    assert err.value.status == k.KMST_EINVAL
    # compression levels beyond the difference codec are rejected before any
    # device work
    ok = k.to_wedges(np.array([0, 1], np.uint32), np.array([1, 0], np.uint32),
                     np.array([5, 5], np.uint32))
    with pytest.raises(k.KmstError) as err:
        k.boruvka(ok, k.AlgorithmConfig(compression=2))
    assert err.value.status == k.KMST_EINVAL


def test_kernel_entries_reject_null_context_without_gpu():
    lib = k.load()
    assert lib.kmst_k_segmented_argmin(None, None, None, None, 0, 0, None, None) == k.KMST_EINVAL
    assert lib.kmst_k_segmented_argmin_packed(None, None, None, None, 8, 0, 0, None,
                                              None) == k.KMST_EINVAL
    assert b"null context" in lib.kmst_last_error()


@pytest.mark.parametrize("n,m,mw,seed", [(2, 1, 254, 0), (16, 120, 3, 1), (300, 2000, 254, 2),
                                         (1 << 14, 1 << 16, 254, 0x4B414D5354410001),
                                         (256, 4000, 254, 9)])
def test_gnm_host_contract(n, m, mw, seed):
    s, d, w = k.gen_gnm_host(n, m, mw, seed)
    assert len(s) == 2 * m
    key = s.astype(np.uint64) << 32 | d
    assert np.all(np.diff(key.astype(np.int64)) > 0), "sorted, no duplicate records"
    assert np.all(s != d) and s.max() < n and d.max() < n
    assert np.all((w >= 1) & (w <= mw))
    fwd = s < d
    assert fwd.sum() == m
    # symmetric with equal weights
    rev = {(int(b), int(a)): int(x) for a, b, x in zip(s[~fwd], d[~fwd], w[~fwd])}
    for a, b, x in zip(s[fwd], d[fwd], w[fwd]):
        assert rev[(int(a), int(b))] == int(x)
    s2, d2, w2 = k.gen_gnm_host(n, m, mw, seed)
    assert np.array_equal(s, s2) and np.array_equal(d, d2) and np.array_equal(w, w2)


def test_gnm_prefix_property():
    # the first m distinct pairs of the counter sequence: a smaller m is a subset
    s1, d1, _ = k.gen_gnm_host(500, 800, 254, 3)
    s2, d2, _ = k.gen_gnm_host(500, 1600, 254, 3)
    a = set(zip(s1.tolist(), d1.tolist()))
    b = set(zip(s2.tolist(), d2.tolist()))
    assert a < b


def test_gnm_rejects_bad_args():
    with pytest.raises(k.KmstError):
        k.gen_gnm_host(4, 7, 254, 1)  # more than n(n-1)/2 pairs
    with pytest.raises(k.KmstError):
        k.gen_gnm_host(4, 2, 0, 1)


def test_driver_help_without_gpu():
    import subprocess
    exe = os.path.join(ROOT, "kamsta_amd", "mst_benchmarks")
    assert os.path.exists(exe), "built by kamsta_amd.build"
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0 and "--filter_threshold" in r.stdout
    r = subprocess.run([exe, "--bogus"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 2


@pytest.mark.parametrize("gen", ["rgg", "rhg"])
def test_geometric_generators_contract(gen):
    import oracle
    if gen == "rgg":
        s, d, w = k.gen_rgg2d_host(5000, 16.0, (1 << 31) - 1, 3)
        assert 12 < len(s) / 5000 < 17 and w.max() > (1 << 20)
    else:
        s, d, w = k.gen_rhg_host(5000, 32.0, 2.8, 254, 4)
        assert 24 < len(s) / 5000 < 36 and np.bincount(s).max() > 200  # heavy tail
    key = s.astype(np.uint64) << 32 | d
    assert np.all(np.diff(key.astype(np.int64)) > 0) and np.all(s != d)
    fwd = s < d
    assert fwd.sum() * 2 == len(s)
    rev = {(int(b), int(a)): int(x) for a, b, x in zip(s[~fwd], d[~fwd], w[~fwd])}
    assert all(rev[(int(a), int(b))] == int(x) for a, b, x in zip(s[fwd], d[fwd], w[fwd]))
    s2, d2, w2 = (k.gen_rgg2d_host(5000, 16.0, (1 << 31) - 1, 3) if gen == "rgg"
                  else k.gen_rhg_host(5000, 32.0, 2.8, 254, 4))
    assert np.array_equal(s, s2) and np.array_equal(d, d2) and np.array_equal(w, w2)
    e = k.to_wedges(s, d, w)
    kr = oracle.kruskal(e)
    for lp in (0, 1):
        b = oracle.boruvka(e, lp)
        assert b[1] == kr[1] and np.array_equal(b[0], kr[0])


def test_tuning_default_matches_python_mirror():
    t = k._Tuning()
    k.load().kmst_tuning_default(ctypes.byref(t))
    assert k.Tuning._from(t) == k.Tuning()
    assert (t.sort_threads, t.balanced_digits, t.sort_wide_min, t.persistent_max) == \
        (0, 1, 1 << 20, 1 << 18)


def test_server_config_default_carries_the_default_tuning():
    c = k._ServerConfig()
    k.load().kmst_server_config_default(ctypes.byref(c))
    assert k.Tuning._from(c.tuning) == k.Tuning()
    assert (c.algorithm, c.filter_threshold, c.buffer_sets) == (k.FILTER_BORUVKA, 4, 3)
