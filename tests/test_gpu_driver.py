"""The C++ drop-in header and command-line driver, on the GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

import graphs
import kamsta_amd as k
import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "kamsta_amd", "mst_benchmarks")


def run(args):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("algo", ["hybridBoruvka", "filter_hybridBoruvka"])
def test_driver_gnm_matches_oracle(gpu, algo):
    out = run(["-a", algo, "--log_n", "12", "--log_m", "16", "--seed", "9", "-i", "2", "-c"])
    res = [dict(kv.split("=") for kv in line.split()[1:]) for line in out.splitlines()
           if line.startswith("RESULT")]
    assert len(res) == 2
    s, d, w = k.gen_gnm_host(1 << 12, 1 << 16, 254, 9)
    exp, exp_w = oracle.kruskal(k.to_wedges(s, d, w))
    for r in res:
        assert int(r["mst_weight"]) == exp_w and int(r["mst_edges"]) == len(exp)
    assert re.search(r"CHECK identical_across_algorithms=1 forest=1", out)


def test_driver_infile_weighted(gpu, tmp_path):
    und = graphs.random_undirected(500, 3000, 254, 4)
    rec = np.zeros(len(und), dtype=[("s", "<u4"), ("d", "<u4"), ("w", "u1")])
    for i, (u, v, w) in enumerate(und):
        rec[i] = (v, u, w) if i % 2 else (u, v, w)  # one direction per edge, any order
    f = tmp_path / "g.bin"
    rec[np.random.default_rng(1).permutation(len(rec))].tofile(f)
    out = run(["-g", "INFILE_WEIGHTED", "-f", str(f), "--symmetrize", "-a", "filter_hybridBoruvka",
               "-c"])
    got = int(re.search(r"mst_weight=(\d+)", out).group(1))
    assert got == oracle.kruskal(graphs.records(und))[1]
