"""Writes the golden fixtures (inputs + canonical MST + u64 weight).

The reference ships no golden vectors (its only check compares total weight
against Kruskal, benchmarks/mst_benchmarks.hpp:287-318), so the expected
outputs here come from brute-force enumeration where the graph is tiny and
from the oracle's Kruskal otherwise (itself pinned by brute force and closed
forms in tests/test_oracle.py).  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import graphs  # noqa: E402
import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def write(name, rec, mst, total):
    np.savez_compressed(os.path.join(OUT, name + ".npz"),
                        src=rec["src"].astype(np.uint64), dst=rec["dst"].astype(np.uint64),
                        weight=rec["weight"].astype(np.uint32),
                        mst_src=mst["src"].astype(np.uint64), mst_dst=mst["dst"].astype(np.uint64),
                        mst_weight=mst["weight"].astype(np.uint32),
                        total_weight=np.uint64(total))


def from_triples(triples):
    m = np.zeros(len(triples), dtype=graphs.WEDGE_DTYPE)
    for i, (u, v, w) in enumerate(triples):
        m[i] = (u, v, w, 0)
    return m


def main():
    # tiny hand graphs: expected output by brute force
    hand = {
        "tiny_ties": (5, [(0, 1, 1), (1, 2, 1), (2, 3, 1), (3, 4, 1), (0, 4, 1), (1, 3, 1), (0, 2, 1)]),
        "tiny_multi": (4, [(0, 1, 5), (0, 1, 2), (1, 2, 2), (1, 2, 2), (2, 3, 9), (0, 3, 9), (0, 2, 4)]),
        "tiny_forest": (8, [(0, 1, 3), (1, 2, 1), (0, 2, 2), (4, 5, 7), (5, 6, 7), (4, 6, 7)]),
    }
    for name, (n, und) in hand.items():
        bf = graphs.brute_force_msf(n, und)
        write(name, graphs.records(und), from_triples(bf), sum(w for *_, w in bf))
    # config-1 shape at 1/16 scale from the repository's own GNM generator
    import kamsta_amd as k
    s, d, w = k.gen_gnm_host(1 << 10, 1 << 14, 254, 0x4B414D5354410001)
    rec = k.to_wedges(s, d, w)
    mst, total = oracle.kruskal(rec)
    write("gnm_n1024_m16384", rec, mst, total)
    print("written to", OUT)


if __name__ == "__main__":
    main()
