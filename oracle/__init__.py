"""CPU oracle for the MST hot path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings for ``oracle/libkmst_oracle.so`` (built from
``oracle/kmst_oracle.cpp`` by ``oracle/Makefile``), a restatement of KaMSTa's
Boruvka / filter-Boruvka on one process with the strict (weight, uid) key.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
it, and only as the checker.  Inputs and outputs are numpy arrays of
``WEDGE_DTYPE``, the 24-byte layout of hybridMST::WEdge
(include/definitions.hpp:66-84).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkmst_oracle.so")

WEDGE_DTYPE = np.dtype([("src", "<u8"), ("dst", "<u8"), ("weight", "<u4"), ("_pad", "<u4")])

STATUS = {0: "ok", 1: "invalid", 2: "unsorted", 3: "asymmetric"}


class OracleError(RuntimeError):
    def __init__(self, status, what):
        super().__init__(f"{what}: status {status} ({STATUS.get(status, '?')})")
        self.status = status


_lib = None


def build():
    """Compile the oracle library with its Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        p, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        pu64 = ctypes.POINTER(ctypes.c_uint64)
        L.orc_pair_records.argtypes = [p, u64, p, pu64]
        L.orc_kruskal.argtypes = [p, u64, p, pu64, pu64]
        L.orc_boruvka.argtypes = [p, u64, i32, i32, p, pu64, pu64]
        L.orc_filter_boruvka.argtypes = [p, u64, i32, u64, i32, p, pu64, pu64]
        L.orc_wedges_from_soa.argtypes = [p, p, p, u64, p]
        L.orc_wedges_from_soa.restype = None
        for f in (L.orc_pair_records, L.orc_kruskal, L.orc_boruvka, L.orc_filter_boruvka):
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def _prep(edges):
    edges = np.ascontiguousarray(edges, dtype=WEDGE_DTYPE)
    return edges, edges.ctypes.data, len(edges)


def _result(fn, edges, *args):
    edges, ptr, m = _prep(edges)
    out = np.zeros(max(m // 2, 1), dtype=WEDGE_DTYPE)
    n_out = ctypes.c_uint64(0)
    total = ctypes.c_uint64(0)
    st = fn(ptr, m, *args, out.ctypes.data, ctypes.byref(n_out), ctypes.byref(total))
    if st != 0:
        raise OracleError(st, fn.__name__)
    return out[: n_out.value].copy(), total.value


def pair_records(edges):
    """uid of every record (position of its (min, max) twin) and n."""
    edges, ptr, m = _prep(edges)
    uid = np.zeros(m, dtype=np.uint32)
    n = ctypes.c_uint64(0)
    st = lib().orc_pair_records(ptr, m, uid.ctypes.data, ctypes.byref(n))
    if st != 0:
        raise OracleError(st, "pair_records")
    return uid, n.value


def wedges_from_soa(src, dst, weight):
    """WEDGE_DTYPE records from uint32 (src, dst, weight) arrays (parallel copy)."""
    s, d, w = (np.ascontiguousarray(a, dtype=np.uint32) for a in (src, dst, weight))
    assert len(s) == len(d) == len(w)
    out = np.empty(len(s), dtype=WEDGE_DTYPE)
    if len(s):
        lib().orc_wedges_from_soa(s.ctypes.data, d.ctypes.data, w.ctypes.data, len(s),
                                  out.ctypes.data)
    return out


def kruskal(edges):
    """Reference check order Kruskal; returns (canonical MST edges, u64 weight)."""
    return _result(lib().orc_kruskal, edges)


def boruvka(edges, local_preprocessing=1, threads=0):
    """Restatement of hybridMST::boruvka on one process."""
    return _result(lib().orc_boruvka, edges, int(local_preprocessing), int(threads))


def filter_boruvka(edges, local_preprocessing=1, filter_threshold=4, threads=0):
    """Restatement of hybridMST::filter_boruvka on one process."""
    return _result(lib().orc_filter_boruvka, edges, int(local_preprocessing),
                   int(filter_threshold), int(threads))
