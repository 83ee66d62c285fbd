"""kamsta_amd -- MI355X-native Boruvka / filter-Boruvka minimum spanning forest.

Python mirror of KaMSTa's algorithm entry points over the C ABI of
``libkmst.so`` (declared in include/kmst.h):

* ``boruvka(edges, config)``        <- hybridMST::boruvka
  (include/algorithms/hybrid_boruvka.hpp:263-301)
* ``filter_boruvka(edges, config)`` <- hybridMST::filter_boruvka
  (include/algorithms/hybrid_boruvka_edgefilter.hpp:420-462)
* ``AlgorithmConfig``               <- AlgorithmConfig
  (include/algorithms/algorithm_configurations.hpp:28-34)

``edges`` is a numpy array of ``WEDGE_DTYPE`` (the 24-byte hybridMST::WEdge,
include/definitions.hpp:66-84): every undirected edge in both directions,
sorted by (src, dst).  The result is the MST as canonical (min, max, weight)
records sorted by (min, max), plus its u64 total weight.  Unlike the
reference, which aborts, contract violations raise ``KmstError`` with the C
status code.  All compute runs on the GPU through libkmst; there is no CPU
fallback, and a missing library raises ``ImportError``.
"""
import atexit
import ctypes
import dataclasses
import json
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkmst.so")

WEDGE_DTYPE = np.dtype([("src", "<u8"), ("dst", "<u8"), ("weight", "<u4"), ("_pad", "<u4")])

KMST_OK, KMST_EINVAL, KMST_EUNSORTED, KMST_EASYMMETRIC, KMST_ENOMEM, KMST_EDEVICE, KMST_EBUSY = range(7)
STATUS_NAMES = {0: "OK", 1: "EINVAL", 2: "EUNSORTED", 3: "EASYMMETRIC", 4: "ENOMEM",
                5: "EDEVICE", 6: "EBUSY"}

BORUVKA, FILTER_BORUVKA = 0, 1

# every symbol include/kmst.h declares
EXPORTED_SYMBOLS = [
    "kmst_gen_rgg2d_host", "kmst_gen_rhg_host", "kmst_gen_rgg2d_device", "kmst_gen_rhg_device",
    "kmst_config_default", "kmst_context_create", "kmst_context_destroy", "kmst_context_trim",
    "kmst_context_reserve",
    "kmst_boruvka", "kmst_filter_boruvka", "kmst_solve_device", "kmst_solve_device_noalloc",
    "kmst_context_set_timing",
    "kmst_tuning_default", "kmst_context_set_tuning", "kmst_context_get_tuning", "kmst_autotune",
    "kmst_encode_host",
    "kmst_last_error", "kmst_context_stats_json", "kmst_gen_gnm_device", "kmst_gen_gnm_host",
    "kmst_k_segmented_argmin", "kmst_k_segmented_argmin_packed", "kmst_k_sort_pairs",
    "kmst_k_pointer_jump",
    "kmst_server_config_default", "kmst_server_create", "kmst_server_submit", "kmst_server_poll",
    "kmst_server_wait", "kmst_server_counters", "kmst_server_destroy",
]


class KmstError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class _Config(ctypes.Structure):
    _fields_ = [("algorithm", ctypes.c_int32), ("local_preprocessing", ctypes.c_int32),
                ("compression", ctypes.c_int32), ("device", ctypes.c_int32),
                ("filter_threshold", ctypes.c_uint64)]


class _DevEdges(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("dst", ctypes.c_void_p), ("weight", ctypes.c_void_p),
                ("m_directed", ctypes.c_uint64), ("n", ctypes.c_uint64)]


class _Tuning(ctypes.Structure):
    _fields_ = [("sort_threads", ctypes.c_uint32), ("balanced_digits", ctypes.c_uint32),
                ("sort_wide_min", ctypes.c_uint64), ("persistent_max", ctypes.c_uint64)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("algorithm", ctypes.c_int32),
                ("filter_threshold", ctypes.c_uint64), ("max_batch_records", ctypes.c_uint64),
                ("max_batch_requests", ctypes.c_uint64), ("hbm_budget_bytes", ctypes.c_uint64),
                ("queue_capacity", ctypes.c_uint32), ("buffer_sets", ctypes.c_uint32),
                ("tuning", _Tuning)]


class _RequestResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("status", ctypes.c_int32), ("done", ctypes.c_int32),
                ("mst_edges", ctypes.c_uint64), ("total_weight", ctypes.c_uint64),
                ("batch_requests", ctypes.c_uint64), ("submit_ns", ctypes.c_uint64),
                ("admit_ns", ctypes.c_uint64), ("start_ns", ctypes.c_uint64),
                ("solved_ns", ctypes.c_uint64), ("done_ns", ctypes.c_uint64)]


@dataclasses.dataclass
class Tuning:
    """Mirror of kmst_tuning: launch shapes of a context's solves (speed only;
    every tuning returns the same MST)."""
    sort_threads: int = 0          # 256, 512, or 0 = by size
    balanced_digits: int = 1       # digit widths balanced over the passes
    sort_wide_min: int = 1 << 20   # with sort_threads 0: 512-thread tiles from here on
    persistent_max: int = 1 << 18  # rounds on at most this many records: one persistent launch

    def _c(self):
        return _Tuning(self.sort_threads, self.balanced_digits, self.sort_wide_min,
                       self.persistent_max)

    @staticmethod
    def _from(t):
        return Tuning(t.sort_threads, t.balanced_digits, t.sort_wide_min, t.persistent_max)


class _DevMst(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("dst", ctypes.c_void_p), ("weight", ctypes.c_void_p),
                ("capacity", ctypes.c_uint64)]


@dataclasses.dataclass
class AlgorithmConfig:
    """Mirror of hybridMST::AlgorithmConfig; local_preprocessing is accepted
    and has no effect on one GPU (the whole graph is local)."""
    local_preprocessing: int = 0
    compression: int = 0
    filter_threshold: int = 4
    device: int = 0

    def _c(self, algorithm):
        return _Config(algorithm, self.local_preprocessing, self.compression, self.device,
                       self.filter_threshold)


_lib = None

# Contexts and servers still open when the interpreter exits are closed by an
# atexit hook, which runs before the HIP runtime's own exit handlers; a
# close from __del__ during interpreter teardown could come after them.
_live = weakref.WeakSet()


@atexit.register
def _close_live():
    for obj in sorted(list(_live), key=lambda o: o._order, reverse=True):
        try:
            obj.close()
        except Exception:
            pass


_order = [0]


def _track(obj):
    _order[0] += 1
    obj._order = _order[0]
    _live.add(obj)


def load():
    """Load libkmst.so (built by kamsta_amd.build); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m kamsta_amd.build` first")
    # torch ships its own HIP runtime under the same soname: when torch is
    # installed it is loaded first, so that libkmst binds to that runtime
    # rather than starting the system one, after which torch finds no GPU
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
    pu64 = ctypes.POINTER(ctypes.c_uint64)
    sig = {
        "kmst_config_default": (None, [vp]),
        "kmst_context_create": (i32, [i32, ctypes.POINTER(vp)]),
        "kmst_context_destroy": (None, [vp]),
        "kmst_context_trim": (None, [vp]),
        "kmst_context_reserve": (i32, [vp, u64, u64]),
        "kmst_boruvka": (i32, [vp, u64, vp, vp, pu64, pu64]),
        "kmst_filter_boruvka": (i32, [vp, u64, vp, vp, pu64, pu64]),
        "kmst_solve_device": (i32, [vp, vp, vp, vp, pu64, pu64]),
        "kmst_solve_device_noalloc": (i32, [vp, vp, vp, pu64, pu64, vp]),
        "kmst_context_set_timing": (None, [vp, i32]),
        "kmst_tuning_default": (None, [vp]),
        "kmst_encode_host": (i32, [vp, vp, vp, u64, u32, vp, u64, vp, vp, pu64]),
        "kmst_context_set_tuning": (i32, [vp, vp]),
        "kmst_context_get_tuning": (i32, [vp, vp]),
        "kmst_autotune": (i32, [vp, vp, vp, ctypes.c_int32, vp, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_uint32)]),
        "kmst_last_error": (ctypes.c_char_p, []),
        "kmst_context_stats_json": (ctypes.c_size_t, [vp, vp, ctypes.c_size_t]),
        "kmst_gen_gnm_device": (i32, [vp, u64, u64, u32, u64, vp, vp, vp]),
        "kmst_gen_gnm_host": (i32, [u64, u64, u32, u64, vp, vp, vp]),
        "kmst_gen_rgg2d_host": (i32, [u64, ctypes.c_double, u32, u64, vp, vp, vp, u64, pu64]),
        "kmst_gen_rhg_host": (i32, [u64, ctypes.c_double, ctypes.c_double, u32, u64, vp, vp, vp,
                                    u64, pu64]),
        "kmst_gen_rgg2d_device": (i32, [vp, u64, ctypes.c_double, u32, u64, vp, vp, vp, u64,
                                        pu64]),
        "kmst_gen_rhg_device": (i32, [vp, u64, ctypes.c_double, ctypes.c_double, u32, u64, vp, vp,
                                      vp, u64, pu64]),
        "kmst_k_segmented_argmin": (i32, [vp, vp, vp, vp, u64, u64, vp, vp]),
        "kmst_k_segmented_argmin_packed": (i32, [vp, vp, vp, vp, u32, u64, u64, vp, vp]),
        "kmst_k_sort_pairs": (i32, [vp, vp, vp, vp, u64, u32]),
        "kmst_k_pointer_jump": (i32, [vp, vp, u64]),
        "kmst_server_config_default": (None, [vp]),
        "kmst_server_create": (i32, [vp, ctypes.POINTER(vp)]),
        "kmst_server_submit": (i32, [vp, vp, vp, vp, u64, u64, vp, vp, vp, u64, pu64]),
        "kmst_server_poll": (i32, [vp, u64, vp]),
        "kmst_server_wait": (i32, [vp, u64, vp]),
        "kmst_server_counters": (i32, [vp, pu64, pu64]),
        "kmst_server_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(status):
    if status != KMST_OK:
        msg = load().kmst_last_error()
        raise KmstError(status, msg.decode() if msg else "")


def _host_solve(fn, edges, config):
    edges = np.ascontiguousarray(edges, dtype=WEDGE_DTYPE)
    m = len(edges)
    out = np.zeros(max(m // 2, 1), dtype=WEDGE_DTYPE)
    n_out, total = ctypes.c_uint64(0), ctypes.c_uint64(0)
    cfg = config._c(BORUVKA)
    _check(fn(edges.ctypes.data if m else None, m, ctypes.byref(cfg), out.ctypes.data,
              ctypes.byref(n_out), ctypes.byref(total)))
    return out[: n_out.value].copy(), total.value


def boruvka(edges, config=None):
    """hybridMST::boruvka: returns (canonical MST records, u64 total weight)."""
    return _host_solve(load().kmst_boruvka, edges, config or AlgorithmConfig())


def filter_boruvka(edges, config=None):
    """hybridMST::filter_boruvka: returns (canonical MST records, u64 total weight)."""
    return _host_solve(load().kmst_filter_boruvka, edges, config or AlgorithmConfig())


def solve_device_noalloc(src, dst, weight, m, n, out_src, out_dst, out_weight, capacity,
                         algorithm=BORUVKA, config=None, stream=None):
    """kmst_solve_device_noalloc: a device-resident solve that allocates no
    device memory (the library's fixed per-GPU scratch; at most 2^20 records
    and 2^19 vertices).  Device arrays are torch tensors or raw addresses;
    stream is a raw hipStream_t (None: the default stream).  Returns (number
    of MST edges, u64 weight)."""
    L = load()
    cfg = (config or AlgorithmConfig())._c(algorithm)
    e = _DevEdges(_ptr(src).value, _ptr(dst).value, _ptr(weight).value, m, n)
    o = _DevMst(_ptr(out_src).value, _ptr(out_dst).value, _ptr(out_weight).value, capacity)
    n_out, total = ctypes.c_uint64(0), ctypes.c_uint64(0)
    _check(L.kmst_solve_device_noalloc(ctypes.byref(e), ctypes.byref(cfg), ctypes.byref(o),
                                       ctypes.byref(n_out), ctypes.byref(total),
                                       ctypes.c_void_p(stream or 0)))
    return n_out.value, total.value


def gen_gnm_host(n, m, max_weight=254, seed=1):
    """Deterministic GNM on the host: (src, dst, weight) uint32 arrays, 2m records."""
    s = np.empty(2 * m, np.uint32)
    d = np.empty(2 * m, np.uint32)
    w = np.empty(2 * m, np.uint32)
    _check(load().kmst_gen_gnm_host(n, m, max_weight, seed, s.ctypes.data, d.ctypes.data,
                                    w.ctypes.data))
    return s, d, w


def _gen_sized(call, guess):
    cap = max(int(guess), 16)
    while True:
        s, d, w = (np.empty(cap, np.uint32) for _ in range(3))
        m = ctypes.c_uint64(0)
        st = call(s.ctypes.data, d.ctypes.data, w.ctypes.data, cap, ctypes.byref(m))
        if st == KMST_OK:
            return s[: m.value], d[: m.value], w[: m.value]
        if st == KMST_EINVAL and m.value > cap:
            cap = m.value
            continue
        _check(st)


def gen_rgg2d_host(n, avg_degree=16.0, max_weight=(1 << 31) - 1, seed=3):
    """Random geometric graph in the unit square (deterministic)."""
    L = load()
    return _gen_sized(lambda s, d, w, cap, m: L.kmst_gen_rgg2d_host(
        n, float(avg_degree), max_weight, seed, s, d, w, cap, m), n * avg_degree * 1.2)


def gen_rhg_host(n, avg_degree=32.0, gamma=2.8, max_weight=254, seed=4):
    """Threshold random hyperbolic graph (deterministic, double precision)."""
    L = load()
    return _gen_sized(lambda s, d, w, cap, m: L.kmst_gen_rhg_host(
        n, float(avg_degree), float(gamma), max_weight, seed, s, d, w, cap, m),
        n * avg_degree * 1.5)


def encode_host(src, dst, weight, chunk_records=32):
    """Difference encoding of compression level 1 (kmst_encode_host): returns
    (bytes uint8, offsets uint64 with chunks + 1 entries, first uint32 of
    3 * chunks entries)."""
    L = load()
    s, d, w = (np.ascontiguousarray(a, dtype=np.uint32) for a in (src, dst, weight))
    m = len(s)
    used = ctypes.c_uint64(0)
    _check(L.kmst_encode_host(s.ctypes.data, d.ctypes.data, w.ctypes.data, m, chunk_records,
                              None, 0, None, None, ctypes.byref(used)))
    chunks = (m + chunk_records - 1) // chunk_records
    b = np.zeros(max(used.value, 1), dtype=np.uint8)
    off = np.zeros(chunks + 1, dtype=np.uint64)
    first = np.zeros(max(3 * chunks, 1), dtype=np.uint32)
    _check(L.kmst_encode_host(s.ctypes.data, d.ctypes.data, w.ctypes.data, m, chunk_records,
                              b.ctypes.data, len(b), off.ctypes.data, first.ctypes.data,
                              ctypes.byref(used)))
    return b[:used.value], off, first[:3 * chunks]


def to_wedges(src, dst, weight):
    """Pack SoA arrays into WEDGE_DTYPE records."""
    e = np.zeros(len(src), dtype=WEDGE_DTYPE)
    e["src"], e["dst"], e["weight"] = src, dst, weight
    return e


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr() if hasattr(t, "data_ptr") else int(t))


class Context:
    """Owns a libkmst context (stream + cached device workspaces) on one GPU.

    Device arrays are torch tensors (int32 / int64 storage reinterpreted as
    uint32 / uint64) or raw device addresses."""

    def __init__(self, device=0):
        self._lib = load()
        self._h = ctypes.c_void_p()
        _check(self._lib.kmst_context_create(device, ctypes.byref(self._h)))
        self.device = device
        _track(self)

    def close(self):
        if self._h:
            self._lib.kmst_context_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reserve(self, m_directed, n):
        """Pre-size the workspaces for solves up to this size."""
        _check(self._lib.kmst_context_reserve(self._h, m_directed, n))

    def set_timing(self, on=True):
        self._lib.kmst_context_set_timing(self._h, 1 if on else 0)

    def stats(self):
        need = self._lib.kmst_context_stats_json(self._h, None, 0)
        buf = ctypes.create_string_buffer(need + 1)
        self._lib.kmst_context_stats_json(self._h, buf, need + 1)
        return json.loads(buf.value.decode())

    def solve_device(self, src, dst, weight, m, n, out_src, out_dst, out_weight, capacity,
                     algorithm=BORUVKA, config=None):
        """Device-resident solve; returns (number of MST edges, u64 weight)."""
        cfg = (config or AlgorithmConfig(device=self.device))._c(algorithm)
        e = _DevEdges(_ptr(src).value, _ptr(dst).value, _ptr(weight).value, m, n)
        o = _DevMst(_ptr(out_src).value, _ptr(out_dst).value, _ptr(out_weight).value, capacity)
        n_out, total = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(self._lib.kmst_solve_device(self._h, ctypes.byref(e), ctypes.byref(cfg),
                                           ctypes.byref(o), ctypes.byref(n_out),
                                           ctypes.byref(total)))
        return n_out.value, total.value

    def set_tuning(self, tuning):
        t = tuning._c()
        _check(self._lib.kmst_context_set_tuning(self._h, ctypes.byref(t)))

    def get_tuning(self):
        t = _Tuning()
        _check(self._lib.kmst_context_get_tuning(self._h, ctypes.byref(t)))
        return Tuning._from(t)

    def autotune(self, src, dst, weight, m, n, algorithm=FILTER_BORUVKA, config=None, repeats=3):
        """Times candidate filter thresholds and sort launch shapes on this
        device-resident input (kmst_autotune); the fastest tuning stays set on
        the context.  Returns {"tuning", "filter_threshold", "best_ms",
        "candidates"}."""
        cfg = (config or AlgorithmConfig(device=self.device))._c(algorithm)
        e = _DevEdges(_ptr(src).value, _ptr(dst).value, _ptr(weight).value, m, n)
        t = _Tuning()
        best = ctypes.c_double(0.0)
        cand = ctypes.c_uint32(0)
        _check(self._lib.kmst_autotune(self._h, ctypes.byref(e), ctypes.byref(cfg), repeats,
                                       ctypes.byref(t), ctypes.byref(best), ctypes.byref(cand)))
        return {"tuning": Tuning._from(t), "filter_threshold": int(cfg.filter_threshold),
                "best_ms": best.value, "candidates": int(cand.value)}

    def gen_gnm_device(self, n, m, max_weight, seed, src, dst, weight):
        _check(self._lib.kmst_gen_gnm_device(self._h, n, m, max_weight, seed, _ptr(src),
                                             _ptr(dst), _ptr(weight)))

    def _gen_device_sized(self, call, guess):
        import torch
        cap = max(int(guess), 16)
        while True:
            arrs = [torch.empty(cap, dtype=torch.int32, device=f"cuda:{self.device}")
                    for _ in range(3)]
            m = ctypes.c_uint64(0)
            st = call(*(_ptr(a) for a in arrs), cap, ctypes.byref(m))
            if st == KMST_OK:
                return [a[: m.value] for a in arrs]
            if st == KMST_EINVAL and m.value > cap:
                del arrs
                cap = m.value
                continue
            _check(st)

    def gen_rgg2d_device(self, n, avg_degree=16.0, max_weight=(1 << 31) - 1, seed=3):
        """RGG2D generated on the device (kmst_gen_rgg2d_device): three int32
        tensors (src, dst, weight) on this context's device, identical to
        gen_rgg2d_host."""
        return self._gen_device_sized(lambda s, d, w, cap, m: self._lib.kmst_gen_rgg2d_device(
            self._h, n, float(avg_degree), max_weight, seed, s, d, w, cap, m),
            n * avg_degree * 1.1 + 1024)

    def gen_rhg_device(self, n, avg_degree=32.0, gamma=2.8, max_weight=254, seed=4):
        """RHG generated on the device (kmst_gen_rhg_device), identical to
        gen_rhg_host."""
        return self._gen_device_sized(lambda s, d, w, cap, m: self._lib.kmst_gen_rhg_device(
            self._h, n, float(avg_degree), float(gamma), max_weight, seed, s, d, w, cap, m),
            n * avg_degree * 1.2 + 1024)

    def segmented_argmin(self, src, dst, val, m, n, best_val, best_dst):
        _check(self._lib.kmst_k_segmented_argmin(self._h, _ptr(src), _ptr(dst), _ptr(val), m, n,
                                                 _ptr(best_val), _ptr(best_dst)))

    def segmented_argmin_packed(self, src, dst, weight, vbits, m, n, best_val, best_dst):
        _check(self._lib.kmst_k_segmented_argmin_packed(self._h, _ptr(src), _ptr(dst),
                                                        _ptr(weight), vbits, m, n,
                                                        _ptr(best_val), _ptr(best_dst)))

    def sort_pairs(self, src, dst, val, m, bits):
        _check(self._lib.kmst_k_sort_pairs(self._h, _ptr(src), _ptr(dst), _ptr(val), m, bits))

    def pointer_jump(self, parent, n):
        _check(self._lib.kmst_k_pointer_jump(self._h, _ptr(parent), n))


class Server:
    """Batched MST serving on one GPU (kmst_server_*).

    submit() takes host uint32 arrays (src, dst, weight; the kmst_boruvka
    input contract) and returns a ticket; wait() returns the result dict and
    the (min, max, weight) arrays.  The arrays are kept alive until the ticket
    completes.  max_batch_requests=1 gives the one-graph-per-launch path."""

    def __init__(self, device=0, algorithm=FILTER_BORUVKA, filter_threshold=4,
                 max_batch_records=None, max_batch_requests=None, hbm_budget_bytes=0,
                 queue_capacity=None, buffer_sets=3, tuning=None):
        self._lib = load()
        cfg = _ServerConfig()
        self._lib.kmst_server_config_default(ctypes.byref(cfg))
        cfg.device, cfg.algorithm, cfg.filter_threshold = device, algorithm, filter_threshold
        if tuning is not None:
            cfg.tuning = tuning._c()
        if max_batch_records is not None:
            cfg.max_batch_records = max_batch_records
        if max_batch_requests is not None:
            cfg.max_batch_requests = max_batch_requests
        cfg.hbm_budget_bytes = hbm_budget_bytes
        if queue_capacity is not None:
            cfg.queue_capacity = queue_capacity
        cfg.buffer_sets = buffer_sets
        self._h = ctypes.c_void_p()
        _check(self._lib.kmst_server_create(ctypes.byref(cfg), ctypes.byref(self._h)))
        self._live = {}
        _track(self)

    def submit(self, src, dst, weight, n, out=None):
        """src/dst/weight: numpy uint32 arrays (host) or torch tensors (host
        or device); out: optional three arrays of >= max(n, 1) 32-bit entries
        that receive the result: numpy arrays (e.g. views of a pinned pool)
        or torch tensors (pinned host or device; the server's kernel writes
        device-writable outputs directly)."""
        def arr(x):
            if hasattr(x, "data_ptr"):
                return x
            return np.ascontiguousarray(x, dtype=np.uint32)

        def ptr(x):
            return x.data_ptr() if hasattr(x, "data_ptr") else x.ctypes.data

        src, dst, weight = arr(src), arr(dst), arr(weight)
        cap = max(int(n), 1)
        outs = list(out) if out is not None else [np.empty(cap, np.uint32) for _ in range(3)]
        assert all(len(o) >= cap for o in outs)
        t = ctypes.c_uint64(0)
        _check(self._lib.kmst_server_submit(
            self._h, ptr(src), ptr(dst), ptr(weight), len(src), int(n),
            ptr(outs[0]), ptr(outs[1]), ptr(outs[2]), cap, ctypes.byref(t)))
        self._live[t.value] = (src, dst, weight, outs)
        return t.value

    def wait(self, ticket):
        r = _RequestResult()
        _check(self._lib.kmst_server_wait(self._h, ticket, ctypes.byref(r)))
        _, _, _, outs = self._live.pop(ticket)
        k = r.mst_edges if r.status == KMST_OK else 0
        res = {f: getattr(r, f) for f, _ in _RequestResult._fields_}
        return res, (outs[0][:k], outs[1][:k], outs[2][:k])

    def counters(self):
        b, q = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(self._lib.kmst_server_counters(self._h, ctypes.byref(b), ctypes.byref(q)))
        return b.value, q.value

    def close(self):
        if self._h:
            self._lib.kmst_server_destroy(self._h)
            self._h = ctypes.c_void_p()
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
