"""Python binding of libpumatch.so, the MI355X subgraph matcher.

The library's C ABI (include/pumatch.h) is the product boundary; this module
is a thin ctypes wrapper used by the tests and the benchmark.  It loads only
the in-tree build (pumatch_amd/lib/libpumatch.so) and raises when it is
missing: there is no CPU fallback.

Host-side names follow the reference driver (test.cu): a Graph is what
GraphPreprocessor loads (src/graph.h:228-254), a Pattern is what
PatternPreprocessor plans (src/pattern.h:48-58), and match() is the timed
launch of test.cu:171-200 returning (kernel ms, count).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpumatch.so")
CLI_PATH = os.path.join(_HERE, "bin", "test")

PM_GRAPH_LABELS = 0x1
PM_GRAPH_HOST_ONLY = 0x2
PM_LABELED, PM_VERTEX_INDUCED = 0x1, 0x2  # pattern modes
PM_OK, PM_ERR_IO, PM_ERR_FORMAT, PM_ERR_PATTERN = 0, -1, -2, -3
PM_ERR_DEVICE, PM_ERR_ARG, PM_ERR_OVERFLOW, PM_ERR_NOMEM = -4, -5, -6, -7
MAX_PATTERN = 7


class PumatchError(RuntimeError):
    pass


class Opts(C.Structure):
    _fields_ = [
        ("waves", C.c_int32),
        ("chunk", C.c_int32),
        ("repeats", C.c_int32),
        ("split", C.c_int32),
        ("hub_factor", C.c_int32),
        ("policy", C.c_int32),
        ("edge_lo", C.c_uint64),
        ("edge_hi", C.c_uint64),
        ("per_edge", C.POINTER(C.c_uint64)),
        ("stream", C.c_void_p),
        ("cache_bytes", C.c_uint64),
        ("package_bytes", C.c_uint32),
        ("prefetch", C.c_int32),
        ("order", C.c_int32),
        ("sample", C.c_int32),
        ("learned", C.c_int32),
        ("model", C.c_void_p),
        ("root_work", C.POINTER(C.c_uint64)),
        ("package_use", C.POINTER(C.c_uint32)),
        ("direct_degree", C.c_int32),
        ("graph", C.c_int32),
        ("resplit_us", C.c_int32),
    ]


class TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float),
                ("pad", C.c_int32), ("seed", C.c_uint64)]


class TuneOpts(C.Structure):
    _fields_ = [("sample", C.c_int32), ("max_orders", C.c_int32), ("repeats", C.c_int32)]


class TuneReport(C.Structure):
    _fields_ = [("base_ms", C.c_double), ("best_ms", C.c_double), ("trials", C.c_int32),
                ("reordered", C.c_int32), ("sample", C.c_int32), ("pad", C.c_int32)]


class Result(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("min_ms", C.c_double),
        ("count", C.c_uint64),
        ("jobs", C.c_uint64),
        ("alg_bytes", C.c_uint64),
        ("set_ops", C.c_uint64),
        ("link_bytes", C.c_uint64),
        ("tasks", C.c_uint64),
        ("waves", C.c_int32),
        ("chunk", C.c_int32),
        ("phase_ms", C.c_double * 3),
        ("phase_alg_bytes", C.c_uint64 * 3),
        ("misses", C.c_uint64),
        ("deferred", C.c_uint64),
        ("rounds", C.c_int32),
        ("packages", C.c_int32),
        ("link_ms", C.c_double),
        ("round_kernel_ms", C.c_double),
    ]


class GraphInfo(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("m", C.c_uint64),
        ("max_degree", C.c_uint32),
        ("has_labels", C.c_int32),
        ("on_device", C.c_int32),
        ("reserved", C.c_int32),
        ("device_bytes", C.c_uint64),
    ]


class PatternInfo(C.Structure):
    _fields_ = [
        ("k", C.c_int32),
        ("labeled", C.c_int32),
        ("automorphisms", C.c_int64),
        ("rc_order", C.c_int32),
        ("nslots", C.c_int32),
        ("degree", C.c_int32 * MAX_PATTERN),
        ("label", C.c_int32 * MAX_PATTERN),
        ("sym_mask", C.c_uint32 * MAX_PATTERN),
        ("nbr_mask", C.c_uint32 * MAX_PATTERN),
        ("vertex_induced", C.c_int32),
    ]


# every symbol include/pumatch.h declares, with its ctypes signature
_SIGS = {
    "pm_last_error": (C.c_char_p, []),
    "pm_version": (C.c_char_p, []),
    "pm_graph_open": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "pm_graph_from_csr": (C.c_int, [C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_uint32, C.POINTER(C.c_void_p)]),
    "pm_graph_generate_rmat": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_uint64, C.c_uint32,
                                         C.POINTER(C.c_void_p)]),
    "pm_graph_generate_chunglu": (C.c_int, [C.c_uint64, C.c_double, C.c_double, C.c_double,
                                            C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "pm_graph_set_uniform_labels": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64]),
    "pm_graph_write": (C.c_int, [C.c_void_p, C.c_char_p]),
    "pm_graph_info_get": (C.c_int, [C.c_void_p, C.POINTER(GraphInfo)]),
    "pm_graph_host_csr": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_void_p)]),
    "pm_graph_to_device": (C.c_int, [C.c_void_p]),
    "pm_graph_destroy": (None, [C.c_void_p]),
    "pm_pattern_load": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "pm_pattern_parse": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "pm_pattern_info_get": (C.c_int, [C.c_void_p, C.POINTER(PatternInfo)]),
    "pm_pattern_destroy": (None, [C.c_void_p]),
    "pm_opts_default": (None, [C.POINTER(Opts)]),
    "pm_match": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(Result)]),
    "pm_match_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int32, C.POINTER(Opts),
                                 C.POINTER(Result)]),
    "pm_model_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "pm_model_destroy": (None, [C.c_void_p]),
    "pm_model_get": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pm_model_set": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pm_graph_features": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "pm_model_predict": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_uint64,
                                   C.POINTER(C.c_float)]),
    "pm_model_grad": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                C.POINTER(C.c_float), C.c_uint64, C.POINTER(C.c_float),
                                C.POINTER(C.c_double)]),
    "pm_model_train": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.POINTER(C.c_float), C.c_uint64, C.POINTER(TrainOpts),
                                 C.POINTER(C.c_double)]),
    "pm_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int32, C.POINTER(Opts),
                              C.POINTER(TuneOpts), C.POINTER(Opts), C.POINTER(TuneReport)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libpumatch.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PumatchError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().pm_last_error()
        raise PumatchError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


class Graph:
    """A CSR graph owned by the library (host copy + HBM copy)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    # -- constructors --------------------------------------------------------
    @classmethod
    def open(cls, prefix: str, labels: bool = True, device: bool = True) -> "Graph":
        h = C.c_void_p()
        flags = (PM_GRAPH_LABELS if labels else 0) | (0 if device else PM_GRAPH_HOST_ONLY)
        _check(lib().pm_graph_open(prefix.encode(), flags, C.byref(h)), f"open {prefix}")
        return cls(h.value)

    @classmethod
    def from_csr(cls, rowptr, colidx, labels=None, device: bool = True) -> "Graph":
        rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
        colidx = np.ascontiguousarray(colidx, dtype=np.uint32)
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.uint32)
        n = rowptr.size - 1
        h = C.c_void_p()
        flags = 0 if device else PM_GRAPH_HOST_ONLY
        _check(lib().pm_graph_from_csr(n, colidx.size, rowptr.ctypes.data,
                                       colidx.ctypes.data if colidx.size else None,
                                       None if lab is None else lab.ctypes.data, flags,
                                       C.byref(h)), "from_csr")
        return cls(h.value)

    @classmethod
    def rmat(cls, scale: int, edge_factor: float = 16.0, a=0.57, b=0.19, c=0.19, seed: int = 1,
             device: bool = True) -> "Graph":
        h = C.c_void_p()
        flags = 0 if device else PM_GRAPH_HOST_ONLY
        _check(lib().pm_graph_generate_rmat(scale, edge_factor, a, b, c, seed, flags,
                                            C.byref(h)), "rmat")
        return cls(h.value)

    @classmethod
    def chung_lu(cls, n: int, gamma: float, dmax: float, avg: float, seed: int,
                 device: bool = True) -> "Graph":
        h = C.c_void_p()
        flags = 0 if device else PM_GRAPH_HOST_ONLY
        _check(lib().pm_graph_generate_chunglu(n, gamma, dmax, avg, seed, flags, C.byref(h)),
               "chung_lu")
        return cls(h.value)

    # -- operations ----------------------------------------------------------
    def set_uniform_labels(self, nlabels: int, seed: int) -> None:
        _check(lib().pm_graph_set_uniform_labels(self._h, nlabels, seed), "labels")

    def write(self, prefix: str) -> None:
        _check(lib().pm_graph_write(self._h, prefix.encode()), f"write {prefix}")

    def to_device(self) -> None:
        _check(lib().pm_graph_to_device(self._h), "to_device")

    def info(self) -> GraphInfo:
        gi = GraphInfo()
        _check(lib().pm_graph_info_get(self._h, C.byref(gi)), "info")
        return gi

    @property
    def n(self) -> int:
        return int(self.info().n)

    @property
    def m(self) -> int:
        return int(self.info().m)

    def csr(self):
        """Zero-copy numpy views of the host CSR (rowptr u64, colidx u32,
        labels u32 or None), valid while the graph lives."""
        gi = self.info()
        rp, ci, lb = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib().pm_graph_host_csr(self._h, C.byref(rp), C.byref(ci), C.byref(lb)), "csr")
        rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_uint64)), shape=(gi.n + 1,))
        colidx = (np.ctypeslib.as_array(C.cast(ci, C.POINTER(C.c_uint32)), shape=(gi.m,))
                  if gi.m else np.zeros(0, np.uint32))
        labels = (np.ctypeslib.as_array(C.cast(lb, C.POINTER(C.c_uint32)), shape=(gi.n,))
                  if lb.value else None)
        return rowptr, colidx, labels

    def close(self) -> None:
        if self._h is not None and self._h.value:
            lib().pm_graph_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pattern:
    """A planned query pattern (.g)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @staticmethod
    def _mode(labeled: bool, induced: bool) -> int:
        return (PM_LABELED if labeled else 0) | (PM_VERTEX_INDUCED if induced else 0)

    @classmethod
    def load(cls, path: str, labeled: bool = True, induced: bool = False) -> "Pattern":
        """induced=True: vertex-induced matches (EDGE_INDUCED=false)."""
        h = C.c_void_p()
        _check(lib().pm_pattern_load(path.encode(), cls._mode(labeled, induced), C.byref(h)),
               f"pattern {path}")
        return cls(h.value)

    @classmethod
    def parse(cls, text: str, labeled: bool = True, induced: bool = False) -> "Pattern":
        h = C.c_void_p()
        _check(lib().pm_pattern_parse(text.encode(), cls._mode(labeled, induced), C.byref(h)),
               "pattern")
        return cls(h.value)

    def info(self) -> PatternInfo:
        pi = PatternInfo()
        _check(lib().pm_pattern_info_get(self._h, C.byref(pi)), "pattern info")
        return pi

    def close(self) -> None:
        if self._h is not None and self._h.value:
            lib().pm_pattern_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class MatchResult:
    kernel_ms: float
    min_ms: float
    count: int
    jobs: int
    alg_bytes: int
    set_ops: int
    link_bytes: int
    tasks: int
    waves: int
    chunk: int
    phase_ms: tuple = ()
    phase_alg_bytes: tuple = ()
    misses: int = 0
    deferred: int = 0
    rounds: int = 1
    packages: int = 0
    link_ms: float = 0.0
    per_edge: np.ndarray | None = None
    round_kernel_ms: float = 0.0
    root_work: np.ndarray | None = None
    package_use: np.ndarray | None = None


ORDER = {"cost": 0, "file": 1}
PM_LEARN_ORDER, PM_LEARN_PAGING = 0x1, 0x2
N_FEATURES, N_PARAMS = 16, 5378


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Model:
    """The 16 -> 64 -> 64 -> 2 MLP (pm_model)."""

    def __init__(self, seed: int = 1):
        h = C.c_void_p()
        _check(lib().pm_model_create(seed, C.byref(h)), "model_create")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().pm_model_destroy(self._h)
            self._h = C.c_void_p()

    @property
    def params(self) -> np.ndarray:
        w = np.zeros(N_PARAMS, np.float32)
        _check(lib().pm_model_get(self._h, _fptr(w)), "model_get")
        return w

    @params.setter
    def params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == N_PARAMS
        _check(lib().pm_model_set(self._h, _fptr(w)), "model_set")

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, N_FEATURES)
        Y = np.zeros((X.shape[0], 2), np.float32)
        _check(lib().pm_model_predict(self._h, _fptr(X), X.shape[0], _fptr(Y)), "model_predict")
        return Y

    def grad(self, X, Y, W=None):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, N_FEATURES)
        Y = np.ascontiguousarray(Y, dtype=np.float32).reshape(-1, 2)
        Wp = None
        if W is not None:
            W = np.ascontiguousarray(W, dtype=np.float32).reshape(-1, 2)
            Wp = _fptr(W)
        g = np.zeros(N_PARAMS, np.float32)
        loss = C.c_double()
        _check(lib().pm_model_grad(self._h, _fptr(X), _fptr(Y), Wp, X.shape[0], _fptr(g),
                                   C.byref(loss)), "model_grad")
        return loss.value, g

    def train(self, X, Y, W=None, *, epochs: int = 0, batch: int = 0, lr: float = 0.0,
              seed: int = 0) -> float:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, N_FEATURES)
        Y = np.ascontiguousarray(Y, dtype=np.float32).reshape(-1, 2)
        Wp = None
        if W is not None:
            W = np.ascontiguousarray(W, dtype=np.float32).reshape(-1, 2)
            Wp = _fptr(W)
        t = TrainOpts(epochs, batch, lr, 0, seed)
        loss = C.c_double()
        _check(lib().pm_model_train(self._h, _fptr(X), _fptr(Y), Wp, X.shape[0], C.byref(t),
                                    C.byref(loss)), "model_train")
        return loss.value


def features(g: "Graph") -> np.ndarray:
    """Per-vertex features (n x 16) computed on the GPU (pm_graph_features)."""
    out = np.zeros((max(g.n, 1), N_FEATURES), np.float32)
    _check(lib().pm_graph_features(g._h, _fptr(out)), "graph_features")
    return out[:g.n]
POLICY = {"lru": 0, "fifo": 1, "learned": 2}


def _make_opts(g: Graph, waves, chunk, repeats, split, hub_factor, edge_lo, edge_hi, per_edge,
               stream, cache_bytes, package_bytes, policy, prefetch, order="cost", sample=0,
               model=None, learned=0, direct_degree=0, graph=False, resplit_us=0):
    o = Opts()
    lib().pm_opts_default(C.byref(o))
    o.waves, o.chunk, o.repeats = waves, chunk, repeats
    o.split, o.hub_factor = split, hub_factor
    o.cache_bytes, o.package_bytes = cache_bytes, package_bytes
    o.policy, o.prefetch = POLICY[policy], 1 if prefetch else 0
    o.edge_lo, o.edge_hi = edge_lo, edge_hi
    o.order = ORDER[order]
    o.sample = sample
    o.direct_degree = direct_degree
    o.graph = 1 if graph else 0
    o.resplit_us = resplit_us
    if model is not None:
        o.model = model._h.value
        o.learned = learned if learned else (PM_LEARN_ORDER | PM_LEARN_PAGING)
    pe = None
    if per_edge:
        pe = np.zeros(g.m, dtype=np.uint64)
        o.per_edge = pe.ctypes.data_as(C.POINTER(C.c_uint64))
    if stream is not None:
        o.stream = stream
    return o, pe


def _result(r: Result, pe=None) -> MatchResult:
    return MatchResult(r.kernel_ms, r.min_ms, int(r.count), int(r.jobs), int(r.alg_bytes),
                       int(r.set_ops), int(r.link_bytes), int(r.tasks), int(r.waves), int(r.chunk),
                       tuple(r.phase_ms), tuple(int(x) for x in r.phase_alg_bytes),
                       int(r.misses), int(r.deferred), int(r.rounds), int(r.packages),
                       float(r.link_ms), pe, float(r.round_kernel_ms))


def match(g: Graph, p: Pattern, *, waves: int = 0, chunk: int = 0, repeats: int = 1,
          split: int = 0, hub_factor: int = 0, edge_lo: int = 0, edge_hi: int = 0,
          per_edge: bool = False, stream: int | None = None, cache_bytes: int = 0,
          package_bytes: int = 0, policy: str = "lru", prefetch: bool = False,
          order: str = "cost", sample: int = 0, model=None, learned: int = 0,
          traces: bool = False, direct_degree: int = 0, graph: bool = False,
          resplit_us: int = 0) -> MatchResult:
    """Count matches of p in g on the GPU (pm_match).  order="file" matches
    labeled patterns in file order as the reference does; "cost" (default)
    lets the planner pick the order of vertices 2.. (same count).  With a
    Model, `learned` (PM_LEARN_* bits, default both) turns on learned job
    order and package ranking.  traces=True records root_work and
    package_use (both per vertex) into the result.  direct_degree (out of
    core): missing rows of at most that many entries are read from pinned
    host memory instead of being deferred (0: 1024, < 0: never).  graph=True
    replays the run's launch sequence as a captured hipGraph (resident rows;
    phase_ms is then not measured)."""
    o, pe = _make_opts(g, waves, chunk, repeats, split, hub_factor, edge_lo, edge_hi, per_edge,
                       stream, cache_bytes, package_bytes, policy, prefetch, order, sample,
                       model, learned, direct_degree, graph, resplit_us)
    rw = pu = None
    if traces:
        rw = np.zeros(max(g.n, 1), dtype=np.uint64)
        pu = np.zeros(max(g.n, 1), dtype=np.uint32)
        o.root_work = rw.ctypes.data_as(C.POINTER(C.c_uint64))
        o.package_use = pu.ctypes.data_as(C.POINTER(C.c_uint32))
    r = Result()
    _check(lib().pm_match(g._h, p._h, C.byref(o), C.byref(r)), "match")
    res = _result(r, pe)
    if traces:
        res.root_work = rw[:g.n]
        res.package_use = pu[:g.n]
    return res


def match_batch(g: Graph, patterns, *, waves: int = 0, chunk: int = 0, repeats: int = 1,
                split: int = 0, hub_factor: int = 0, edge_lo: int = 0, edge_hi: int = 0,
                stream: int | None = None, cache_bytes: int = 0, package_bytes: int = 0,
                policy: str = "lru", prefetch: bool = False, order: str = "cost",
                sample: int = 0, model=None, learned: int = 0, direct_degree: int = 0,
                graph: bool = False, resplit_us: int = 0) -> list:
    """Count several patterns in one persistent launch (pm_match_batch).
    Returns one MatchResult per pattern; count/jobs are the pattern's own,
    the timing fields describe the whole launch."""
    patterns = list(patterns)
    n = len(patterns)
    o, _ = _make_opts(g, waves, chunk, repeats, split, hub_factor, edge_lo, edge_hi, False,
                      stream, cache_bytes, package_bytes, policy, prefetch, order, sample,
                      model, learned, direct_degree, graph, resplit_us)
    hs = (C.c_void_p * max(n, 1))(*[q._h.value for q in patterns])
    rs = (Result * max(n, 1))()
    _check(lib().pm_match_batch(g._h, hs, n, C.byref(o), rs), "match_batch")
    return [_result(rs[i]) for i in range(n)]


def autotune(g: Graph, patterns, *, sample: int = 0, max_orders: int = 0, repeats: int = 0,
             **base):
    """pm_autotune: time sampled launches and return (options, report).
    The options are a dict for match()/match_batch(); the patterns remember
    their measured-fastest core order for this graph."""
    patterns = list(patterns)
    n = len(patterns)
    o, _ = _make_opts(g, base.get("waves", 0), base.get("chunk", 0), 1, base.get("split", 0),
                      base.get("hub_factor", 0), base.get("edge_lo", 0), base.get("edge_hi", 0),
                      False, base.get("stream"), base.get("cache_bytes", 0),
                      base.get("package_bytes", 0), base.get("policy", "lru"),
                      base.get("prefetch", False), base.get("order", "cost"))
    t = TuneOpts(sample, max_orders, repeats)
    best = Opts()
    rep = TuneReport()
    hs = (C.c_void_p * max(n, 1))(*[q._h.value for q in patterns])
    _check(lib().pm_autotune(g._h, hs, n, C.byref(o), C.byref(t), C.byref(best), C.byref(rep)),
           "autotune")
    opts = dict(base)
    opts.update(chunk=best.chunk, split=best.split, hub_factor=best.hub_factor)
    report = {"base_ms": rep.base_ms, "best_ms": rep.best_ms, "trials": rep.trials,
              "reordered": rep.reordered, "sample": rep.sample}
    return opts, report


def pattern_text(k: int, edges, labels=None) -> str:
    """Write a pattern in the reference's .g layout: '# t 0', 'v i label'
    lines, then every edge listed in both directions (pattern/*.g)."""
    labels = labels if labels is not None else [1] * k
    lines = ["# t 0"] + [f"v {i} {labels[i]}" for i in range(k)]
    for a, b in edges:
        lines.append(f"e {a} {b} 1")
        lines.append(f"e {b} {a} 1")
    return "\n".join(lines) + "\n"
