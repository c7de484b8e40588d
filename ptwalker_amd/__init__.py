"""ptwalker_amd -- MI355X probabilistic-termination random-walk engine.

Python mirror of the reference's in-process interface, over the C ABI of
``libptw.so`` (include/ptw.h):

* ``Graph`` -- ``Graph(MultiThreadConfig)`` + ``Graph::load(path, format,
  as_undirected, bodary)`` (src/core/graph.hpp:151,277); also a CSR handed
  in and the seeded R-MAT generator.
* ``walk(graph, walker_num, app, ...)`` -- ``walk(FMobSolver*, walker_num,
  walk_len, mem_quota, Alg&)`` (src/core/solver.hpp:1049-1074).
* apps -- the reference's Alg plug-ins as plain data: ``PageRank``
  (src/core/pagerank.cpp:10-90), ``PPR`` (src/core/ppr.cpp:10-119),
  ``RWR`` (src/core/RWR.cpp:15-107), ``DeepWalk`` (src/core/deepwalk.cpp:10-42).

There is no CPU fallback: importing works without a GPU, but every call
that walks needs the HIP library and a device, and fails loudly otherwise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libptw.so")
BIN_DIR = os.path.join(_HERE, "bin")

# status codes, apps, flags: include/ptw.h
OK = 0
ERR_INVALID_ARG = -1
ERR_IO = -2
ERR_HIP = -3
ERR_NO_GRAPH = -4
ERR_OOM = -5
ERR_DANGLING = -6
ERR_FORMAT = -7
ERR_UNSUPPORTED = -8
ERR_ADMISSION = -9
ERR_CANCELLED = -10
ERR_TIMEOUT = -11

FORMAT_BINARY = 0
FORMAT_TEXT = 1

APP_PAGERANK = 0
APP_PPR = 1
APP_RWR = 2
APP_DEEPWALK = 3
APP_NODE2VEC = 4

SEM_REFERENCE = 0
SEM_TEXTBOOK = 1

START_UNIFORM = 0
START_SOURCE = 1
START_ROUNDROBIN = 2

DANGLING_ERROR = 0
DANGLING_STOP = 1

FLAG_OUT_DEVICE = 1
FLAG_ACCUMULATE = 2
FLAG_WEIGHTED = 4

RMAT_KEEP_ISOLATED = 1
GRAPH_NO_RELABEL = 2
GRAPH_INGEST_HOST = 4
GRAPH_INGEST_DEVICE = 8

INVALID_VERTEX = 0xFFFFFFFF


class PtwError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"ptw error {code}: {message}")
        self.code = code


class GraphInfo(ctypes.Structure):
    _fields_ = [
        ("num_vertices", ctypes.c_uint32),
        ("max_degree", ctypes.c_uint32),
        ("num_edges", ctypes.c_uint64),
        ("max_degree_vertex", ctypes.c_uint32),
        ("num_dangling", ctypes.c_uint32),
        ("hub_vertices", ctypes.c_uint32),
        ("relabelled", ctypes.c_uint32),
        ("device_bytes", ctypes.c_uint64),
        ("weighted", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class WalkDesc(ctypes.Structure):
    _fields_ = [
        ("app", ctypes.c_int32),
        ("semantics", ctypes.c_int32),
        ("start_mode", ctypes.c_int32),
        ("dangling_policy", ctypes.c_int32),
        ("source", ctypes.c_uint32),
        ("walk_len", ctypes.c_uint32),
        ("tp", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("wid_base", ctypes.c_uint64),
        ("walkers", ctypes.c_uint64),
        ("stream", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("p", ctypes.c_double),
        ("q", ctypes.c_double),
    ]


class WalkOut(ctypes.Structure):
    _fields_ = [
        ("counts", ctypes.c_void_p),
        ("ends", ctypes.c_void_p),
        ("paths", ctypes.c_void_p),
        ("starts", ctypes.c_void_p),
        ("steps", ctypes.c_uint64),
        ("visits", ctypes.c_uint64),
        ("dangling", ctypes.c_uint64),
        ("kernel_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("launch_block", ctypes.c_uint32),
        ("launch_blocks", ctypes.c_uint32),
        ("launch_chunk", ctypes.c_uint32),
        ("launch_lds_hubs", ctypes.c_uint32),
        ("launch_cache_slots", ctypes.c_uint32),
        ("launch_visit_log", ctypes.c_uint32),
    ]

    def launch_shape(self) -> dict:
        """The launch the walk ran with (after the automatic rules)."""
        return {"block": self.launch_block, "blocks": self.launch_blocks, "chunk": self.launch_chunk,
                "lds_hubs": self.launch_lds_hubs, "cache_slots": self.launch_cache_slots,
                "visit_log": self.launch_visit_log}


CACHE_AUTO = -1


class Query(ctypes.Structure):
    _fields_ = [
        ("source", ctypes.c_uint32),
        ("stream", ctypes.c_uint32),
        ("walkers", ctypes.c_uint64),
        ("tp", ctypes.c_double),
        ("kind", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


IDS_GRAPH = 0     # outputs in graph ids
IDS_ORIGINAL = 1  # outputs in the input's own ids (ptw_graph_export_names)

QUERY_VISITS = 0     # PPR: visit counts
QUERY_ENDPOINTS = 1  # RWR: end-vertex counts


# the same layout as a numpy record, so query tables are filled without a
# Python loop
QUERY_DTYPE = np.dtype({"names": [f for f, _ in Query._fields_],
                        "formats": [np.uint32, np.uint32, np.uint64, np.float64, np.uint32, np.uint32],
                        "offsets": [getattr(Query, f).offset for f, _ in Query._fields_],
                        "itemsize": ctypes.sizeof(Query)})


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("queries", ctypes.c_void_p),
        ("num_queries", ctypes.c_uint32),
        ("topk", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("semantics", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("dense_threshold", ctypes.c_uint64),
    ]


class BatchOut(ctypes.Structure):
    _fields_ = [
        ("topk_vertex", ctypes.c_void_p),
        ("topk_count", ctypes.c_void_p),
        ("steps", ctypes.c_void_p),
        ("entropy", ctypes.c_void_p),
        ("total_steps", ctypes.c_uint64),
        ("slice_bytes", ctypes.c_uint64),
        ("dense_queries", ctypes.c_uint32),
        ("hash_queries", ctypes.c_uint32),
        ("kernel_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("logged_queries", ctypes.c_uint32),
        ("log_rerun", ctypes.c_uint32),
    ]


BATCH_GRAPH = 1
BATCH_WEIGHTED = 2


class Tuning(ctypes.Structure):
    _fields_ = [
        ("block", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("lds_hubs", ctypes.c_int32),
        ("cache_slots", ctypes.c_int32),
        ("batch_chunk", ctypes.c_int32),
        ("topk_tile", ctypes.c_int32),
        ("dense_chunk", ctypes.c_int32),
        ("visit_log", ctypes.c_int32),
        ("edge_records", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


BUDGET_FIXED = 0
BUDGET_LEARNED = 1


class Request(ctypes.Structure):
    _fields_ = [
        ("source", ctypes.c_uint32),
        ("id", ctypes.c_uint32),
        ("eps", ctypes.c_float),
        ("reserved", ctypes.c_uint32),
        ("arrival_s", ctypes.c_double),
    ]


class ServeConfig(ctypes.Structure):
    _fields_ = [
        ("tp", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("budget_mode", ctypes.c_int32),
        ("topk", ctypes.c_uint32),
        ("fixed_c2", ctypes.c_double),
        ("margin", ctypes.c_double),
        ("w_min", ctypes.c_uint64),
        ("w_max", ctypes.c_uint64),
        ("feat_lo", ctypes.c_float * 6),
        ("feat_hi", ctypes.c_float * 6),
        ("audit_rate", ctypes.c_double),
        ("max_batch", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("hbm_limit", ctypes.c_uint64),
    ]


class RequestResult(ctypes.Structure):
    _fields_ = [
        ("budget", ctypes.c_uint64),
        ("admit_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("fallback", ctypes.c_uint32),
        ("audited", ctypes.c_uint32),
        ("topped_up", ctypes.c_uint32),
        ("launches", ctypes.c_uint32),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("makespan_s", ctypes.c_double),
        ("queries_per_s", ctypes.c_double),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double),
        ("walk_steps", ctypes.c_uint64),
        ("launches", ctypes.c_uint32),
        ("fallbacks", ctypes.c_uint32),
        ("audited", ctypes.c_uint32),
        ("topped_up", ctypes.c_uint32),
        ("admission_waits", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


# every symbol include/ptw.h declares, with its ctypes signature
_P = ctypes.c_void_p
_SIGNATURES = {
    "ptw_version": (ctypes.c_int, []),
    "ptw_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int]),
    "ptw_destroy": (None, [_P]),
    "ptw_last_error": (ctypes.c_char_p, [_P]),
    "ptw_graph_load": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32]),
    "ptw_graph_from_csr": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32]),
    "ptw_graph_rmat": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32]),
    "ptw_graph_set_weights": (ctypes.c_int, [_P, _P, ctypes.c_uint64]),
    "ptw_graph_random_weights": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "ptw_graph_export_weights": (ctypes.c_int, [_P, _P, _P, _P]),
    "ptw_graph_info_get": (ctypes.c_int, [_P, ctypes.POINTER(GraphInfo)]),
    "ptw_graph_export_csr": (ctypes.c_int, [_P, _P, _P]),
    "ptw_graph_export_names": (ctypes.c_int, [_P, _P]),
    "ptw_graph_write": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int]),
    "ptw_walk": (ctypes.c_int, [_P, ctypes.POINTER(WalkDesc), ctypes.POINTER(WalkOut)]),
    "ptw_walk_static": (ctypes.c_int, [_P, ctypes.POINTER(WalkDesc), ctypes.POINTER(WalkOut)]),
    "ptw_batch_ppr": (ctypes.c_int, [_P, ctypes.POINTER(BatchDesc), ctypes.POINTER(BatchOut)]),
    "ptw_batch_plan_bytes": (ctypes.c_int, [_P, ctypes.POINTER(BatchDesc), ctypes.POINTER(ctypes.c_uint64)]),
    "ptw_batch_counts": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
    "ptw_batch_gather": (ctypes.c_int, [_P, ctypes.c_uint64, _P, _P, _P]),
    "ptw_batch_write_topk": (ctypes.c_int, [_P, ctypes.c_char_p, _P, _P, ctypes.c_int]),
    "ptw_ppr_exact": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, ctypes.c_int32, ctypes.c_double, _P,
                                     ctypes.c_uint32, _P, _P, ctypes.POINTER(ctypes.c_uint32)]),
    "ptw_mlp_create": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "ptw_mlp_destroy": (None, [_P]),
    "ptw_mlp_num_params": (ctypes.c_int, []),
    "ptw_mlp_set_normalization": (ctypes.c_int, [_P, _P, _P, ctypes.c_float, ctypes.c_float]),
    "ptw_mlp_get_params": (ctypes.c_int, [_P, _P, _P]),
    "ptw_mlp_set_params": (ctypes.c_int, [_P, _P]),
    "ptw_mlp_forward": (ctypes.c_int, [_P, _P, ctypes.c_uint32, _P]),
    "ptw_mlp_gradients": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P, ctypes.POINTER(ctypes.c_float)]),
    "ptw_mlp_train": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                     ctypes.c_uint64, ctypes.POINTER(ctypes.c_float)]),
    "ptw_serve_replay": (ctypes.c_int, [_P, _P, ctypes.POINTER(ServeConfig), _P, ctypes.c_uint32, _P, _P, _P,
                                        ctypes.POINTER(ServeStats)]),
    "ptw_tuning_get": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "ptw_tuning_set": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "ptw_buffer_alloc": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "ptw_buffer_free": (ctypes.c_int, [_P, _P]),
    "ptw_buffer_upload": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64]),
    "ptw_buffer_download": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64]),
    "ptw_buffer_zero": (ctypes.c_int, [_P, _P, ctypes.c_uint64]),
    "ptw_device_memory": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    "ptw_bench_step": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, _P, _P]),
    "ptw_check_primitives": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64]),
    "ptw_bench_gather": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "ptw_bench_copy": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]),
}

_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libptw.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return sorted(_SIGNATURES)


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data


class Context:
    """One HIP device, one stream, one resident graph (ptw_create)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _P()
        rc = self.lib.ptw_create(ctypes.byref(h), device)
        if rc != OK:
            raise PtwError(rc, f"cannot create a context on HIP device {device}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.ptw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def check(self, rc: int):
        if rc != OK:
            raise PtwError(rc, self.lib.ptw_last_error(self.h).decode(errors="replace"))
        return rc

    # ---- graph -----------------------------------------------------------
    def load(self, path: str, fmt: int = FORMAT_TEXT, as_undirected: bool = False, flags: int = 0):
        self.check(self.lib.ptw_graph_load(self.h, path.encode(), fmt, int(as_undirected), flags))
        return self.info()

    def from_csr(self, row_ptr: np.ndarray, col: np.ndarray, flags: int = 0):
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        self.check(self.lib.ptw_graph_from_csr(self.h, _ptr(row_ptr), _ptr(col), len(row_ptr) - 1, len(col), flags))
        return self.info()

    def rmat(self, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1, flags: int = 0):
        self.check(self.lib.ptw_graph_rmat(self.h, scale, edge_factor, a, b, c, seed, flags))
        return self.info()

    def info(self) -> GraphInfo:
        gi = GraphInfo()
        self.check(self.lib.ptw_graph_info_get(self.h, ctypes.byref(gi)))
        return gi

    def export_csr(self):
        gi = self.info()
        row_ptr = np.empty(gi.num_vertices + 1, dtype=np.uint32)
        col = np.empty(max(gi.num_edges, 1), dtype=np.uint32)
        self.check(self.lib.ptw_graph_export_csr(self.h, _ptr(row_ptr), _ptr(col)))
        return row_ptr, col[: gi.num_edges]

    def export_names(self) -> np.ndarray:
        """names[v] = the id graph id v had in the input (file id; R-MAT generator id)."""
        names = np.empty(max(self.info().num_vertices, 1), dtype=np.uint32)
        self.check(self.lib.ptw_graph_export_names(self.h, _ptr(names)))
        return names[: self.info().num_vertices]

    def set_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self.check(self.lib.ptw_graph_set_weights(self.h, _ptr(w), len(w)))

    def random_weights(self, seed: int = 1):
        self.check(self.lib.ptw_graph_random_weights(self.h, seed))

    def export_weights(self):
        E = self.info().num_edges
        w = np.empty(max(E, 1), dtype=np.float32)
        thr = np.empty(max(E, 1), dtype=np.uint32)
        ali = np.empty(max(E, 1), dtype=np.uint32)
        self.check(self.lib.ptw_graph_export_weights(self.h, _ptr(w), _ptr(thr), _ptr(ali)))
        return w[:E], thr[:E], ali[:E]

    def write(self, path: str, fmt: int = FORMAT_TEXT):
        self.check(self.lib.ptw_graph_write(self.h, path.encode(), fmt))

    # ---- walks -----------------------------------------------------------
    def walk(self, desc: WalkDesc, counts=None, ends=None, paths=None, device_ptrs: bool = False,
             starts=None) -> WalkOut:
        """ptw_walk.  Host outputs must be C-contiguous numpy arrays: u64 counts of at
        least V entries, u32 ends and starts of at least `walkers`, u32 paths of at least
        walkers * walk_len (the library writes that many).  With device_ptrs the
        arguments are device addresses (ints) and are passed through unchecked."""
        out = WalkOut()
        if device_ptrs:
            out.counts, out.ends, out.paths, out.starts = counts, ends, paths, starts
        else:
            V = self.info().num_vertices
            need = {"counts": (counts, np.uint64, V),
                    "ends": (ends, np.uint32, desc.walkers),
                    "starts": (starts, np.uint32, desc.walkers),
                    "paths": (paths, np.uint32, desc.walkers * desc.walk_len)}
            for name, (a, dt, n) in need.items():
                if a is None:
                    continue
                if not isinstance(a, np.ndarray) or a.dtype != dt or not a.flags.c_contiguous or a.size < n:
                    raise ValueError(f"{name} must be a C-contiguous {np.dtype(dt).name} array of at least {n} "
                                     f"entries")
            out.counts, out.ends, out.paths, out.starts = _ptr(counts), _ptr(ends), _ptr(paths), _ptr(starts)
        self.check(self.lib.ptw_walk(self.h, ctypes.byref(desc), ctypes.byref(out)))
        return out

    def walk_static(self, desc: WalkDesc, counts=None, ends=None, paths=None, starts=None) -> WalkOut:
        """ptw_walk_static: the walk without any allocation.  desc.flags must hold
        FLAG_OUT_DEVICE and the outputs are device addresses (ints)."""
        out = WalkOut()
        out.counts, out.ends, out.paths, out.starts = counts, ends, paths, starts
        self.check(self.lib.ptw_walk_static(self.h, ctypes.byref(desc), ctypes.byref(out)))
        return out

    # ---- batched queries -------------------------------------------------
    @staticmethod
    def _batch_desc(sources, walkers, tps, streams, topk, seed, semantics, flags, dense_threshold, kinds=None):
        n = len(sources)
        qs = np.zeros(max(n, 1), dtype=QUERY_DTYPE)
        if n:
            if not (len(walkers) == len(tps) == n and (streams is None or len(streams) == n)
                    and (kinds is None or np.ndim(kinds) == 0 or len(kinds) == n)):
                raise ValueError("sources, walkers, tps, streams and kinds must have one entry per query")
            qs["source"][:n] = np.asarray(sources, dtype=np.int64)
            qs["walkers"][:n] = np.asarray(walkers, dtype=np.int64)
            qs["tp"][:n] = np.asarray(tps, dtype=np.float64)
            qs["stream"][:n] = np.arange(n) if streams is None else np.asarray(streams, dtype=np.int64)
            qs["kind"][:n] = QUERY_VISITS if kinds is None else np.asarray(kinds, dtype=np.int64)
        d = BatchDesc()
        d.queries = qs.ctypes.data
        d.num_queries = n
        d.topk = topk
        d.seed = seed
        d.semantics = semantics
        d.flags = flags
        d.dense_threshold = dense_threshold
        return d, qs

    def batch_ppr(self, sources, walkers, tps, *, streams=None, topk: int = 100, seed: int = 1,
                  semantics: int = SEM_REFERENCE, graph: bool = False, weighted: bool = False,
                  dense_threshold: int = 0, kinds=None) -> "BatchResult":
        """Many single-source queries in one launch (ptw_batch_ppr).  Query i uses Philox
        stream `streams[i]` (default i) and counts by `kinds[i]`: QUERY_VISITS (PPR, the
        default) or QUERY_ENDPOINTS (RWR end vertices); `kinds` may be one value for all."""
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_WEIGHTED if weighted else 0)
        d, qs = self._batch_desc(sources, walkers, tps, streams, topk, seed, semantics, flags, dense_threshold,
                                 kinds)
        n = len(sources)
        tv = np.zeros(max(n * topk, 1), dtype=np.uint32)
        tc = np.zeros(max(n * topk, 1), dtype=np.uint64)
        steps = np.zeros(max(n, 1), dtype=np.uint64)
        ent = np.zeros(max(n, 1), dtype=np.float64)
        out = BatchOut()
        out.topk_vertex, out.topk_count, out.steps, out.entropy = _ptr(tv), _ptr(tc), _ptr(steps), _ptr(ent)
        self.check(self.lib.ptw_batch_ppr(self.h, ctypes.byref(d), ctypes.byref(out)))
        del qs
        return BatchResult(tv[: n * topk].reshape(n, topk), tc[: n * topk].reshape(n, topk), steps[:n],
                           out.total_steps, out.slice_bytes, out.dense_queries, out.hash_queries, out.kernel_ms,
                           out.total_ms, ent[:n], out.logged_queries, bool(out.log_rerun))

    def batch_write_topk(self, path: str, sources, walkers, tps, result: "BatchResult", *, streams=None,
                         kinds=None, id_space: int = 1) -> None:
        """Write a batch's top-k lists as text (ptw_batch_write_topk): one line per entry,
        "query source kind rank vertex count", in the input's own ids (id_space 1,
        IDS_ORIGINAL) or graph ids (0, IDS_GRAPH)."""
        n, k = result.topk_vertex.shape
        d, qs = self._batch_desc(sources, walkers, tps, streams, k, 1, SEM_REFERENCE, 0, 0, kinds)
        tv = np.ascontiguousarray(result.topk_vertex, dtype=np.uint32)
        tc = np.ascontiguousarray(result.topk_count, dtype=np.uint64)
        out = BatchOut()
        out.topk_vertex, out.topk_count = _ptr(tv), _ptr(tc)
        self.check(self.lib.ptw_batch_write_topk(self.h, os.fsencode(path), ctypes.byref(d), ctypes.byref(out),
                                                 id_space))
        del qs

    def batch_gather(self, query_index, vertices) -> np.ndarray:
        """Counts of (query, graph id) pairs of the last batch."""
        q = np.ascontiguousarray(query_index, dtype=np.uint32)
        v = np.ascontiguousarray(vertices, dtype=np.uint32)
        out = np.zeros(max(len(q), 1), dtype=np.uint64)
        self.check(self.lib.ptw_batch_gather(self.h, len(q), _ptr(q), _ptr(v), _ptr(out)))
        return out[: len(q)]

    def batch_plan_bytes(self, sources, walkers, tps, *, topk: int = 100, dense_threshold: int = 0) -> int:
        d, qs = self._batch_desc(sources, walkers, tps, None, topk, 1, SEM_REFERENCE, 0, dense_threshold)
        b = ctypes.c_uint64()
        self.check(self.lib.ptw_batch_plan_bytes(self.h, ctypes.byref(d), ctypes.byref(b)))
        del qs
        return b.value

    def batch_counts(self, index: int) -> np.ndarray:
        out = np.zeros(self.info().num_vertices, dtype=np.uint64)
        self.check(self.lib.ptw_batch_counts(self.h, index, _ptr(out)))
        return out

    # ---- FP64 solver ------------------------------------------------------
    def ppr_exact(self, sources, tps, *, semantics: int = SEM_REFERENCE, tol: float = 1e-10,
                  full: bool = True, k: int = 0):
        """Expected visits per walker from each source (FP64, on the device).
        Returns (vectors [n, V] or None, topk_vertex [n, k] or None, topk_value, iterations)."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        tp = np.ascontiguousarray(tps, dtype=np.float64)
        n = len(src)
        V = self.info().num_vertices
        out = np.zeros((n, V), dtype=np.float64) if full else None
        tv = np.zeros((n, k), dtype=np.uint32) if k else None
        tval = np.zeros((n, k), dtype=np.float64) if k else None
        it = ctypes.c_uint32()
        self.check(self.lib.ptw_ppr_exact(self.h, _ptr(src), _ptr(tp), n, semantics, tol, _ptr(out), k, _ptr(tv),
                                          _ptr(tval), ctypes.byref(it)))
        return out, tv, tval, it.value

    # ---- serving -----------------------------------------------------------
    def serve(self, sources, eps, arrival_s, *, ids=None, mlp: Optional["BudgetMLP"] = None, tp: float = 0.2,
              seed: int = 1, budget_mode: int = BUDGET_FIXED, topk: int = 100, fixed_c2: float = 2.0,
              margin: float = 1.0, w_min: int = 1000, w_max: int = 10 ** 7, feat_lo=None, feat_hi=None,
              audit_rate: float = 0.0, max_batch: int = 512, hbm_limit: int = 0) -> "ServeResult":
        """Serve requests arriving at `arrival_s` (seconds after the call) by continuous
        batching (ptw_serve_replay).  Request i has id `ids[i]` (default i)."""
        n = len(sources)
        reqs = (Request * max(n, 1))()
        ids = range(n) if ids is None else ids
        for i, (s_, e_, a_, id_) in enumerate(zip(sources, eps, arrival_s, ids)):
            reqs[i].source, reqs[i].eps, reqs[i].arrival_s, reqs[i].id = int(s_), float(e_), float(a_), int(id_)
        c = ServeConfig()
        c.tp, c.seed, c.budget_mode, c.topk = tp, seed, budget_mode, topk
        c.fixed_c2, c.margin, c.w_min, c.w_max = fixed_c2, margin, w_min, w_max
        lo = [-3.0e38] * 6 if feat_lo is None else feat_lo
        hi = [3.0e38] * 6 if feat_hi is None else feat_hi
        for k in range(6):
            c.feat_lo[k], c.feat_hi[k] = float(lo[k]), float(hi[k])
        c.audit_rate, c.max_batch, c.hbm_limit = audit_rate, max_batch, hbm_limit
        res = (RequestResult * max(n, 1))()
        tv = np.zeros(max(n * topk, 1), dtype=np.uint32)
        tc = np.zeros(max(n * topk, 1), dtype=np.uint64)
        st = ServeStats()
        self.check(self.lib.ptw_serve_replay(self.h, mlp.h if mlp is not None else None, ctypes.byref(c), reqs, n,
                                             res, _ptr(tv), _ptr(tc), ctypes.byref(st)))
        field = lambda f, dt: np.array([getattr(res[i], f) for i in range(n)], dtype=dt)
        return ServeResult(tv[: n * topk].reshape(n, topk), tc[: n * topk].reshape(n, topk),
                           field("budget", np.uint64), field("admit_s", np.float64), field("done_s", np.float64),
                           field("fallback", bool), field("audited", bool), field("topped_up", bool),
                           field("launches", np.uint32), st.as_dict())

    # ---- launch tuning ---------------------------------------------------
    def tuning(self) -> Tuning:
        t = Tuning()
        self.check(self.lib.ptw_tuning_get(self.h, ctypes.byref(t)))
        return t

    def set_tuning(self, **kw) -> Tuning:
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        self.check(self.lib.ptw_tuning_set(self.h, ctypes.byref(t)))
        return t

    # ---- device buffers --------------------------------------------------
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def device_memory(self):
        f, t = ctypes.c_uint64(), ctypes.c_uint64()
        self.check(self.lib.ptw_device_memory(self.h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def bench_copy(self, nbytes: int):
        gbs = ctypes.c_double()
        ms = ctypes.c_double()
        self.check(self.lib.ptw_bench_copy(self.h, nbytes, ctypes.byref(gbs), ctypes.byref(ms)))
        return gbs.value, ms.value

    def bench_gather(self, table_bytes: int, gathers: int, seed: int = 1):
        gps = ctypes.c_double()
        ms = ctypes.c_double()
        self.check(self.lib.ptw_bench_gather(self.h, table_bytes, gathers, seed, ctypes.byref(gps), ctypes.byref(ms)))
        return gps.value, ms.value

    def bench_step(self, steps: int, seed: int = 1):
        """Independent walk steps from degree-drawn vertices (ptw_bench_step): (steps/s, ms)."""
        sps = ctypes.c_double()
        ms = ctypes.c_double()
        self.check(self.lib.ptw_bench_step(self.h, steps, seed, ctypes.byref(sps), ctypes.byref(ms)))
        return sps.value, ms.value

    def check_primitives(self, n: int, bits: int = 32, seed: int = 1):
        """ptw_check_primitives: the ingest scans, sort, select and max on n random
        items against a host computation; raises PtwError on the first difference."""
        self.check(self.lib.ptw_check_primitives(self.h, n, bits, seed))


class BudgetMLP:
    """The 6-64-64-1 budget predictor (ptw_mlp_*): BF16 weights, FP32 accumulation,
    trained on the device."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        h = _P()
        ctx.check(ctx.lib.ptw_mlp_create(ctx.h, seed, ctypes.byref(h)))
        self.h = h
        self.num_params = ctx.lib.ptw_mlp_num_params()

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.ptw_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_normalization(self, mean, std, y_mean: float, y_std: float):
        m = np.ascontiguousarray(mean, dtype=np.float32)
        s = np.ascontiguousarray(std, dtype=np.float32)
        self.ctx.check(self.ctx.lib.ptw_mlp_set_normalization(self.h, _ptr(m), _ptr(s), y_mean, y_std))

    def params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        n = np.zeros(14, dtype=np.float32)
        self.ctx.check(self.ctx.lib.ptw_mlp_get_params(self.h, _ptr(p), _ptr(n)))
        return p, n

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert len(p) == self.num_params
        self.ctx.check(self.ctx.lib.ptw_mlp_set_params(self.h, _ptr(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        y = np.zeros(max(len(x), 1), dtype=np.float32)
        self.ctx.check(self.ctx.lib.ptw_mlp_forward(self.h, _ptr(x), len(x), _ptr(y)))
        return y[: len(x)]

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = ctypes.c_float()
        self.ctx.check(self.ctx.lib.ptw_mlp_gradients(self.h, _ptr(x), _ptr(t), len(x), _ptr(g), ctypes.byref(loss)))
        return g, loss.value

    def train(self, x, t, *, epochs: int = 50, batch: int = 256, lr: float = 1e-3, seed: int = 1) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_float()
        self.ctx.check(self.ctx.lib.ptw_mlp_train(self.h, _ptr(x), _ptr(t), len(x), epochs, batch, lr, seed,
                                                  ctypes.byref(loss)))
        return loss.value


class DeviceBuffer:
    """Device memory on a context's device (ptw_buffer_*)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _P()
        ctx.check(ctx.lib.ptw_buffer_alloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def zero(self):
        self.ctx.check(self.ctx.lib.ptw_buffer_zero(self.ctx.h, self.ptr, self.nbytes))
        return self

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        assert host.nbytes <= self.nbytes
        self.ctx.check(self.ctx.lib.ptw_buffer_upload(self.ctx.h, self.ptr, host.ctypes.data, host.nbytes))
        return self

    def download(self, dtype, count: Optional[int] = None, offset: int = 0) -> np.ndarray:
        """Copy `count` elements starting at element `offset` to the host."""
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize - offset
        if offset < 0 or count < 0 or (offset + count) * dtype.itemsize > self.nbytes:
            raise ValueError("range outside the buffer")
        out = np.empty(count, dtype=dtype)
        self.ctx.check(self.ctx.lib.ptw_buffer_download(self.ctx.h, self.ptr + offset * dtype.itemsize,
                                                        out.ctypes.data, out.nbytes))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            self.ctx.lib.ptw_buffer_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---- the reference's apps as plain data --------------------------------------
@dataclass
class PageRank:
    """src/core/pagerank.cpp:10-90: uniform starts, 1 + Geometric(tp) steps, visit counts."""
    tp: float
    semantics: int = SEM_REFERENCE


@dataclass
class PPR:
    """src/core/ppr.cpp:10-119: every walker starts at `source`, visit counts."""
    tp: float
    source: int
    semantics: int = SEM_REFERENCE


@dataclass
class RWR:
    """src/core/RWR.cpp:15-107: end vertex per walker (source, or uniform if None)."""
    tp: float
    source: Optional[int] = None
    semantics: int = SEM_REFERENCE


@dataclass
class DeepWalk:
    """src/core/deepwalk.cpp:10-42: fixed walk_len vertices per path (walk_len - 1 steps)."""
    walk_len: int
    start: int = START_UNIFORM
    source: Optional[int] = None


@dataclass
class Node2Vec:
    """Second-order walks of walk_len vertices (src/core/node2vec.cpp, flags -p -q): from cur,
    reached from prev, neighbour x weighs 1/p if x == prev, 1 if x neighbours prev, 1/q otherwise."""
    walk_len: int
    p: float
    q: float
    start: int = START_UNIFORM
    source: Optional[int] = None


def make_desc(app, walkers: int, seed: int = 1, wid_base: int = 0, stream: int = 0,
              dangling: int = DANGLING_ERROR, flags: int = 0) -> WalkDesc:
    d = WalkDesc()
    d.walkers = walkers
    d.seed = seed
    d.wid_base = wid_base
    d.stream = stream
    d.dangling_policy = dangling
    d.flags = flags
    if isinstance(app, PageRank):
        d.app, d.tp, d.semantics, d.start_mode = APP_PAGERANK, app.tp, app.semantics, START_UNIFORM
    elif isinstance(app, PPR):
        d.app, d.tp, d.semantics, d.start_mode, d.source = APP_PPR, app.tp, app.semantics, START_SOURCE, app.source
    elif isinstance(app, RWR):
        d.app, d.tp, d.semantics = APP_RWR, app.tp, app.semantics
        if app.source is None:
            d.start_mode = START_UNIFORM
        else:
            d.start_mode, d.source = START_SOURCE, app.source
    elif isinstance(app, DeepWalk):
        d.app, d.walk_len, d.start_mode = APP_DEEPWALK, app.walk_len, app.start
        if app.source is not None:
            d.start_mode, d.source = START_SOURCE, app.source
    elif isinstance(app, Node2Vec):
        d.app, d.walk_len, d.start_mode, d.p, d.q = APP_NODE2VEC, app.walk_len, app.start, app.p, app.q
        if app.source is not None:
            d.start_mode, d.source = START_SOURCE, app.source
    else:
        raise TypeError(f"unknown app {app!r}")
    return d


@dataclass
class BatchResult:
    topk_vertex: np.ndarray  # (queries, k) graph ids
    topk_count: np.ndarray   # (queries, k)
    steps: np.ndarray        # (queries,)
    total_steps: int
    slice_bytes: int
    dense_queries: int
    hash_queries: int
    kernel_ms: float
    total_ms: float
    entropy: np.ndarray      # (queries,) nats
    logged_queries: int = 0  # dense queries counted through the visit log
    log_rerun: bool = False  # the log filled and the batch ran again with slices


@dataclass
class ServeResult:
    topk_vertex: np.ndarray  # (requests, k) graph ids
    topk_count: np.ndarray   # (requests, k) visit counts of the answering walk
    budget: np.ndarray       # walkers of each main walk
    admit_s: np.ndarray
    done_s: np.ndarray
    fallback: np.ndarray
    audited: np.ndarray
    topped_up: np.ndarray
    launches: np.ndarray
    stats: dict


@dataclass
class WalkResult:
    counts: Optional[np.ndarray]
    ends: Optional[np.ndarray]
    paths: Optional[np.ndarray]
    steps: int
    visits: int
    dangling: int
    kernel_ms: float
    total_ms: float
    shape: Optional[dict] = None  # the launch the walk ran with (WalkOut.launch_shape)
    starts: Optional[np.ndarray] = None  # RWR with starts=True: start vertex per walker


def walk(ctx: Context, walker_num: int, app, seed: int = 1, wid_base: int = 0, stream: int = 0,
         dangling: int = DANGLING_ERROR, starts: bool = False) -> WalkResult:
    """Run `walker_num` walkers of `app` on the context's graph (host outputs).  With
    `starts`, an RWR walk also returns each walker's start vertex."""
    gi = ctx.info()
    d = make_desc(app, walker_num, seed=seed, wid_base=wid_base, stream=stream, dangling=dangling)
    counts = ends = paths = st = None
    if d.app in (APP_PAGERANK, APP_PPR):
        counts = np.zeros(gi.num_vertices, dtype=np.uint64)
    elif d.app == APP_RWR:
        ends = np.zeros(max(walker_num, 1), dtype=np.uint32)
        if starts:
            st = np.zeros(max(walker_num, 1), dtype=np.uint32)
    else:
        paths = np.zeros(max(walker_num * d.walk_len, 1), dtype=np.uint32)
    out = ctx.walk(d, counts=counts, ends=ends, paths=paths, starts=st)
    if st is not None:
        st = st[:walker_num]
    if ends is not None:
        ends = ends[:walker_num]
    if paths is not None:
        paths = paths[: walker_num * d.walk_len].reshape(walker_num, d.walk_len)
    return WalkResult(counts, ends, paths, out.steps, out.visits, out.dangling, out.kernel_ms, out.total_ms,
                      out.launch_shape(), st)
