"""CPU oracle for the walk engine -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker (or, for the baseline, as the timed
CPU path).  The product (ptwalker_amd, libptw.so, the CLI binaries) never
loads it.

Contents (reference files followed are cited per function):
* ctypes bindings of liboracle.so (oracle.c): Philox4x32-10, the Philox
  walk replay, the reference process with its own xorshift* RNG, the FP64
  power iteration;
* pure numpy restatements used to pin those: the reference loader's id
  compaction and CSR order, a dense FP64 linear solve for the expected
  visit and endpoint vectors.

Parity status of each oracle function and what pins it is listed in the
module-level PINS table.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

APP_PAGERANK, APP_PPR, APP_RWR, APP_DEEPWALK, APP_NODE2VEC = 0, 1, 2, 3, 4
SEM_REFERENCE, SEM_TEXTBOOK = 0, 1
START_UNIFORM, START_SOURCE, START_ROUNDROBIN = 0, 1, 2
INVALID = 0xFFFFFFFF

PINS = {
    "philox4x32_10": "Random123 known-answer vectors + rocRAND's philox4x32_10 engine (tests/golden/philox_kat.json)",
    "walk_replay": "closed forms (sum of lengths, K2 alternation, bipartite parity), FP64 oracle within 4 sigma, "
                   "transition-matrix test",
    "ref_walk": "exact pmf of the quantized geometric draw (chi-square), FP64 oracle within 4 sigma",
    "power_iteration": "dense FP64 linear solve on small graphs, closed forms on K2 and cycles",
    "load_edge_list": "hand-written edge files with known compaction; edge multiset round trip",
}


class WalkDesc(ctypes.Structure):
    _fields_ = [
        ("app", ctypes.c_int),
        ("semantics", ctypes.c_int),
        ("start_mode", ctypes.c_int),
        ("source", ctypes.c_uint32),
        ("tp", ctypes.c_double),
        ("walk_len", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("stream", ctypes.c_uint32),
        ("wid_base", ctypes.c_uint64),
        ("walkers", ctypes.c_uint64),
    ]


class WalkStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("visits", ctypes.c_uint64), ("dangling", ctypes.c_uint64),
                ("proposals", ctypes.c_uint64), ("exact_draws", ctypes.c_uint64)]


_lib = None
_P = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_philox4x32_10.argtypes = [_P, _P, _P]
        L.orc_philox4x32_10.restype = None
        L.orc_mulhi64.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_mulhi64.restype = ctypes.c_uint32
        L.orc_continue_threshold.argtypes = [ctypes.c_double]
        L.orc_continue_threshold.restype = ctypes.c_uint32
        L.orc_walk_replay.argtypes = [_P, _P, ctypes.c_uint32, ctypes.POINTER(WalkDesc), _P, _P, _P,
                                      ctypes.POINTER(WalkStats), ctypes.c_int]
        L.orc_walk_replay.restype = ctypes.c_int
        L.orc_walk_replay_weighted.argtypes = [_P, _P, ctypes.c_uint32, _P, _P, ctypes.POINTER(WalkDesc), _P, _P,
                                               _P, ctypes.POINTER(WalkStats), ctypes.c_int]
        L.orc_walk_replay_weighted.restype = ctypes.c_int
        L.orc_node2vec_replay.argtypes = [_P, _P, ctypes.c_uint32, ctypes.POINTER(WalkDesc), ctypes.c_double,
                                          ctypes.c_double, _P, ctypes.POINTER(WalkStats), ctypes.c_int]
        L.orc_node2vec_replay.restype = ctypes.c_int
        L.orc_build_alias.argtypes = [_P, ctypes.c_uint32, _P, _P, _P]
        L.orc_build_alias.restype = ctypes.c_int
        L.orc_walk_length.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_int]
        L.orc_walk_length.restype = ctypes.c_uint32
        L.orc_ref_walk.argtypes = [_P, _P, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                   _P, _P, ctypes.POINTER(WalkStats)]
        L.orc_ref_walk.restype = ctypes.c_int
        L.orc_ref_step_histogram.argtypes = [ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64, _P, ctypes.c_uint32]
        L.orc_ref_step_histogram.restype = None
        L.orc_xorshift_gen.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, _P]
        L.orc_xorshift_gen.restype = None
        L.orc_ref_partitioned_walk.argtypes = [_P, _P, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                               ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                               ctypes.c_uint64, ctypes.c_uint32, _P, _P, ctypes.POINTER(WalkStats),
                                               ctypes.POINTER(ctypes.c_double), _P]
        L.orc_ref_partitioned_walk.restype = ctypes.c_int
        L.orc_xorshift_gen_float.argtypes = [ctypes.c_uint64, ctypes.c_float, ctypes.c_uint64, _P]
        L.orc_xorshift_gen_float.restype = None
        L.orc_power_iteration.argtypes = [_P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_int, _P]
        L.orc_power_iteration.restype = ctypes.c_int
        _lib = L
    return _lib


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def philox(ctr, key) -> np.ndarray:
    """Philox4x32-10 of one counter block."""
    c = _u32(ctr)
    k = _u32(key)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def mulhi64(lo: int, hi: int, n: int) -> int:
    return int(lib().orc_mulhi64(lo, hi, n))


def continue_threshold(tp: float) -> int:
    return int(lib().orc_continue_threshold(tp))


def walk_length(seed: int, stream: int, wid: int, tp: float, semantics: int = SEM_REFERENCE) -> int:
    return int(lib().orc_walk_length(seed, stream, wid, tp, semantics))


def node2vec_replay(row_ptr, col, walkers: int, walk_len: int, p: float, q: float, *, seed: int = 1,
                    source: int = 0, start_mode: int = START_UNIFORM, stream: int = 0, wid_base: int = 0,
                    threads: int = 0):
    """Philox replay of the engine's node2vec walks (oracle.c orc_node2vec_replay).
    Returns dict(paths [walkers, walk_len], steps, dangling, proposals,
    exact_draws): the rejection sampler's proposals and the steps its exact
    draw decided."""
    rp, cl = _u32(row_ptr), _u32(col)
    V = len(rp) - 1
    d = WalkDesc(APP_NODE2VEC, SEM_REFERENCE, start_mode, source, 0.5, walk_len, seed, stream, wid_base, walkers)
    paths = np.zeros(max(walkers * walk_len, 1), dtype=np.uint32)
    st = WalkStats()
    rc = lib().orc_node2vec_replay(rp.ctypes.data, cl.ctypes.data if len(cl) else None, V, ctypes.byref(d),
                                   ctypes.c_double(p), ctypes.c_double(q), paths.ctypes.data, ctypes.byref(st),
                                   threads)
    if rc != 0:
        raise ValueError("invalid oracle node2vec arguments")
    return dict(paths=paths[: walkers * walk_len].reshape(walkers, walk_len), steps=st.steps, dangling=st.dangling,
                proposals=st.proposals, exact_draws=st.exact_draws)


def node2vec_law(row_ptr, col, prev: int, cur: int, p: float, q: float):
    """The second-order transition law from the definition (KnightKing,
    third_party/knightking/test_node2vec.hpp:47-79): over cur's adjacency
    positions, weight 1/p for prev, 1 for a neighbour of prev, 1/q otherwise,
    normalized.  Returns {vertex: probability} (exact real weights)."""
    rp = np.asarray(row_ptr, dtype=np.int64)
    nb_prev = set(int(x) for x in col[rp[prev]:rp[prev + 1]])
    law = {}
    for x in col[rp[cur]:rp[cur + 1]]:
        x = int(x)
        w = 1.0 / p if x == prev else (1.0 if x in nb_prev else 1.0 / q)
        law[x] = law.get(x, 0.0) + w
    tot = sum(law.values())
    return {k: v / tot for k, v in law.items()}


def walk_replay(row_ptr, col, app: int, walkers: int, *, tp: float = 0.15, seed: int = 1, source: int = 0,
                start_mode: int = START_UNIFORM, semantics: int = SEM_REFERENCE, walk_len: int = 0,
                stream: int = 0, wid_base: int = 0, threads: int = 0, alias=None):
    """Philox replay of the engine's process (oracle.c orc_walk_replay).
    alias: optional (thr, idx) tables for weighted neighbour choice.

    Returns dict(counts, ends, paths, steps, visits, dangling)."""
    rp, cl = _u32(row_ptr), _u32(col)
    V = len(rp) - 1
    d = WalkDesc(app, semantics, start_mode, source, tp, walk_len, seed, stream, wid_base, walkers)
    counts = np.zeros(V, dtype=np.uint64) if app in (APP_PAGERANK, APP_PPR) else None
    ends = np.zeros(max(walkers, 1), dtype=np.uint32) if app == APP_RWR else None
    paths = np.zeros(max(walkers * walk_len, 1), dtype=np.uint32) if app == APP_DEEPWALK else None
    st = WalkStats()
    at = ai = None
    if alias is not None:
        at, ai = _u32(alias[0]), _u32(alias[1])
    rc = lib().orc_walk_replay_weighted(rp.ctypes.data, cl.ctypes.data if len(cl) else None, V,
                                        at.ctypes.data if at is not None else None,
                                        ai.ctypes.data if ai is not None else None, ctypes.byref(d),
                                        counts.ctypes.data if counts is not None else None,
                                        ends.ctypes.data if ends is not None else None,
                                        paths.ctypes.data if paths is not None else None, ctypes.byref(st), threads)
    if rc != 0:
        raise ValueError("invalid oracle walk arguments")
    if ends is not None:
        ends = ends[:walkers]
    if paths is not None:
        paths = paths[: walkers * walk_len].reshape(walkers, walk_len)
    return dict(counts=counts, ends=ends, paths=paths, steps=st.steps, visits=st.visits, dangling=st.dangling)


def build_alias(row_ptr, w):
    """Integer alias tables from float weights (oracle.c orc_build_alias)."""
    rp = _u32(row_ptr)
    w = np.ascontiguousarray(w, dtype=np.float32)
    E = int(rp[-1])
    thr = np.zeros(max(E, 1), dtype=np.uint32)
    ali = np.zeros(max(E, 1), dtype=np.uint32)
    if lib().orc_build_alias(rp.ctypes.data, len(rp) - 1, w.ctypes.data, thr.ctypes.data, ali.ctypes.data) != 0:
        raise ValueError("weights must be finite and > 0")
    return thr[:E], ali[:E]


def alias_distribution(row_ptr, thr, ali, v):
    """Exact (rational) neighbour distribution implied by vertex v's alias table."""
    from fractions import Fraction
    b, e = int(row_ptr[v]), int(row_ptr[v + 1])
    n = e - b
    p = [Fraction(0)] * n
    for j in range(n):
        own = Fraction(int(thr[b + j]), 2 ** 32)
        p[j] += own / n
        p[int(ali[b + j])] += (1 - own) / n
    return p


def ref_walk(row_ptr, col, app: int, walkers: int, *, tp: float = 0.15, source: int = 0, start_uniform: bool = True,
             walk_len: int = 0, seed: int = 1, threads: int = 0):
    """The reference process with its own RNG (xorshift*, quantized geometric)."""
    rp, cl = _u32(row_ptr), _u32(col)
    V = len(rp) - 1
    counts = np.zeros(V, dtype=np.uint64)
    ends = np.zeros(max(walkers, 1), dtype=np.uint32) if app == APP_RWR else None
    st = WalkStats()
    rc = lib().orc_ref_walk(rp.ctypes.data, cl.ctypes.data, V, app, source, int(start_uniform), tp, walk_len,
                            walkers, seed, threads, counts.ctypes.data,
                            ends.ctypes.data if ends is not None else None, ctypes.byref(st))
    if rc != 0:
        raise ValueError("invalid reference-walk arguments")
    return dict(counts=counts, ends=None if ends is None else ends[:walkers], steps=st.steps, visits=st.visits,
                dangling=st.dangling)


def l2_cache_bytes() -> int:
    """The host's per-core L2 size, which the reference reads to size its partitions."""
    try:
        n = os.sysconf("SC_LEVEL2_CACHE_SIZE")
        if n and n > 0:
            return int(n)
    except (ValueError, OSError):
        pass
    return 1 << 20


def ref_partitioned_walk(row_ptr, col, app: int, walkers: int, *, tp: float = 0.15, source: int = 0,
                         start_uniform: bool = True, seed: int = 1, threads: int = 0, l2_bytes: int = 0,
                         degree_bound: int = 32):
    """The reference's own schedule (partitions, walker pools, migration, buffered hub
    sampling) with its own RNG; returns counts (PAGERANK/PPR) or ends (RWR), the step
    total, the walk time (walker init + walk) and the partition counts."""
    rp, cl = _u32(row_ptr), _u32(col)
    V = len(rp) - 1
    counts = np.zeros(V, dtype=np.uint64) if app != APP_RWR else None
    ends = np.zeros(max(walkers, 1), dtype=np.uint32) if app == APP_RWR else None
    st = WalkStats()
    secs = ctypes.c_double()
    parts = np.zeros(2, dtype=np.uint32)
    rc = lib().orc_ref_partitioned_walk(rp.ctypes.data, cl.ctypes.data, V, app, source, int(start_uniform), tp,
                                        walkers, seed, threads, l2_bytes or l2_cache_bytes(), degree_bound,
                                        counts.ctypes.data if counts is not None else None,
                                        ends.ctypes.data if ends is not None else None, ctypes.byref(st),
                                        ctypes.byref(secs), parts.ctypes.data)
    if rc != 0:
        raise ValueError("invalid partitioned-walk arguments")
    return dict(counts=counts, ends=None if ends is None else ends[:walkers], steps=st.steps, visits=st.visits,
                dangling=st.dangling, seconds=secs.value, hub_partitions=int(parts[0]),
                low_partitions=int(parts[1]))


def ref_step_histogram(tp: float, n: int, seed: int = 1, hist_len: int = 512) -> np.ndarray:
    h = np.zeros(hist_len, dtype=np.uint64)
    lib().orc_ref_step_histogram(tp, seed, n, h.ctypes.data, hist_len)
    return h


def xorshift_gen(state: int, bound: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_xorshift_gen(state, bound, n, out.ctypes.data)
    return out


def xorshift_gen_float(state: int, upper: float, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.float32)
    lib().orc_xorshift_gen_float(state, upper, n, out.ctypes.data)
    return out


def power_iteration(row_ptr, col, *, tp: float, start: int | None, semantics: int = SEM_REFERENCE,
                    mode: str = "visits", tol: float = 1e-15, max_iter: int = 100000) -> np.ndarray:
    """FP64 expected visits per walker (mode='visits') or endpoint distribution (mode='ends')."""
    rp, cl = _u32(row_ptr), _u32(col)
    V = len(rp) - 1
    out = np.zeros(V, dtype=np.float64)
    it = lib().orc_power_iteration(rp.ctypes.data, cl.ctypes.data, V, INVALID if start is None else start, tp,
                                   semantics, 0 if mode == "visits" else 1, tol, max_iter, out.ctypes.data)
    if it < 0:
        raise MemoryError("power iteration allocation failed")
    return out


# ---- numpy restatements (pins) -------------------------------------------------
def transition_matrix(row_ptr, col) -> np.ndarray:
    """Dense row-stochastic uniform-neighbour matrix (multi-edges weigh by multiplicity)."""
    V = len(row_ptr) - 1
    P = np.zeros((V, V))
    for u in range(V):
        b, e = int(row_ptr[u]), int(row_ptr[u + 1])
        for k in range(b, e):
            P[u, col[k]] += 1.0 / (e - b)
    return P


def dense_expected(row_ptr, col, *, tp: float, start: int | None, semantics: int = SEM_REFERENCE,
                   mode: str = "visits") -> np.ndarray:
    """Closed form by a dense FP64 solve, for small graphs.

    visits (reference): e_s + e_s P (I - (1-tp) P)^-1
    visits (textbook):  e_s + (1-tp) e_s P (I - (1-tp) P)^-1
    ends (reference):   tp e_s P (I - (1-tp) P)^-1
    ends (textbook):    tp e_s (I - (1-tp) P)^-1
    """
    V = len(row_ptr) - 1
    P = transition_matrix(row_ptr, col)
    e = np.full(V, 1.0 / V) if start is None else np.eye(V)[start]
    M = np.linalg.inv(np.eye(V) - (1 - tp) * P)
    if mode == "visits":
        w = 1.0 if semantics == SEM_REFERENCE else (1 - tp)
        return e + w * (e @ P @ M)
    if semantics == SEM_REFERENCE:
        return tp * (e @ P @ M)
    return tp * (e @ M)


def load_edge_list(path: str, fmt: str = "text", as_undirected: bool = False):
    """The reference loader's semantics (include/io.hpp:39-53,76-90; src/core/graph.hpp:293-366):
    ids by first appearance (src then dst), adjacency in edge order, reversed copies after all
    forward edges.  Returns (row_ptr, col, name2id dict)."""
    src, dst = [], []
    if fmt == "text":
        with open(path, "rb") as f:
            for raw in f:
                line = raw.decode()
                if line.startswith("#"):
                    continue
                parts = line.split()
                if len(parts) < 2:
                    continue
                src.append(int(parts[0]))
                dst.append(int(parts[1]))
    else:
        a = np.fromfile(path, dtype=np.uint32)
        a = a[: (len(a) // 2) * 2].reshape(-1, 2)
        src, dst = a[:, 0].tolist(), a[:, 1].tolist()
    name2id = {}
    cs, cd = [], []
    for s, d in zip(src, dst):
        if s not in name2id:
            name2id[s] = len(name2id)
        if d not in name2id:
            name2id[d] = len(name2id)
        cs.append(name2id[s])
        cd.append(name2id[d])
    if as_undirected:
        cs, cd = cs + cd, cd + cs
    V = len(name2id)
    adj = [[] for _ in range(V)]
    for s, d in zip(cs, cd):
        adj[s].append(d)
    row_ptr = np.zeros(V + 1, dtype=np.uint32)
    row_ptr[1:] = np.cumsum([len(x) for x in adj])
    col = np.array([x for lst in adj for x in lst], dtype=np.uint32)
    return row_ptr, col, name2id


def topk(counts, row_ptr, k):
    """Top-k of a count vector: count descending, ties by degree descending then
    vertex id ascending; zero counts never listed.  Returns (vertices, counts)
    padded with (INVALID, 0) to length k."""
    counts = np.asarray(counts, dtype=np.uint64)
    rp = np.asarray(row_ptr, dtype=np.int64)
    ids = np.flatnonzero(counts)  # zero counts are never listed
    deg = rp[ids + 1] - rp[ids]
    order = ids[np.lexsort((ids, -deg, -counts[ids].astype(np.float64)))][:k]
    v = np.full(k, INVALID, dtype=np.uint32)
    c = np.zeros(k, dtype=np.uint64)
    v[: len(order)] = order
    c[: len(order)] = counts[order]
    return v, c


def topk_l1(estimate, exact, k=100):
    """Top-k L1 error: sum over the k largest entries of the normalized exact vector of
    |estimate/sum(estimate) - exact/sum(exact)|."""
    est = np.asarray(estimate, dtype=np.float64)
    ex = np.asarray(exact, dtype=np.float64)
    est = est / est.sum()
    ex = ex / ex.sum()
    top = np.argsort(-ex, kind="stable")[:k]
    return float(np.abs(est[top] - ex[top]).sum())
