"""Autotuner for the walk engine's launch shape.

The knobs (ptw_tuning and the batch planner's dense threshold) change how
work is laid out on the device, never what is computed: every walker keeps
its Philox stream, so counts are identical under every setting.  The tuner
therefore times candidates on a sample of the caller's own workload and
checks, for each one, that the results equal those of the starting point.

Search: coordinate descent -- one knob at a time over its candidate values,
keeping the best, for up to `rounds` passes or until a pass changes nothing.
Each candidate is timed by the median kernel time of `repeats` runs after
one warm-up run.

Knobs:
  walks   chunk (walker ids per wavefront refill), block, lds_hubs (direct LDS
          counters), cache_slots (LDS visit cache)
  batches batch_chunk (walker ids per queue entry of a hash query),
          dense_chunk (walker ids per workgroup chunk of a dense query; 0 =
          the planner's per-query rule), block, lds_hubs, cache_slots, dense_threshold
          (walkers above which a query is counted densely -- through the visit
          log, or in a dense slice when visit_log is 0 -- rather than in a hash
          slice), topk_tile (LDS candidate slots of the top-k kernel)

Results are cached in a JSON file keyed by the graph's shape and a workload
signature; `apply` restores a cached choice without timing anything.
"""
from __future__ import annotations

import json
import os
import statistics
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "ptwalker_amd", "tune.json")

WALK_SPACE: Dict[str, List[int]] = {
    "chunk": [1024, 2048, 4096, 8192, 16384],
    "block": [256, 512, 1024],
    "lds_hubs": [16384, 36864, 55296, 77824],
    "cache_slots": [-1, 0, 512, 1024, 2048],  # -1: the library's per-walk choice
    "visit_log": [1, 0],
    "edge_records": [1, 0],
}

BATCH_SPACE: Dict[str, List[int]] = {
    "batch_chunk": [2048, 4096, 8192, 16384, 32768, 65536],
    "dense_chunk": [0, 16384, 32768, 49152],
    "block": [256, 512, 1024],
    # the cache before the hub counts: 79872 hubs fit the LDS only without a cache
    "cache_slots": [0, 512],
    "lds_hubs": [36864, 55296, 77824, 79872],
    # walkers above which a query is dense (logged or in a slice, by visit_log)
    "dense_threshold": [1, 1_500, 4_096, 12_500, 25_000, 50_000, 100_000, 200_000],
    "topk_tile": [256, 512, 1024, 2048, 4096],
}


@dataclass
class TuneResult:
    best: Dict[str, int]
    best_ms: float
    start_ms: float
    trials: List[dict] = field(default_factory=list)
    # walks: the launch the best setting ran with, after the library's automatic
    # rules (blocks_per_cu = 0, cache_slots = -1)
    effective: Optional[Dict[str, int]] = None

    @property
    def speedup(self) -> float:
        return self.start_ms / self.best_ms if self.best_ms > 0 else 1.0


def coordinate_descent(measure: Callable[[Dict[str, int]], float], start: Dict[str, int],
                       space: Dict[str, List[int]], *, rounds: int = 2,
                       valid: Optional[Callable[[Dict[str, int]], bool]] = None) -> TuneResult:
    """Minimise measure(config) one knob at a time.  `measure` returns a time
    (ms) or raises ValueError for a candidate the engine rejects."""
    valid = valid or (lambda c: True)
    cur = dict(start)
    cache: Dict[tuple, float] = {}
    trials: List[dict] = []

    def cost(c):
        key = tuple(sorted(c.items()))
        if key not in cache:
            try:
                cache[key] = measure(dict(c)) if valid(c) else float("inf")
            except ValueError:
                cache[key] = float("inf")
            trials.append({**c, "ms": cache[key]})
        return cache[key]

    start_ms = cost(cur)
    best_ms = start_ms
    for _ in range(rounds):
        changed = False
        for knob, values in space.items():
            for val in values:
                if val == cur.get(knob):
                    continue
                cand = dict(cur)
                cand[knob] = val
                t = cost(cand)
                if t < best_ms:
                    best_ms, cur, changed = t, cand, True
        if not changed:
            break
    return TuneResult(best=cur, best_ms=best_ms, start_ms=start_ms, trials=trials)


def _median_ms(run: Callable[[], float], repeats: int) -> float:
    run()  # warm-up: pools, code objects, caches
    return statistics.median(run() for _ in range(repeats))


def graph_signature(ctx) -> str:
    gi = ctx.info()
    return f"V{gi.num_vertices}-E{gi.num_edges}-dmax{gi.max_degree}-w{int(gi.weighted)}"


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, key: str, value: dict) -> None:
    d = _load(path)
    d[key] = value
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _batch_signature(walkers, tps, topk) -> str:
    w = np.asarray(walkers, dtype=np.float64)
    return (f"batch-n{len(w)}-wlog{np.log10(max(w.sum(), 1)):.1f}-wmax{np.log10(max(w.max(), 1)):.1f}"
            f"-tp{float(np.mean(tps)):.2f}-k{topk}")


def tune_batch(ctx, sources, walkers, tps, *, topk: int = 100, seed: int = 1, repeats: int = 3,
               rounds: int = 2, space: Optional[Dict[str, List[int]]] = None,
               cache_path: Optional[str] = DEFAULT_CACHE) -> TuneResult:
    """Tune batched PPR launches on the given query sample; leaves the context
    tuned with the best choice and returns it (dense_threshold is a planner
    argument: pass result.best['dense_threshold'] to batch_ppr)."""
    space = dict(BATCH_SPACE if space is None else space)
    if topk:
        space["topk_tile"] = [t for t in space.get("topk_tile", []) if t >= 2 * topk]
    base = ctx.tuning().as_dict()
    start = {k: base[k] for k in space if k in base}
    if "dense_threshold" in space:
        # the planner's default for the context's counting
        start["dense_threshold"] = 1 if base.get("visit_log", 1) else 100_000
    ref = ctx.batch_ppr(sources, walkers, tps, topk=topk, seed=seed, dense_threshold=start.get("dense_threshold", 0))

    def measure(c):
        tun = {k: v for k, v in c.items() if k != "dense_threshold"}
        try:
            ctx.set_tuning(**tun)
        except Exception as e:  # the engine rejects the combination
            raise ValueError(str(e))
        out = {}

        def run():
            r = ctx.batch_ppr(sources, walkers, tps, topk=topk, seed=seed,
                              dense_threshold=c.get("dense_threshold", 0))
            out["r"] = r
            return r.kernel_ms

        t = _median_ms(run, repeats)
        r = out["r"]
        if not (np.array_equal(r.topk_vertex, ref.topk_vertex) and np.array_equal(r.topk_count, ref.topk_count)
                and np.array_equal(r.steps, ref.steps)):
            raise AssertionError(f"tuning {c} changed batch results")
        return t

    res = coordinate_descent(measure, start, space, rounds=rounds)
    ctx.set_tuning(**{k: v for k, v in res.best.items() if k != "dense_threshold"})
    if cache_path:
        _store(cache_path, f"{graph_signature(ctx)}|{_batch_signature(walkers, tps, topk)}",
               {"best": res.best, "best_ms": res.best_ms, "start_ms": res.start_ms})
    return res


def tune_walk(ctx, app, walkers: int, *, seed: int = 1, repeats: int = 3, rounds: int = 2,
              space: Optional[Dict[str, List[int]]] = None,
              cache_path: Optional[str] = DEFAULT_CACHE) -> TuneResult:
    """Tune single-source / whole-graph walks of `app` with `walkers` walkers."""
    from . import walk

    space = dict(WALK_SPACE if space is None else space)
    base = ctx.tuning().as_dict()
    start = {k: base[k] for k in space}
    ref = walk(ctx, walkers, app, seed=seed)

    def measure(c):
        try:
            ctx.set_tuning(**c)
        except Exception as e:
            raise ValueError(str(e))
        out = {}

        def run():
            r = walk(ctx, walkers, app, seed=seed)
            out["r"] = r
            return r.kernel_ms

        t = _median_ms(run, repeats)
        r = out["r"]
        same = r.steps == ref.steps
        for a, b in ((r.counts, ref.counts), (r.ends, ref.ends), (r.paths, ref.paths)):
            if a is not None:
                same = same and np.array_equal(a, b)
        if not same:
            raise AssertionError(f"tuning {c} changed walk results")
        return t

    res = coordinate_descent(measure, start, space, rounds=rounds)
    ctx.set_tuning(**res.best)
    res.effective = walk(ctx, walkers, app, seed=seed).shape
    if cache_path:
        _store(cache_path, f"{graph_signature(ctx)}|walk-{type(app).__name__}-{walkers}",
               {"best": res.best, "best_ms": res.best_ms, "start_ms": res.start_ms, "effective": res.effective})
    return res


def apply(ctx, key_suffix: str, *, cache_path: str = DEFAULT_CACHE) -> Optional[Dict[str, int]]:
    """Restore a cached choice for this graph and workload signature, if any."""
    entry = _load(cache_path).get(f"{graph_signature(ctx)}|{key_suffix}")
    if not entry:
        return None
    best = entry["best"]
    ctx.set_tuning(**{k: v for k, v in best.items() if k != "dense_threshold"})
    return best


def batch_key(walkers, tps, topk: int) -> str:
    return _batch_signature(walkers, tps, topk)
