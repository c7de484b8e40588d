#!/usr/bin/env python3
"""Walk-engine benchmark: walk-steps/s on one MI355X.

Workload (one bench step): single-source PPR (visit counts) and RWR
(endpoints) on an R-MAT scale-22 graph (a/b/c/d = .57/.19/.19/.05, edge
factor 16, symmetrized, deduplicated, self-loops and isolated ids removed,
u32 CSR), from two sources -- the highest-degree vertex and one uniform
random vertex (seed 1) -- with 1e8 walkers per source and app, termination
probability 0.2, Philox seed 1.  Graph and outputs stay in HBM; the timed
region covers the four walks.

Printed: one JSON line with walk-steps/s, the roofline of the dominant
kernel (PPR visit counting: 20 algorithmic bytes per step), the measured
random-gather rate at the CSR's footprint, and the reference process timed
on the host cores (xorshift* CPU path) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the serving lanes need hardware queues of their own (HIP's default is 4,
# and the variable may already be set to that); set before HIP starts
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--walkers", type=float, default=1e8)
    ap.add_argument("--tp", type=float, default=0.2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "walk_count_rmat22_traffic.json"),
                    help="HBM bytes per launch measured with rocprofv3 --pmc (JSON), if present")
    ap.add_argument("--workload", default="ppr-rwr",
                    choices=["ppr-rwr", "deepwalk", "batch", "serve", "cli", "ingest", "node2vec", "gather", "shapes"],
                    help="ppr-rwr: the headline workload; deepwalk: R-MAT scale 24 path corpus; "
                         "batch: 4096 PPR queries in one launch on R-MAT scale 24; "
                         "serve: learned- vs fixed-budget PPR serving on R-MAT scale 22; "
                         "cli: the reference's pagerank command on R-MAT scale 16, binary vs CPU path; "
                         "ingest: loading an R-MAT scale 22 edge list, device vs host parsing; "
                         "node2vec: second-order walks on R-MAT scale 22; "
                         "gather: random u32 gather rate against table size (the Infinity Cache knee), and "
                         "the stationary-step rate on R-MAT scale 22; shapes: walk-kernel time over LDS "
                         "hub-counter / visit-cache shapes on the headline graph (the sweep behind the defaults)")
    ap.add_argument("--no-tune", action="store_true", help="batch: skip the autotuner")
    ap.add_argument("--requests", type=int, default=3000)
    ap.add_argument("--rates", default="500,20000",
                    help="serve: Poisson arrival rates (queries/s) of the mixed-eps stream; the last one saturates")
    ap.add_argument("--latency-rate", type=float, default=2000.0,
                    help="serve: arrival rate of the eps-1e-3 stream below capacity, for latency")
    ap.add_argument("--audit-rate", type=float, default=0.02)
    ap.add_argument("--eval", type=int, default=64, help="serve: requests per mode scored against FP64")
    ap.add_argument("--calibration-sources", type=int, default=512,
                    help="serve: sources of the serving graph labelled against FP64 to calibrate both budgets")
    ap.add_argument("--hoeffding-requests", type=int, default=200,
                    help="serve: requests of stream A served with the Hoeffding budget (146e6 walkers each)")
    ap.add_argument("--queries", type=int, default=4096)
    ap.add_argument("--walks-per-vertex", type=int, default=10)
    ap.add_argument("--walk-len", type=int, default=80)
    ap.add_argument("--tuning", default="",
                    help="launch tuning for the ppr-rwr and deepwalk workloads, as key=value pairs separated by "
                         "commas (ptw_tuning fields, e.g. block=256,blocks_per_cu=0)")
    return ap.parse_args()


def apply_tuning(ctx, spec):
    """Sets the ptw_tuning fields named in `spec` ("key=value,..."); returns them."""
    kw = {}
    for item in filter(None, (x.strip() for x in spec.split(","))):
        k, v = item.split("=")
        kw[k.strip()] = int(v)
    if kw:
        ctx.set_tuning(**kw)
    return kw


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the engine runs on a single device "
              "(walkers would shard by id range with one counter all-reduce; not implemented)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "deepwalk":
        return bench_deepwalk(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "cli":
        return bench_cli(args)
    if args.workload == "ingest":
        return bench_ingest(args)
    if args.workload == "node2vec":
        return bench_node2vec(args)
    if args.workload == "gather":
        return bench_gather_sweep(args)
    if args.workload == "shapes":
        return bench_shape_sweep(args)

    import numpy as np

    import ptwalker_amd as ptw

    W = int(args.walkers)
    ctx = ptw.Context(0)
    tuning = apply_tuning(ctx, args.tuning)
    t0 = time.time()
    gi = ctx.rmat(scale=args.scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1)
    build_s = time.time() - t0
    from ptwalker_amd.workloads import walk_sources

    sources = walk_sources(gi, seed=1)
    hub = sources[0]
    V = gi.num_vertices
    counts = ctx.alloc(8 * V)
    ends = ctx.alloc(4 * W)

    def one_step(record):
        steps = 0
        for s in sources:
            d = ptw.make_desc(ptw.PPR(args.tp, s), W, seed=1, flags=ptw.FLAG_OUT_DEVICE)
            o = ctx.walk(d, counts=counts.ptr, device_ptrs=True)
            steps += o.steps
            if record is not None:
                record["ppr_ms"].append(o.kernel_ms)
                record["ppr_steps"].append(o.steps)
            d = ptw.make_desc(ptw.RWR(args.tp, s), W, seed=1, flags=ptw.FLAG_OUT_DEVICE)
            o = ctx.walk(d, ends=ends.ptr, device_ptrs=True)
            steps += o.steps
            if record is not None:
                record["rwr_ms"].append(o.kernel_ms)
                record["rwr_steps"].append(o.steps)
        return steps

    for _ in range(args.warmup):
        one_step(None)
    rec = {"ppr_ms": [], "ppr_steps": [], "rwr_ms": [], "rwr_steps": []}
    # ptw_walk returns after its stream has drained, so the region is bracketed
    t1 = time.perf_counter()
    total_steps = 0
    for _ in range(args.steps):
        total_steps += one_step(rec)
    elapsed = time.perf_counter() - t1
    value = total_steps / elapsed

    # roofline of the dominant kernel: PPR visit counting, 20 B per step
    ppr_ms = float(np.mean(rec["ppr_ms"]))
    ppr_steps = float(np.mean(rec["ppr_steps"]))
    ppr_bytes = 20.0 * ppr_steps
    achieved = ppr_bytes / (ppr_ms * 1e-3) / 1e9
    rwr_ms = float(np.mean(rec["rwr_ms"]))
    rwr_steps = float(np.mean(rec["rwr_steps"]))
    rwr_gbs = (12.0 * rwr_steps + 4.0 * W) / (rwr_ms * 1e-3) / 1e9
    traffic = traffic_source = None
    req_per_step = lines_per_step = None
    if os.path.exists(args.traffic):
        with open(args.traffic) as f:
            tj = json.load(f)
        traffic = tj.get("bytes_per_launch")
        req_per_step = tj.get("memory_requests_per_step")
        lines_per_step = tj.get("l2_miss_lines_per_step")
        traffic_source = (f"not measured in this run: read from {os.path.relpath(args.traffic, ROOT)}, the "
                          "rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE of these kernels on this workload scaled by the "
                          "copy-kernel calibration (profiles/walk_count_rmat22_traffic.md)")

    # measured random-gather roofline at the CSR's col[] footprint, and the
    # stream-copy rate (also the known-size kernel for counter calibration)
    col_bytes = 4 * gi.num_edges
    gps, _ = ctx.bench_gather(col_bytes, 1 << 30, seed=1)
    copy_gbs, copy_ms = ctx.bench_copy(1 << 30)
    ppr_sps = ppr_steps / (ppr_ms * 1e-3)
    gather_frac = 2.0 * ppr_sps / gps
    # the step's own access pattern from degree-drawn vertices, steps independent
    step_sps, step_ms = ctx.bench_step(200_000_000, seed=1)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, hub, args)

    line = {
        "metric": "walk-steps/s",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic R-MAT graph generated in-process (seeded Philox); no dataset",
        "config": {
            "workload": f"single-source PPR + RWR, R-MAT scale {args.scale} ef16 symmetrized, "
                        f"sources = max-degree + 1 uniform, {W:.0e} walkers per source per app, tp {args.tp}",
            "vertices": V,
            "edges": gi.num_edges,
            "walkers_per_source": W,
            "sources": sources,
            "tp": args.tp,
            "walk_steps_per_bench_step": total_steps // args.steps,
            "graph_build_s": round(build_s, 2),
            **({"tuning": tuning} if tuning else {}),
        },
        "roofline": {
            "bound": "hbm",
            # ptw_walk's device time of a PPR launch: the walk and the visit
            # log's counting passes that finish its counts
            "kernel": "walk_kernel<count> + vlog_plan/scatter/count_kernel (PPR)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_source,
            "bytes_per_step": 20,
            "kernel_ms": ppr_ms,
            "steps_per_launch": ppr_steps,
        },
        "random_gather": {
            "gathers_per_s": gps,
            "table_bytes": col_bytes,
            "frac": gather_frac,
            "note": "(PPR steps/s x 2 random gathers per step) / measured random u32 gathers/s",
        },
        # the bound the counters show: memory requests (128 B L2-miss lines,
        # 32 B memory-side atomics and write requests), against the
        # random-gather request rate measured in this run (one 128 B line per
        # gather)
        "request_roofline": None if req_per_step is None else {
            "unit": "memory requests/s",
            "requests_per_step": req_per_step,
            "lines_per_step": lines_per_step,
            "achieved": req_per_step * ppr_sps,
            "peak": gps,
            "frac": req_per_step * ppr_sps / gps,
            "lines_frac": lines_per_step * ppr_sps / gps,
            "source": f"per-step counts from {os.path.relpath(args.traffic, ROOT)} (rocprofv3 TCC counters, "
                      "profiles/walk_kernels_rmat22_counters.md); rates from this run",
        },
        "stationary_step": {
            "steps_per_s": step_sps,
            "kernel_ms": step_ms,
            "frac": ppr_sps / step_sps,
            "rwr_frac": (rwr_steps / (rwr_ms * 1e-3)) / step_sps,
            "note": "PPR (and RWR) kernel steps/s / the rate of independent steps (row_ptr pair + one random col "
                    "entry) from 2e8 vertices drawn in proportion to degree, ptw_bench_step",
        },
        "stream_copy": {"GBs": copy_gbs, "bytes": 2 * (1 << 30), "kernel_ms": copy_ms},
        "rwr": {"kernel_ms": rwr_ms, "achieved_GBs": rwr_gbs, "bytes_per_step": 12},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    counts.free()
    ends.free()
    ctx.close()


def bench_gather_sweep(args):
    """Random u32 gathers (1e9, Philox indices, 8 in flight per lane) against the
    table size, 64 MiB to 16 GiB: tables that fit the 256 MB Infinity Cache and
    tables far beyond it show the cache's share of a random gather; then the
    stationary-step rate on the R-MAT graph (ptw_bench_step)."""
    import ptwalker_amd as ptw

    ctx = ptw.Context(0)
    rows = []
    for mib in (64, 128, 256, 513, 1024, 2048, 4096, 16384):
        gps, ms = ctx.bench_gather(mib << 20, 1 << 30, seed=1)
        rows.append({"table_MiB": mib, "gathers_per_s": gps, "kernel_ms": ms})
        print(f"# gather {mib} MiB: {gps:.3e}/s", file=sys.stderr, flush=True)
    gi = ctx.rmat(scale=args.scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1)
    sps, sms = ctx.bench_step(200_000_000, seed=1)
    copy_gbs, _ = ctx.bench_copy(1 << 30)
    line = {"metric": "random gathers/s", "value": rows[-1]["gathers_per_s"], "unit": "gathers/s", "n_gpus": 1,
            "steps": 1, "warmup": 1, "higher_is_better": True, "data": "synthetic",
            "config": {"workload": "random u32 gathers by table size; stationary steps on R-MAT scale "
                                   f"{args.scale}", "vertices": gi.num_vertices, "edges": gi.num_edges},
            "gather": rows, "stationary_step": {"steps_per_s": sps, "kernel_ms": sms},
            "stream_copy_GBs": copy_gbs}
    print(json.dumps(line), flush=True)
    ctx.close()


# (u16 LDS hub counters, visit-cache slots); the cache's slots and the hub
# counters share the LDS, so the larger hub counts leave room for fewer slots
# (lds_hubs, cache_slots, block, visit_log); blocks_per_cu stays 0, so the
# occupancy query sets the workgroups per CU from each shape's LDS
SHAPES = [(8192, 4096, 1024, 0), (36864, 1024, 1024, 0), (77824, 512, 1024, 0), (79872, 0, 1024, 0),
          (77824, -1, 1024, 0),
          (79872, 0, 1024, 1), (77824, -1, 1024, 1), (55296, 0, 1024, 1), (36864, 0, 1024, 1),
          (36864, 0, 512, 1), (24576, 0, 512, 1), (16384, 0, 512, 1), (16384, 0, 256, 1), (8192, 0, 256, 1),
          (0, 0, 256, 1), (0, 0, 1024, 1)]
SHAPE_KEYS = ("lds_hubs", "cache_slots", "block", "visit_log")
# (block, blocks_per_cu) of the endpoint kernel; 0: the occupancy query
RWR_SHAPES = [(1024, 0), (1024, 1), (512, 0), (512, 2), (512, 3), (256, 0), (256, 4)]


def bench_shape_sweep(args):
    """Walk-kernel time over launch shapes on the headline graph: PPR from the
    max-degree and from the uniform source (1e8 walkers, tp 0.2) and PageRank
    (1e8 walkers, tp 0.15); median kernel ms of 3 runs per shape, after one
    warm-up.  A shape is LDS hub counters, visit-cache slots, workgroup size
    and the visit log on or off.  Counts are compared across shapes (a shape
    changes where counts accumulate, never what is counted)."""
    import hashlib
    import statistics

    import numpy as np

    import ptwalker_amd as ptw
    from ptwalker_amd.workloads import walk_sources

    W = int(args.walkers)
    ctx = ptw.Context(0)
    gi = ctx.rmat(scale=args.scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1)
    hub, uni = walk_sources(gi, seed=1)[:2]
    default = ctx.tuning()
    apps = {"ppr_max_degree": ptw.PPR(args.tp, hub), "ppr_uniform": ptw.PPR(args.tp, uni),
            "pagerank": ptw.PageRank(0.15)}
    counts = np.zeros(gi.num_vertices, dtype=np.uint64)
    rows, digests = [], {}
    for shape_row in SHAPES:
        setting = dict(zip(SHAPE_KEYS, shape_row))
        ctx.set_tuning(**setting)
        row = dict(setting)
        for name, app in apps.items():
            ms = []
            for i in range(4):
                o = ctx.walk(ptw.make_desc(app, W, seed=1), counts=counts)
                if i:
                    ms.append(o.kernel_ms)
            shape = o.launch_shape()
            row[name] = statistics.median(ms)
            row[name + "_shape"] = {k: shape[k] for k in ("lds_hubs", "cache_slots", "chunk", "block", "blocks",
                                                          "visit_log")}
            h = hashlib.sha1(counts.tobytes()).hexdigest()
            row[name + "_identical"] = digests.setdefault(name, h) == h
        rows.append(row)
        print("# shapes " + "/".join(str(x) for x in shape_row) + ": " +
              " ".join(f"{n} {row[n]:.2f} ms" for n in apps), file=sys.stderr, flush=True)
    # the endpoint kernel has no counters: its shape is the workgroup size and
    # how many workgroups a CU holds
    ends = np.zeros(W, dtype=np.uint32)
    rwr_apps = {"rwr_max_degree": ptw.RWR(args.tp, hub), "rwr_uniform": ptw.RWR(args.tp, uni)}
    rwr_rows = []
    for block, per_cu in RWR_SHAPES:
        ctx.set_tuning(block=block, blocks_per_cu=per_cu)
        row = {"block": block, "blocks_per_cu": per_cu}
        for name, app in rwr_apps.items():
            ms = []
            for i in range(4):
                o = ctx.walk(ptw.make_desc(app, W, seed=1), ends=ends)
                if i:
                    ms.append(o.kernel_ms)
            row[name] = statistics.median(ms)
            row[name + "_shape"] = {k: v for k, v in o.launch_shape().items() if k in ("block", "blocks", "chunk")}
            h = hashlib.sha1(ends.tobytes()).hexdigest()
            row[name + "_identical"] = digests.setdefault(name, h) == h
        rwr_rows.append(row)
        print(f"# rwr shapes {block}/{per_cu}: " + " ".join(f"{n} {row[n]:.2f} ms" for n in rwr_apps),
              file=sys.stderr, flush=True)
    ctx.set_tuning(**{k: getattr(default, k) for k in SHAPE_KEYS + ("blocks_per_cu",)})
    line = {"metric": "walk kernel ms", "value": min(r["ppr_max_degree"] for r in rows), "unit": "ms", "n_gpus": 1,
            "steps": 3, "warmup": 1, "higher_is_better": False, "data": "synthetic",
            "config": {"workload": f"PPR (max-degree, uniform source; tp {args.tp}) and PageRank (tp 0.15), "
                                   f"{W:.0e} walkers, R-MAT scale {args.scale} ef16",
                       "vertices": gi.num_vertices, "edges": gi.num_edges,
                       "default": {k: getattr(default, k) for k in SHAPE_KEYS}},
            "shapes": rows, "rwr_shapes": rwr_rows}
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_deepwalk(args):
    """DeepWalk corpus: R-MAT scale 24 (default), walks_per_vertex walks from every
    vertex, walk_len vertices each, uniform and alias-weighted neighbour choice;
    paths stay in HBM.  One bench step = the unweighted corpus + the weighted one."""
    import numpy as np

    import ptwalker_amd as ptw

    scale = args.scale if args.scale != 22 else 24
    ctx = ptw.Context(0)
    tuning = apply_tuning(ctx, args.tuning)
    t0 = time.time()
    gi = ctx.rmat(scale=scale, edge_factor=16, seed=1)
    ctx.random_weights(seed=1)
    build_s = time.time() - t0
    V, L = gi.num_vertices, args.walk_len
    W = args.walks_per_vertex * V
    paths = ctx.alloc(4 * W * L)

    def run(weighted, rec):
        flags = ptw.FLAG_OUT_DEVICE | (ptw.FLAG_WEIGHTED if weighted else 0)
        d = ptw.make_desc(ptw.DeepWalk(L, start=ptw.START_ROUNDROBIN), W, seed=1, flags=flags)
        o = ctx.walk(d, paths=paths.ptr, device_ptrs=True)
        if rec is not None:
            rec.append((weighted, o.kernel_ms, o.steps))
        return o.steps

    for _ in range(args.warmup):
        run(False, None)
        run(True, None)
    rec = []
    t1 = time.perf_counter()
    total = 0
    for _ in range(args.steps):
        total += run(False, rec) + run(True, rec)
    elapsed = time.perf_counter() - t1
    uw = [r for r in rec if not r[0]]
    wt = [r for r in rec if r[0]]
    uw_ms = float(np.mean([r[1] for r in uw]))
    wt_ms = float(np.mean([r[1] for r in wt]))
    steps = float(uw[0][2])
    gps, _ = ctx.bench_gather(4 * gi.num_edges, 1 << 30, seed=1)
    line = {
        "metric": "walk-steps/s",
        "value": total / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic R-MAT graph and U(0,1] weights generated in-process (seeded Philox); no dataset",
        "config": {"workload": f"DeepWalk corpus, R-MAT scale {scale} ef16, {args.walks_per_vertex} walks per vertex "
                               f"x {L} vertices, uniform + alias-weighted, paths resident in HBM",
                   "vertices": V, "edges": gi.num_edges, "walkers": W, "path_bytes": 4 * W * L,
                   "graph_and_alias_build_s": round(build_s, 1), **({"tuning": tuning} if tuning else {})},
        "roofline": {"bound": "hbm", "kernel": "walk_kernel<path> (unweighted)",
                     "achieved": 16.0 * steps / (uw_ms * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": 16.0 * steps / (uw_ms * 1e-3) / 1e9 / 8000.0, "traffic": None, "bytes_per_step": 16,
                     "kernel_ms": uw_ms, "steps_per_launch": steps},
        # weighted: row_ptr (8 B) + one 16 B weighted-edge record + 4 B path store
        "weighted": {"kernel_ms": wt_ms, "bytes_per_step": 28,
                     "achieved": 28.0 * steps / (wt_ms * 1e-3) / 1e9},
        "random_gather": {"gathers_per_s": gps, "frac_unweighted": 2.0 * steps / (uw_ms * 1e-3) / gps,
                          "frac_weighted": 2.0 * steps / (wt_ms * 1e-3) / gps},
        "cpu_baseline": None if args.no_cpu else cpu_baseline_deepwalk(ctx, L, args),
    }
    print(json.dumps(line), flush=True)
    paths.free()
    ctx.close()


def cpu_cores():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 16))


def cpu_baseline_deepwalk(ctx, L, args):
    """The reference process (oracle: xorshift*, OpenMP over walkers) walking
    unweighted L-vertex walks from uniform starts on the same graph, on a
    bounded walker sample.  It stores no paths, so it bounds the CPU rate
    from above."""
    import oracle

    rp, col = ctx.export_csr()
    cores = cpu_cores()
    def run(n):
        t = time.perf_counter()
        r = oracle.ref_walk(rp, col, oracle.APP_DEEPWALK, n, walk_len=L, seed=1, threads=cores)
        return r, time.perf_counter() - t

    # a probe long enough to time, and one resize if the sample falls short
    probe = 200_000
    r, dt = run(probe)
    n = int(min(2e8, max(probe, r["steps"] / max(dt, 1e-9) * args.cpu_seconds / max(L - 1, 1))))
    r, dt = run(n)
    if dt < 0.5 * args.cpu_seconds and n < 2e8:
        n = int(min(2e8, n * args.cpu_seconds / max(dt, 1e-3)))
        r, dt = run(n)
    return {"value": r["steps"] / dt, "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"{n} unweighted walks of {L} vertices from uniform starts ({r['steps']} steps, {dt:.1f} s), "
                      f"xorshift* reference CPU path without path stores, {cores} threads"}


def batch_queries(gi, rp, n, seed=1):
    from ptwalker_amd.workloads import batch_queries as bq

    return bq(gi, rp, n, seed)


def bench_batch(args):
    """Batched PPR: `queries` single-source queries in one persistent launch with
    per-query counter slices (dense above 1e5 walkers, hash below) and top-100
    per query.  One bench step = one batch; run eagerly and as a replayed hipGraph."""
    import numpy as np

    import ptwalker_amd as ptw

    scale = args.scale if args.scale != 22 else 24
    ctx = ptw.Context(0)
    tuning = apply_tuning(ctx, args.tuning)
    t0 = time.time()
    gi = ctx.rmat(scale=scale, edge_factor=16, seed=1)
    build_s = time.time() - t0
    rp, _ = ctx.export_csr()
    src, W, tp = batch_queries(gi, rp, args.queries, seed=1)

    def measure(dense_threshold=0):
        res = {}
        for mode in ("eager", "graph"):
            for _ in range(args.warmup):
                ctx.batch_ppr(src, W, tp, topk=100, seed=1, graph=(mode == "graph"), dense_threshold=dense_threshold)
            t1 = time.perf_counter()
            ks = []
            for _ in range(args.steps):
                r = ctx.batch_ppr(src, W, tp, topk=100, seed=1, graph=(mode == "graph"),
                                  dense_threshold=dense_threshold)
                ks.append(r.kernel_ms)
            el = (time.perf_counter() - t1) / args.steps
            res[mode] = (el, float(np.mean(ks)), r)
        return res

    ctx_default_shape = ctx.tuning().as_dict()
    untuned = measure()
    tuned_cfg = None
    res = untuned
    if not args.no_tune:
        from ptwalker_amd import tune

        t0 = time.time()
        # the median of three runs per candidate: shapes differ by about 1 %,
        # as much as one run's noise
        tr = tune.tune_batch(ctx, src, W, tp, topk=100, repeats=3, rounds=1, cache_path=None)
        tuned_cfg = {"choice": tr.best, "tuning_s": round(time.time() - t0, 1), "candidates": len(tr.trials)}
        res = measure(dense_threshold=tr.best.get("dense_threshold", 0))
    el, kms, r = res["graph"]
    audit = audit_batch(ctx, src, W, tp, r)
    gps, _ = ctx.bench_gather(4 * gi.num_edges, 1 << 30, seed=1)
    line = {
        "metric": "walk-steps/s",
        "value": r.total_steps / el,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic R-MAT graph and query mix generated in-process (seeded); no dataset",
        "config": {"workload": f"{args.queries} PPR queries, R-MAT scale {scale} ef16, sources degree-weighted, "
                               "walkers log-U[1e3,1e6], tp U[0.15,0.3], top-100 per query, one captured launch",
                   "vertices": gi.num_vertices, "edges": gi.num_edges, "walkers": int(W.sum()),
                   "dense_queries": r.dense_queries, "hash_queries": r.hash_queries,
                   "slice_bytes": r.slice_bytes, "graph_build_s": round(build_s, 1)},
        "queries_per_s": args.queries / el,
        "kernel_ms": kms,
        "random_gather": {"gathers_per_s": gps, "table_bytes": 4 * gi.num_edges,
                          "frac": 2.0 * r.total_steps / el / gps,
                          "note": "(batch steps/s x 2 random gathers per step) / measured random u32 gathers/s"},
        "eager": {"ms_per_batch": 1e3 * res["eager"][0], "kernel_ms": res["eager"][1]},
        "fp64_audit": audit,
        "autotune": tuned_cfg,
        "untuned": {"steps_per_s": untuned["graph"][2].total_steps / untuned["graph"][0],
                    "kernel_ms": untuned["graph"][1], "shape": ctx_default_shape},
        # algorithmic bytes of the batched walk: 12 B gather + 4 B u32 slice add
        # per step, over the whole captured batch (walks, compaction, top-k, reset)
        "cpu_baseline": None if args.no_cpu else cpu_baseline_batch(ctx, src, W, tp, args),
        "roofline": {"bound": "hbm", "kernel": "whole batch (batch_dense_kernel, batch_walk_kernel, compaction, "
                                              "top-k, reset)",
                     "achieved": 16.0 * r.total_steps / (kms * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": 16.0 * r.total_steps / (kms * 1e-3) / 1e9 / 8000.0, "traffic": None,
                     "bytes_per_step": 16},
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def audit_batch(ctx, src, W, tp, r, n=64, k=100):
    """Scores n queries spread over the last batch against the FP64 solver:
    the top-k L1 of the visit distribution, read from the batch's own slices at
    the exact top-k vertices, and how many of the batch's top-k vertices are in
    the exact top-k."""
    import numpy as np

    idx = np.linspace(0, len(src) - 1, min(n, len(src))).astype(np.int64)
    t0 = time.perf_counter()
    _, tv, tval, iters = ctx.ppr_exact(src[idx], tp[idx], tol=1e-7, full=False, k=k)
    solve_s = time.perf_counter() - t0
    cnt = ctx.batch_gather(np.repeat(idx.astype(np.uint32), k), tv.reshape(-1)).reshape(len(idx), k)
    visits = W[idx].astype(np.float64) + r.steps[idx].astype(np.float64)
    est = cnt.astype(np.float64) / visits[:, None]
    l1 = np.abs(est - tval / (1.0 + 1.0 / tp[idx])[:, None]).sum(axis=1)
    overlap = np.array([len(np.intersect1d(r.topk_vertex[q], tv[j])) / k for j, q in enumerate(idx)])
    return {"queries": int(len(idx)), "k": k, "solver_iterations": int(iters), "solver_s": round(solve_s, 2),
            "walkers_min": int(W[idx].min()), "walkers_max": int(W[idx].max()),
            "topk_l1_median": float(np.median(l1)), "topk_l1_max": float(l1.max()),
            "topk_l1_median_above_1e5_walkers": float(np.median(l1[W[idx] > 100000])) if (W[idx] > 100000).any()
            else None,
            "topk_overlap_mean": round(float(overlap.mean()), 4)}


def cpu_baseline_batch(ctx, src, W, tp, args):
    """The reference process runs one PPR query at a time (oracle: xorshift*,
    OpenMP over the query's walkers): the batch's queries in order until the
    sample reaches about cpu_seconds."""
    import oracle

    rp, col = ctx.export_csr()
    cores = cpu_cores()
    steps = done = 0
    t = time.perf_counter()
    for i in range(len(src)):
        r = oracle.ref_walk(rp, col, oracle.APP_PPR, int(W[i]), tp=float(tp[i]), source=int(src[i]),
                            start_uniform=False, seed=1, threads=cores)
        steps += r["steps"]
        done += 1
        if time.perf_counter() - t > args.cpu_seconds:
            break
    dt = time.perf_counter() - t
    return {"value": steps / dt, "unit": "steps/s", "cores": cores, "kind": "port",
            "queries_per_s": done / dt,
            "sample": f"the first {done} of the batch's queries one after another ({steps} steps, {dt:.1f} s), "
                      f"xorshift* reference-semantics CPU path, {cores} threads"}


def serve_requests(rp, n, rate, seed=1, eps_choices=(1e-3, 3e-4)):
    from ptwalker_amd.workloads import serve_requests as sr

    return sr(rp, n, rate, seed, eps_choices)


def score_served(ctx, res, src, eps, tp, seed, sample, k=100):
    """Top-k L1 of the served answers against the FP64 solver: each sampled
    request's answering walk (main, or its 10x audit when topped up) is rerun
    alone -- bit-identical to the served one -- and its counts are read at
    the exact top-k vertices."""
    import numpy as np

    idx = np.linspace(0, len(src) - 1, min(sample, len(src))).astype(np.int64)
    s = src[idx]
    w = np.where(res.topped_up[idx], 10 * res.budget[idx], res.budget[idx]).astype(np.int64)
    streams = np.where(res.topped_up[idx], (1 << 31) + idx, idx)
    _, tv, tval, _ = ctx.ppr_exact(s, [tp] * len(s), tol=1e-7, full=False, k=k)
    r = ctx.batch_ppr(s, w, [tp] * len(s), streams=streams, topk=0, seed=seed)
    cnt = ctx.batch_gather(np.repeat(np.arange(len(s), dtype=np.uint32), k), tv.reshape(-1)).reshape(len(s), k)
    est = cnt.astype(np.float64) / (w + r.steps.astype(np.float64))[:, None]
    l1 = np.abs(est - tval / (1.0 + 1.0 / tp)).sum(axis=1)
    e = eps[idx]
    return {"requests": int(len(s)), "mean_l1_over_eps": round(float(np.mean(l1 / e)), 4),
            "p95_l1_over_eps": round(float(np.quantile(l1 / e, 0.95)), 4),
            "frac_within_eps": round(float(np.mean(l1 <= e)), 4), "max_l1_over_eps": round(float(np.max(l1 / e)), 4)}


def bench_serve(args):
    """Serving: requests (source, eps) arrive by a Poisson process and are served by
    continuous batching (ptw_serve_replay).  Budgets: learned (committed policy:
    pilot walk + MLP on the device, calibrated to cover 95 % of held-out queries),
    calibrated fixed (fixed_c2 / eps^2 covering the same 95 % of training queries)
    and Hoeffding fixed (the distribution-free L1 bound, no training data).

    Stream A: eps = 1e-3 for every request, the metric's own target (top-100
    L1 <= 1e-3 vs FP64); learned vs both fixed budgets.  Stream B: eps uniform
    over {1e-3, 3e-4} at each rate in --rates; learned vs calibrated fixed (the
    Hoeffding budget at 3e-4, 1.6e9 walkers, exceeds the u32 counter slices)."""
    import numpy as np

    import ptwalker_amd as ptw
    from ptwalker_amd.budget import BudgetPolicy, calibrate, hoeffding_c2

    trained = BudgetPolicy.load(os.path.join(ROOT, "ptwalker_amd", "models", "budget_policy.json"))
    tp = float(trained.meta["tp"])
    ctx = ptw.Context(0)
    t0 = time.time()
    gi = ctx.rmat(scale=args.scale, edge_factor=16, seed=1)
    build_s = time.time() - t0
    rp, _ = ctx.export_csr()
    net = trained.network(ctx)
    # calibrate both budgets to 95 % coverage on this graph (sources disjoint
    # in seed from the served stream), outside the timed region
    t0 = time.time()
    cal_src, _, _ = serve_requests(rp, args.calibration_sources, 1.0, seed=99)
    policy, cal_stats = calibrate(ctx, trained, net, cal_src)
    cal_stats["seconds"] = round(time.time() - t0, 1)
    print(f"# calibration on {len(cal_src)} sources: margin {trained.margin:.3f} -> {policy.margin:.3f}, fixed_c2 "
          f"{trained.fixed_c2:.3f} -> {policy.fixed_c2:.3f}, work ratio {cal_stats['work_ratio_fixed_over_learned']:.3f}",
          file=sys.stderr, flush=True)
    h_c2 = hoeffding_c2()
    common = dict(tp=tp, seed=1, topk=100, max_batch=512)
    modes = {
        "learned": dict(mlp=net, budget_mode=ptw.BUDGET_LEARNED, margin=policy.margin, feat_lo=policy.feat_lo,
                        feat_hi=policy.feat_hi, fixed_c2=policy.fixed_c2, w_min=policy.w_min, w_max=policy.w_max,
                        audit_rate=args.audit_rate),
        "fixed": dict(budget_mode=ptw.BUDGET_FIXED, fixed_c2=policy.fixed_c2, w_min=policy.w_min,
                      w_max=policy.w_max, audit_rate=args.audit_rate),
        "hoeffding": dict(budget_mode=ptw.BUDGET_FIXED, fixed_c2=h_c2, w_min=policy.w_min,
                          w_max=int(np.ceil(h_c2 / 1e-3 ** 2)), audit_rate=0.0),
    }

    def one(name, src, eps, arr, score):
        kw = modes[name]
        for _ in range(args.warmup):
            m = min(64, len(src))
            ctx.serve(src[:m], eps[:m], np.zeros(m), **common, **kw)
        res = ctx.serve(src, eps, arr, **common, **kw)
        st = res.stats
        rec = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}
        rec["requests"] = len(src)
        rec["mean_budget"] = round(float(res.budget.mean()), 1)
        if score and args.eval:
            rec["accuracy"] = score_served(ctx, res, src, eps, tp, 1, args.eval)
        print(f"# {name}: {len(src)} requests, {st['queries_per_s']:.1f} q/s, p50 {st['p50_ms']:.1f} ms, "
              f"p99 {st['p99_ms']:.1f} ms, {st['launches']} launches, mean budget {rec['mean_budget']:.3g}",
              file=sys.stderr, flush=True)
        return rec

    rates = [float(x) for x in args.rates.split(",")]
    top = rates[-1]
    stream_a = {}
    # saturated: every request waiting from the start of its burst, so
    # queries/s is the capacity at equal accuracy
    src, eps, arr = serve_requests(rp, args.requests, top, seed=2, eps_choices=(1e-3,))
    for name in ("learned", "fixed"):
        stream_a[name] = one(name, src, eps, arr, True)
    nh = max(1, min(args.requests, args.hoeffding_requests))
    stream_a["hoeffding"] = one("hoeffding", src[:nh], eps[:nh], arr[:nh], True)
    # below capacity: latency of requests that do not queue behind a backlog
    stream_a_latency = {}
    src, eps, arr = serve_requests(rp, args.requests, args.latency_rate, seed=3, eps_choices=(1e-3,))
    for name in ("learned", "fixed"):
        stream_a_latency[name] = one(name, src, eps, arr, False)
    stream_b = {}
    for rate in rates:
        src, eps, arr = serve_requests(rp, args.requests, rate, seed=1)
        for name in ("learned", "fixed"):
            stream_b[f"{name}@{int(rate)}"] = one(name, src, eps, arr, rate == top)
    L = stream_a["learned"]
    gather_gps, _ = ctx.bench_gather(int(rp[-1]) * 4 + 4 * len(rp), 1 << 28)
    line = {
        "metric": "PPR queries/s at top-100 L1 <= 1e-3 vs FP64",
        "value": L["queries_per_s"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * L["makespan_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic R-MAT graph and request stream generated in-process (seeded); no dataset",
        "config": {"workload": f"PPR serving, R-MAT scale {args.scale} ef16 resident, top-100, tp {tp}, continuous "
                               f"batching over two lanes (max 512 per batch); stream A: {args.requests} requests, "
                               f"eps 1e-3, Poisson {int(top)} q/s (saturated) and {int(args.latency_rate)} q/s "
                               f"(below capacity); stream B: eps U{{1e-3, 3e-4}}, Poisson {args.rates} q/s; "
                               f"audit rate {args.audit_rate}",
                   "vertices": gi.num_vertices, "edges": gi.num_edges, "graph_build_s": round(build_s, 1),
                   "policy": {"margin": policy.margin, "fixed_c2": policy.fixed_c2, "hoeffding_c2": h_c2,
                              "trained_margin": trained.margin, "trained_fixed_c2": trained.fixed_c2,
                              "train_pairs": trained.meta.get("pairs"),
                              "val_rmse_log": trained.meta.get("val_rmse_log")},
                   "calibration": cal_stats},
        # equal accuracy: both budgets calibrated on this graph to the same 95 %
        # coverage of top-100 L1 <= eps; the Hoeffding budget is the plain
        # method's distribution-free bound
        "learned_over_fixed_equal_accuracy": round(L["queries_per_s"] / stream_a["fixed"]["queries_per_s"], 3),
        "learned_over_hoeffding": round(L["queries_per_s"] / stream_a["hoeffding"]["queries_per_s"], 3),
        "stream_a_eps_1e-3": stream_a,
        "stream_a_eps_1e-3_below_capacity": stream_a_latency,
        "stream_b_mixed_eps": stream_b,
        "walk_steps_per_s": L["walk_steps"] / L["makespan_s"],
        "gather_frac": 2.0 * L["walk_steps"] / L["makespan_s"] / gather_gps,
        # 16 B per batched step (12 B gather + 4 B u32 slice add) over the
        # serving makespan, host work between launches included
        "roofline": {"bound": "hbm", "kernel": "serving loop (batched walk kernels, compaction, top-k per launch)",
                     "achieved": 16.0 * L["walk_steps"] / L["makespan_s"] / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": 16.0 * L["walk_steps"] / L["makespan_s"] / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_step": 16},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    net.close()
    ctx.close()


def bench_cli(args):
    """The reference's own command on its CPU-sized case: R-MAT scale 16 (ef 16,
    symmetrized, seed 1) written as a text edge list, then
    `pagerank -f text -g G -w 655360 --sp 0.15 -s 1 -t 20` -- each run a fresh
    process of the GPU binary, timed over the reference's printed scope (walker
    init + walk, graph load excluded).  The CPU baseline is the reference
    process (xorshift*, quantized geometric, OpenMP) on the same graph, at the
    command's 20 threads and at all usable cores.  This process never touches
    the GPU itself: the graph is generated by a host-only context."""
    import statistics
    import subprocess
    import tempfile

    import oracle
    import ptwalker_amd as ptw

    h = ptw.Context(-1)
    gi = h.rmat(scale=16, edge_factor=16, seed=1)
    rp, col = h.export_csr()
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "rmat16.txt")
    h.write(path, ptw.FORMAT_TEXT)
    h.close()
    binp = os.path.join(ROOT, "ptwalker_amd", "bin", "pagerank")
    cmd = [binp, "-f", "text", "-g", path, "-w", "655360", "--sp", "0.15", "-s", "1", "-t", "20", "--json"]
    recs = []
    for i in range(args.warmup + args.steps):
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        if out.returncode != 0:
            print(out.stdout + out.stderr, file=sys.stderr)
            sys.exit(out.returncode)
        rec = json.loads([ln for ln in out.stdout.splitlines() if ln.startswith("{")][-1])
        if i >= args.warmup:
            recs.append(rec)
    walk_s = statistics.median(r["walk_s"] for r in recs)
    kernel_ms = statistics.median(r["kernel_ms"] for r in recs)
    steps = recs[0]["steps"]
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    cpu = {}
    for th in sorted({20, cores}):
        ts = []
        for _ in range(3):
            t = time.perf_counter()
            r = oracle.ref_walk(rp, col, oracle.APP_PAGERANK, 655360, tp=0.15, start_uniform=True, seed=1, threads=th)
            ts.append(time.perf_counter() - t)
        cpu[th] = (r["steps"] / statistics.median(ts), statistics.median(ts), r["steps"])
    best_th = max(cpu, key=lambda k: cpu[k][0])
    # the reference's own schedule (partitions from this host's L2, walker
    # pools, migration, buffered hub sampling), walker init + walk scope
    part = {}
    for th in sorted({20, cores}):
        ts = []
        for _ in range(3):
            p = oracle.ref_partitioned_walk(rp, col, oracle.APP_PAGERANK, 655360, tp=0.15, start_uniform=True, seed=1,
                                            threads=th)
            ts.append(p["seconds"])
        part[str(th)] = p["steps"] / statistics.median(ts)
    line = {
        "metric": "walk-steps/s",
        "value": steps / walk_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * walk_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic R-MAT graph generated in-process (seeded) and written as a text edge list; no dataset",
        "config": {"workload": "pagerank -f text -g rmat16.txt -w 655360 --sp 0.15 -s 1 -t 20 (R-MAT scale 16 ef16 "
                               "symmetrized, seed 1), one process per run, reference timing scope",
                   "vertices": gi.num_vertices, "edges": gi.num_edges, "walk_steps": steps},
        "kernel_ms": kernel_ms,
        "kernel_steps_per_s": steps / (kernel_ms * 1e-3),
        "walk_s_runs": [r["walk_s"] for r in recs],
        "roofline": {"bound": "hbm", "kernel": "walk_kernel<count>", "achieved": 20.0 * steps / (kernel_ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": 20.0 * steps / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_step": 20, "note": "the graph (4 MB CSR) is cache-resident; launch-bound"},
        "cpu_baseline": {"value": cpu[best_th][0], "unit": "steps/s", "cores": best_th, "kind": "port",
                         "sample": f"the whole command: 655360 walkers, {cpu[best_th][2]} steps, "
                                   f"{cpu[best_th][1] * 1e3:.1f} ms (median of 3), reference process "
                                   f"(xorshift*, quantized geometric), {best_th} threads",
                         "by_threads": {str(k): v[0] for k, v in cpu.items()},
                         "partitioned_by_threads": part},
        "vs_cpu": (steps / walk_s) / cpu[best_th][0],
    }
    print(json.dumps(line), flush=True)


def bench_ingest(args):
    """Graph load (ptw_graph_load) of an R-MAT edge list written as text and as
    binary: the file parsed, the CSR built and degree-relabelled on the device
    (the graph-id CSR and the relabel copied back as the host mirror) against
    the host threads (mmap + parallel parse, relabel, upload).  The reference
    reads the text with one thread
    (include/io.hpp:76-90).  Files are written to a temporary directory first
    (not timed); each load is timed from the call to the graph being resident."""
    import shutil
    import statistics
    import tempfile

    import ptwalker_amd as ptw

    tmp = tempfile.mkdtemp()
    try:
        h = ptw.Context(-1)
        gi = h.rmat(scale=args.scale, edge_factor=16, seed=1)
        paths = {"text": os.path.join(tmp, "g.txt"), "binary": os.path.join(tmp, "g.bin")}
        h.write(paths["text"], ptw.FORMAT_TEXT)
        h.write(paths["binary"], ptw.FORMAT_BINARY)
        h.close()
        ctx = ptw.Context(0)
        res = {}
        for fmt_name, fmt in (("text", ptw.FORMAT_TEXT), ("binary", ptw.FORMAT_BINARY)):
            for where, flag in (("device", ptw.GRAPH_INGEST_DEVICE), ("host", ptw.GRAPH_INGEST_HOST)):
                ts = []
                for i in range(args.warmup + args.steps):
                    t = time.perf_counter()
                    g2 = ctx.load(paths[fmt_name], fmt, flags=flag)
                    if i >= args.warmup:
                        ts.append(time.perf_counter() - t)
                assert g2.num_edges == gi.num_edges
                res[f"{fmt_name}_{where}_s"] = statistics.median(ts)
                print(f"# {fmt_name} {where}: {statistics.median(ts):.3f} s", file=sys.stderr, flush=True)
        size = os.path.getsize(paths["text"])
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        line = {
            "metric": "edges/s loaded",
            "value": gi.num_edges / res["text_device_s"],
            "unit": "edges/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": 1e3 * res["text_device_s"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "u32",
            "data": "synthetic R-MAT graph written as text and binary edge lists in-process (seeded); no dataset",
            "config": {"workload": f"ptw_graph_load of R-MAT scale {args.scale} ef16 symmetrized "
                                   f"({gi.num_edges} lines, {size / 1e9:.2f} GB of text), load to resident graph",
                       "vertices": gi.num_vertices, "edges": gi.num_edges, "text_bytes": size},
            "seconds": {k: round(v, 3) for k, v in res.items()},
            "device_over_host": {"text": res["text_host_s"] / res["text_device_s"],
                                 "binary": res["binary_host_s"] / res["binary_device_s"]},
            "roofline": {"bound": "hbm", "kernel": "whole load (file read and copy, parse, sorts, relabel, host mirror)",
                         "achieved": size / res["text_device_s"] / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": size / res["text_device_s"] / 1e9 / HBM_PEAK_GBS, "traffic": None,
                         "note": "end-to-end input bytes per second; the host file read, PCIe and the host "
                                 "mirror's download bound it"},
            "cpu_baseline": {"value": gi.num_edges / res["text_host_s"], "unit": "edges/s",
                             "cores": min(cores, 16), "kind": "port",
                             "sample": "the same load with the file parsed and the CSR built by host threads "
                                       "(mmap + parallel parse), relabelled on the host and uploaded"},
        }
        print(json.dumps(line), flush=True)
        ctx.close()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def bench_node2vec(args):
    """Node2vec corpus: 10 walks per vertex of 80 vertices on R-MAT scale 22,
    paths resident in HBM, for (p, q) = (0.5, 2), (2, 0.5) and (1, 1) (the
    last equals DeepWalk), beside DeepWalk on the same graph.  The CPU
    baseline is the oracle's node2vec replay (the same decisions, OpenMP) on
    a bounded walker sample."""
    import statistics

    import numpy as np

    import oracle
    import ptwalker_amd as ptw

    ctx = ptw.Context(0)
    gi = ctx.rmat(scale=args.scale, edge_factor=16, seed=1)
    W = 10 * gi.num_vertices
    L = 80
    paths = ctx.alloc(W * L * 4)
    runs = {}
    for name, app in (("p0.5_q2", ptw.Node2Vec(L, 0.5, 2.0)), ("p2_q0.5", ptw.Node2Vec(L, 2.0, 0.5)),
                      ("p1_q1", ptw.Node2Vec(L, 1.0, 1.0)), ("deepwalk", ptw.DeepWalk(L))):
        d = ptw.make_desc(app, W, seed=1, flags=ptw.FLAG_OUT_DEVICE)
        ks, ts = [], []
        for i in range(args.warmup + args.steps):
            t = time.perf_counter()
            o = ctx.walk(d, paths=paths.ptr, device_ptrs=True)
            el = time.perf_counter() - t
            if i >= args.warmup:
                ks.append(o.kernel_ms)
                ts.append(el)
        runs[name] = {"kernel_ms": statistics.median(ks), "ms": 1e3 * statistics.median(ts), "steps": o.steps,
                      "steps_per_s": o.steps / (statistics.median(ks) * 1e-3)}
        print(f"# {name}: {runs[name]['kernel_ms']:.1f} ms, {runs[name]['steps_per_s']:.3g} steps/s",
              file=sys.stderr, flush=True)
    main = runs["p0.5_q2"]
    cpu = None
    if not args.no_cpu:
        rp, col = ctx.export_csr()
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        cores = max(1, min(cores, 16))
        # every replay first sorts the rows and checks the graph (the device
        # does this once, outside its timing): time it alone and leave it out
        t = time.perf_counter()
        oracle.node2vec_replay(rp, col, 1, L, 0.5, 2.0, seed=1, threads=cores)
        setup = time.perf_counter() - t
        n = 200000  # long enough that the setup's run-to-run spread does not skew the rate
        t = time.perf_counter()
        r = oracle.node2vec_replay(rp, col, n, L, 0.5, 2.0, seed=1, threads=cores)
        dt = max(time.perf_counter() - t - setup, 1e-3)
        n = int(min(W, max(n, n * args.cpu_seconds / dt)))
        t = time.perf_counter()
        r = oracle.node2vec_replay(rp, col, n, L, 0.5, 2.0, seed=1, threads=cores)
        dt = max(time.perf_counter() - t - setup, 1e-3)
        cpu = {"value": r["steps"] / dt, "unit": "steps/s", "cores": cores, "kind": "port",
               "sample": f"{n} node2vec walks (p 0.5, q 2) of {L} vertices, {r['steps']} steps, {dt:.1f} s "
                         f"after {setup:.1f} s of row sorting left out, the oracle's replay of the same "
                         f"decisions, {cores} threads"}
    line = {
        "metric": "walk-steps/s",
        "value": main["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic R-MAT graph generated in-process (seeded); no dataset",
        "config": {"workload": f"node2vec corpus, R-MAT scale {args.scale} ef16, 10 walks per vertex x {L} vertices, "
                               "p 0.5, q 2, paths resident in HBM",
                   "vertices": gi.num_vertices, "edges": gi.num_edges, "walkers": W},
        "runs": runs,
        "roofline": {"bound": "hbm", "kernel": "walk_kernel<node2vec>",
                     "achieved": 16.0 * main["steps_per_s"] / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": 16.0 * main["steps_per_s"] / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_step": 16,
                     "note": "16 B per accepted step as for DeepWalk; rejected proposals and membership "
                             "searches add reads not counted here"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    paths.free()
    ctx.close()


def cpu_baseline(ctx, hub, args):
    """The reference process with its own RNG (oracle: xorshift*, quantized geometric,
    OpenMP over walkers), PPR from the same source, on a bounded walker sample."""
    import oracle

    rp, col = ctx.export_csr()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    def run(n):
        t = time.perf_counter()
        r = oracle.ref_walk(rp, col, oracle.APP_PPR, n, tp=args.tp, source=hub, start_uniform=False, seed=1,
                            threads=cores)
        return r, time.perf_counter() - t

    # size the sample from a probe long enough to time (thread start-up
    # dominates a short one), and once more if it still falls well short
    probe = 2_000_000
    r, dt = run(probe)
    n = int(min(5e8, max(probe, r["steps"] / max(dt, 1e-9) * args.cpu_seconds * args.tp)))
    r, dt = run(n)
    if dt < 0.5 * args.cpu_seconds and n < 5e8:
        n = int(min(5e8, n * args.cpu_seconds / max(dt, 1e-3)))
        r, dt = run(n)
    naive = {"steps_per_s": r["steps"] / dt, "walkers": n, "steps": r["steps"], "seconds": dt}
    # the reference's own schedule on the same sample: partitions sized from
    # this host's L2, walker pools, migration, buffered hub sampling; timed
    # over walker init + walk, the scope the reference prints
    t = time.perf_counter()
    p = oracle.ref_partitioned_walk(rp, col, oracle.APP_PPR, n, tp=args.tp, source=hub, start_uniform=False, seed=1,
                                    threads=cores)
    part = {"steps_per_s": p["steps"] / p["seconds"], "walkers": n, "steps": p["steps"], "seconds": p["seconds"],
            "with_partitioning_s": time.perf_counter() - t, "hub_partitions": p["hub_partitions"],
            "low_partitions": p["low_partitions"], "l2_bytes": oracle.l2_cache_bytes()}
    best = "partitioned" if part["steps_per_s"] > naive["steps_per_s"] else "walker_parallel"
    b = part if best == "partitioned" else naive
    return {
        "value": b["steps_per_s"],
        "unit": "steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"PPR from the max-degree vertex, {n} walkers ({b['steps']} steps, {b['seconds']:.1f} s), "
                  f"reference RNG (xorshift*, quantized geometric), {cores} threads, the faster of two CPU "
                  f"schedules: {best}",
        "schedules": {"walker_parallel": naive, "partitioned": part},
    }


if __name__ == "__main__":
    main()
