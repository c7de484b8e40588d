"""Synthetic workloads of the measured configurations, shared by bench.py and
the tests so that both run exactly the same inputs.

  single-source walks  R-MAT scale 22 (a/b/c = .57/.19/.19, edge factor 16,
                       generator seed 1); sources: the max-degree vertex and
                       one uniform vertex (numpy seed 1); 1e8 walkers per
                       source, tp 0.2, PPR (visit counts) and RWR (endpoints)
  DeepWalk corpus      R-MAT scale 24; 10 walks per vertex (round-robin
                       starts) of 80 vertices; uniform and alias-weighted
                       (U(0, 1] weights, seed 1)
  batched PPR          4096 queries on scale 24: sources drawn in proportion
                       to degree, walker budgets log-uniform in [1e3, 1e6],
                       tp uniform in [0.15, 0.3]
  serving              requests on scale 22: half degree-weighted, half
                       uniform sources, eps uniform over {1e-3, 3e-4},
                       Poisson arrivals

The reference has one command per app and one query per process
(src/core/ppr.cpp:121-139, src/core/RWR.cpp:109-128); these shapes are the
measured configurations built on that path.
"""
from __future__ import annotations

import numpy as np

RMAT_ABC = (0.57, 0.19, 0.19)


def walk_sources(gi, seed: int = 1):
    """The max-degree vertex and one uniform vertex of the graph."""
    rng = np.random.default_rng(seed)
    return [int(gi.max_degree_vertex), int(rng.integers(0, gi.num_vertices))]


def batch_queries(gi, rp, n, seed=1):
    """Sources drawn with probability proportional to degree, walker budgets
    log-uniform in [1e3, 1e6], termination probabilities uniform in [0.15, 0.3]."""
    rng = np.random.default_rng(seed)
    deg = np.diff(rp.astype(np.int64)).astype(np.float64)
    src = rng.choice(gi.num_vertices, size=n, p=deg / deg.sum())
    W = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n)).astype(np.int64)
    tp = rng.uniform(0.15, 0.3, n)
    return src, W, tp


def serve_requests(rp, n, rate, seed=1, eps_choices=(1e-3, 3e-4)):
    """Half degree-weighted, half uniform sources; eps uniform over `eps_choices`;
    Poisson arrivals at `rate` queries/s."""
    rng = np.random.default_rng(seed)
    deg = np.diff(rp.astype(np.int64)).astype(np.float64)
    src = np.concatenate([rng.choice(len(deg), n // 2, p=deg / deg.sum()), rng.integers(0, len(deg), n - n // 2)])
    rng.shuffle(src)
    eps = rng.choice(list(eps_choices), n)
    arr = np.cumsum(rng.exponential(1.0 / rate, n))
    return src, eps, arr
