"""Statistical checks shared by the tests.

* transition_check: the KnightKing static-walk validator
  (third_party/knightking/test_walk.hpp:145-179): every consecutive pair of
  a path is an edge, the empirical transition matrix has the same support
  as the exact one, and the summed squared row error is below 1.0
  (cmp_trans_matrix, :79-123).
* bd_tvd_check: per-vertex Bhattacharyya and total-variation distances of
  the empirical first-order transitions, with the thresholds of
  src/test/release/test_dataset.cpp:174-193 (avg/median/p99 BD < 0.005 /
  0.005 / 0.015, TVD < 0.01 / 0.01 / 0.03).
* batch_zscores: per-vertex z-scores of a Monte-Carlo mean against an
  expected vector, with the standard error from batch means.
"""
import numpy as np


def exact_transitions(row_ptr, col):
    V = len(row_ptr) - 1
    P = np.zeros((V, V))
    for u in range(V):
        b, e = int(row_ptr[u]), int(row_ptr[u + 1])
        for k in range(b, e):
            P[u, col[k]] += 1.0 / (e - b)
    return P


def empirical_transitions(paths, V):
    a = paths[:, :-1].ravel()
    b = paths[:, 1:].ravel()
    ok = (a != 0xFFFFFFFF) & (b != 0xFFFFFFFF)
    a, b = a[ok].astype(np.int64), b[ok].astype(np.int64)
    C = np.zeros((V, V))
    np.add.at(C, (a, b), 1.0)
    return C


def transition_check(paths, row_ptr, col, bound=1.0):
    V = len(row_ptr) - 1
    P = exact_transitions(row_ptr, col)
    C = empirical_transitions(paths, V)
    illegal = (C > 0) & (P == 0)
    assert not illegal.any(), "a path step is not an edge"
    rows = C.sum(axis=1, keepdims=True)
    R = np.divide(C, rows, out=np.zeros_like(C), where=rows > 0)
    visited = rows[:, 0] > 0
    support_mismatch = ((R > 0) != (P > 0)) & visited[:, None]
    assert not support_mismatch.any(), "empirical transition support differs from the graph"
    err = ((R - P) ** 2)[visited].sum()
    assert err < bound, f"squared transition error {err} >= {bound}"
    return err


def bd_tvd_check(paths, row_ptr, col, min_samples=2000):
    V = len(row_ptr) - 1
    P = exact_transitions(row_ptr, col)
    C = empirical_transitions(paths, V)
    rows = C.sum(axis=1)
    bd, tvd = [], []
    for u in np.nonzero(rows >= min_samples)[0]:
        q = C[u] / rows[u]
        p = P[u]
        bc = np.sum(np.sqrt(p * q))
        bd.append(-np.log(max(bc, 1e-300)))
        tvd.append(0.5 * np.abs(p - q).sum())
    bd, tvd = np.array(bd), np.array(tvd)
    assert len(bd) > 0, "no vertex has enough samples"
    assert bd.mean() < 0.005 and np.median(bd) < 0.005 and np.percentile(bd, 99) < 0.015, bd
    assert tvd.mean() < 0.01 and np.median(tvd) < 0.01 and np.percentile(tvd, 99) < 0.03, tvd
    return bd, tvd


def batch_zscores(batches, expected):
    """batches: (B, V) per-batch per-walker means; expected: (V,)."""
    batches = np.asarray(batches, dtype=np.float64)
    B = batches.shape[0]
    mean = batches.mean(axis=0)
    se = batches.std(axis=0, ddof=1) / np.sqrt(B)
    z = np.zeros_like(mean)
    nz = se > 0
    z[nz] = (mean[nz] - expected[nz]) / se[nz]
    # where every batch agrees exactly, the mean must equal the expectation
    z[~nz] = np.where(np.abs(mean[~nz] - expected[~nz]) > 1e-9, np.inf, 0.0)
    return z, mean


def random_graph(V, p, seed, symmetric=True, dangling=0):
    """Small random graph as a CSR; `dangling` vertices get no out-edges."""
    rng = np.random.default_rng(seed)
    A = rng.random((V, V)) < p
    np.fill_diagonal(A, False)
    if symmetric:
        A = A | A.T
    # keep every non-dangling vertex connected to its successor (no other dangling)
    for u in range(V - dangling):
        A[u, (u + 1) % V] = True
        if symmetric:
            A[(u + 1) % V, u] = True
    for u in range(V - dangling, V):
        A[u, :] = False
    row_ptr = np.zeros(V + 1, dtype=np.uint32)
    cols = []
    for u in range(V):
        nb = np.nonzero(A[u])[0]
        rng.shuffle(nb)  # adjacency order is significant; make it non-sorted
        cols.extend(nb.tolist())
        row_ptr[u + 1] = len(cols)
    return row_ptr, np.array(cols, dtype=np.uint32)
