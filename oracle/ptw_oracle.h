/*
 * ptw_oracle.h -- CPU oracle for the probabilistic-termination walk engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libptw, the CLI
 * binaries, the Python package) links, loads or calls this code.  It is
 * used by tests/, by __graft_entry__.smoke() as the checker, and by
 * bench.py's cpu_baseline leg as the timed CPU path.
 *
 * Three things live here:
 *
 *  1. Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 *     as easy as 1, 2, 3", SC'11), written from the paper's definition.
 *
 *  2. A scalar replay of the engine's walk process on a CSR graph.  The
 *     process is the reference walker loop (src/core/solver.hpp:242-262 for
 *     visit counting, :284-305 for the endpoint callback; step budget
 *     L = 1 + K, K ~ Geometric(tp), src/core/pagerank.cpp:41-68; DeepWalk's
 *     fixed budget walk_len - 1, src/core/deepwalk.cpp:25-28) with every
 *     random decision taken from Philox keyed by (seed, walker id, step):
 *
 *       key  = (seed_lo, seed_hi)
 *       init = Philox(ctr = (wid_lo, wid_hi, 0xFFFFFFFF, stream))
 *              start = mulhi64(init.x | init.y << 32, V)     (uniform start)
 *       step i (i = 0, 1, ...):  r = Philox(ctr = (wid_lo, wid_hi, i, stream))
 *              stop if the budget is spent:
 *                 PT apps, reference semantics: i >= 1 and r.z >= thr
 *                 PT apps, textbook semantics:  r.z >= thr
 *                 fixed-length apps:            i == walk_len - 1
 *              stop if deg(v) == 0 (dangling; counted in *dangling)
 *              v = col[row_ptr[v] + mulhi64(r.x | r.y << 32, deg(v))]
 *       thr  = min(floor((1 - tp) * 2^32), 2^32 - 1)   (continue iff r.z < thr)
 *
 *     so that one Philox block decides one step, and the result does not
 *     depend on which lane, thread or order advanced the walker.
 *
 *  3. The reference's own CPU process with its own RNG: xorshift* with the
 *     output taken before the state update and a 16-bit float
 *     (include/random.hpp:144-180), the 128-entry geometric CDF table
 *     (src/core/pagerank.cpp:22-36,41-68), uniform neighbour choice by
 *     modulo (src/core/sampler.hpp:214-217).  OpenMP over walkers.
 *
 *  4. An FP64 power-iteration oracle for the expected visit vector, the
 *     expected endpoint vector, and the per-step position distributions.
 */
#ifndef PTW_ORACLE_H
#define PTW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* app kinds and semantics: identical numbering to include/ptw.h */
enum { ORC_APP_PAGERANK = 0, ORC_APP_PPR = 1, ORC_APP_RWR = 2, ORC_APP_DEEPWALK = 3, ORC_APP_NODE2VEC = 4 };
enum { ORC_SEM_REFERENCE = 0, ORC_SEM_TEXTBOOK = 1 };
enum { ORC_START_UNIFORM = 0, ORC_START_SOURCE = 1, ORC_START_ROUNDROBIN = 2 };

#define ORC_INVALID 0xFFFFFFFFu

/* Philox4x32-10 single block. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* floor(r64 * n / 2^64) for r64 = hi:lo */
uint32_t orc_mulhi64(uint32_t lo, uint32_t hi, uint32_t n);

/* continuation threshold for termination probability tp */
uint32_t orc_continue_threshold(double tp);

typedef struct {
    int app;            /* ORC_APP_* */
    int semantics;      /* ORC_SEM_* (PT apps only) */
    int start_mode;     /* ORC_START_* */
    uint32_t source;    /* start vertex when start_mode == SOURCE */
    double tp;          /* termination probability (PT apps) */
    uint32_t walk_len;  /* vertices per path (DEEPWALK) */
    uint64_t seed;
    uint32_t stream;    /* query / stream id (4th counter word) */
    uint64_t wid_base;  /* first walker id */
    uint64_t walkers;   /* number of walkers */
} orc_walk_desc;

typedef struct {
    uint64_t steps;     /* sum over walkers of transitions taken */
    uint64_t visits;    /* sum of count increments (count apps) */
    uint64_t dangling;  /* walkers stopped at a degree-0 vertex with budget left */
    uint64_t proposals;   /* node2vec: proposals drawn by the rejection sampler */
    uint64_t exact_draws; /* node2vec: steps decided by the exact draw */
} orc_walk_stats;

/* Philox replay.  counts: u64[V] accumulated (+=), may be NULL;
 * ends: u32[walkers] (RWR), may be NULL; paths: u32[walkers * walk_len]
 * (DEEPWALK), may be NULL.  threads <= 0 means all OpenMP threads. */
int orc_walk_replay(const uint32_t *row_ptr, const uint32_t *col, uint32_t V,
                    const orc_walk_desc *d, uint64_t *counts, uint32_t *ends,
                    uint32_t *paths, orc_walk_stats *st, int threads);

/* Same, with weighted neighbour choice through per-vertex alias tables
 * (thr, alias in CSR order): slot j = mulhi64(r.x | r.y << 32, deg), then
 * neighbour j if r.w < thr[j] else alias[j]. */
int orc_walk_replay_weighted(const uint32_t *row_ptr, const uint32_t *col, uint32_t V,
                             const uint32_t *alias_thr, const uint32_t *alias_idx,
                             const orc_walk_desc *d, uint64_t *counts, uint32_t *ends,
                             uint32_t *paths, orc_walk_stats *st, int threads);

/* Second-order (node2vec) walks of d->walk_len vertices, replayed with the
 * engine's decisions (node2vec app of include/ptw.h).  From cur, reached
 * from prev, neighbour x weighs 1/p (x == prev), 1 (x a neighbour of prev) or
 * 1/q; the integer weights are floor(2^32 w / max(1/p, 1, 1/q)).  Step i
 * (i >= 1): attempt a = 0..63 draws Philox(wid, i + a * 2^20), proposes
 * col[beg + mulhi64(r.x, r.y, deg)] and accepts iff r.z < its weight; then
 * Philox(wid, i + 64 * 2^20) picks exactly: floor((r.y:r.x) * total / 2^64)
 * against the running sum of the weights in adjacency order.  Step 0 is a
 * uniform pick.  Requires 1 <= walk_len <= 2^20 and p, q in [2^-16, 2^16].
 * The membership test uses a sorted copy of each adjacency. */
int orc_node2vec_replay(const uint32_t *row_ptr, const uint32_t *col, uint32_t V, const orc_walk_desc *d,
                        double p, double q, uint32_t *paths, orc_walk_stats *st, int threads);

/* Integer alias tables (Vose's method, exact integer masses) from float
 * weights in CSR order.  Per vertex: q_k = clamp(floor(w_k / w_max 2^24 +
 * 0.5), 1, 2^24), Q = sum q, mass m_k = q_k deg; slots with m < Q go on a
 * "small" stack, others on a "large" stack, both filled in index order and
 * popped from the top; pairing small s with large l sets
 * thr[s] = floor(m_s 2^32 / Q), alias[s] = l and m_l -= Q - m_s; slots
 * left over get thr = 2^32 - 1, alias = self.  Returns -1 on a weight that
 * is not finite and > 0. */
int orc_build_alias(const uint32_t *row_ptr, uint32_t V, const float *w, uint32_t *thr, uint32_t *alias);

/* Walk length of one walker under the Philox process (for closed-form checks). */
uint32_t orc_walk_length(uint64_t seed, uint32_t stream, uint64_t wid, double tp, int semantics);

/* Reference-semantics CPU path with xorshift* (own RNG).  Per-thread
 * generators are seeded from `seed` and the thread index.  counts u64[V]
 * (+=, atomic), ends u32[walkers] or NULL. */
int orc_ref_walk(const uint32_t *row_ptr, const uint32_t *col, uint32_t V,
                 int app, uint32_t source, int start_uniform, double tp, uint32_t walk_len,
                 uint64_t walkers, uint64_t seed, int threads,
                 uint64_t *counts, uint32_t *ends, orc_walk_stats *st);

/* The reference's own schedule on the CPU, with its own RNG: hub/low
 * partitions from a degree bound and the L2 size (l2_bytes 0: 1 MiB;
 * degree_bound 0: 32), per-thread per-partition walker pools, partitions
 * claimed in the reference's order under locks, migration, buffered hub
 * sampling (src/core/graph.hpp:400-628, src/core/solver.hpp:194-498,
 * 758-1028, src/core/sampler.hpp:43-167).  Apps PAGERANK/PPR (counts,
 * accumulated into u64[V]) and RWR (ends u32[walkers]).  walk_seconds: walker
 * init + walk (the reference's printed scope); partitions[2]: hub and low
 * partition counts.  walkers <= 2^32 - 1. */
int orc_ref_partitioned_walk(const uint32_t *row_ptr, const uint32_t *col, uint32_t V, int app, uint32_t source,
                             int start_uniform, double tp, uint64_t walkers, uint64_t seed, int threads,
                             uint64_t l2_bytes, uint32_t degree_bound, uint64_t *counts, uint32_t *ends,
                             orc_walk_stats *st, double *walk_seconds, uint32_t *partitions);

/* The reference's quantized geometric draw: 1 + K with the 128-entry CDF and
 * xorshift* gen_float.  Runs `n` draws from one generator seeded with `seed`
 * and writes a histogram of min(L, hist_len - 1) into hist (u64[hist_len]). */
void orc_ref_step_histogram(double tp, uint64_t seed, uint64_t n, uint64_t *hist, uint32_t hist_len);

/* xorshift* generator (include/random.hpp:161-176) exposed for tests:
 * writes n outputs of gen(bound) starting from `state`. */
void orc_xorshift_gen(uint64_t state, uint32_t bound, uint64_t n, uint32_t *out);
void orc_xorshift_gen_float(uint64_t state, float upper, uint64_t n, float *out);

/* FP64 oracle.  mode 0: expected visits per walker (count apps), mode 1:
 * expected endpoint distribution (RWR).  start: source vertex, or
 * ORC_INVALID for the uniform start.  out: double[V].  Iterates until the
 * remaining mass is below tol or max_iter is reached; returns iterations. */
int orc_power_iteration(const uint32_t *row_ptr, const uint32_t *col, uint32_t V,
                        uint32_t start, double tp, int semantics, int mode,
                        double tol, int max_iter, double *out);

#ifdef __cplusplus
}
#endif

#endif
