/*
 * oracle.c -- CPU oracle for the walk engine.  TEST INFRASTRUCTURE ONLY:
 * see ptw_oracle.h for what is here and who may call it.
 *
 * Reference files followed (PTWalker):
 *   walker loop, visit counting:      src/core/solver.hpp:242-262
 *   walker loop, endpoint callback:   src/core/solver.hpp:284-305, src/core/RWR.cpp:99-105
 *   start vertex per walker:          src/core/solver.hpp:767-776, src/core/pagerank.cpp:70-80
 *   step budget 1 + Geometric(tp):    src/core/pagerank.cpp:22-36,41-68
 *   DeepWalk fixed budget:            src/core/deepwalk.cpp:25-28
 *   uniform neighbour choice:         src/core/sampler.hpp:214-217
 *   xorshift* generator:              include/random.hpp:144-180
 */
#include "ptw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11, section 4.2): S-box round on   */
/* (x0,x1,x2,x3) with multipliers M0, M1 and a Weyl key schedule.       */
/* ------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * x0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

uint32_t orc_mulhi64(uint32_t lo, uint32_t hi, uint32_t n)
{
    /* floor((hi * 2^32 + lo) * n / 2^64), exact in 64-bit pieces */
    uint64_t a = (uint64_t)hi * n;
    uint64_t b = ((uint64_t)lo * n) >> 32;
    return (uint32_t)((a + b) >> 32);
}

uint32_t orc_continue_threshold(double tp)
{
    double c = floor((1.0 - tp) * 4294967296.0);
    if (c < 0.0) c = 0.0;
    if (c > 4294967295.0) c = 4294967295.0;
    return (uint32_t)c;
}

static inline void draw(uint64_t seed, uint32_t stream, uint64_t wid, uint32_t step, uint32_t r[4])
{
    uint32_t ctr[4] = {(uint32_t)wid, (uint32_t)(wid >> 32), step, stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, r);
}

uint32_t orc_walk_length(uint64_t seed, uint32_t stream, uint64_t wid, double tp, int semantics)
{
    uint32_t thr = orc_continue_threshold(tp);
    uint32_t r[4];
    for (uint32_t i = 0;; ++i) {
        draw(seed, stream, wid, i, r);
        int coin = (semantics == ORC_SEM_TEXTBOOK) || i >= 1;
        if (coin && r[2] >= thr) return i;
    }
}

static int is_pt_app(int app) { return app != ORC_APP_DEEPWALK; }

int orc_walk_replay(const uint32_t *row_ptr, const uint32_t *col, uint32_t V,
                    const orc_walk_desc *d, uint64_t *counts, uint32_t *ends,
                    uint32_t *paths, orc_walk_stats *st, int threads)
{
    return orc_walk_replay_weighted(row_ptr, col, V, NULL, NULL, d, counts, ends, paths, st, threads);
}

int orc_walk_replay_weighted(const uint32_t *row_ptr, const uint32_t *col, uint32_t V,
                             const uint32_t *alias_thr, const uint32_t *alias_idx,
                             const orc_walk_desc *d, uint64_t *counts, uint32_t *ends,
                             uint32_t *paths, orc_walk_stats *st, int threads)
{
    if (V == 0 || d->walkers == 0) {
        if (st) memset(st, 0, sizeof(*st));
        return 0;
    }
    if (d->start_mode == ORC_START_SOURCE && d->source >= V) return -1;
    if (d->app == ORC_APP_DEEPWALK && d->walk_len == 0) return -1;
    const uint32_t thr = orc_continue_threshold(d->tp);
    const int pt = is_pt_app(d->app);
    const int count_app = (d->app == ORC_APP_PAGERANK || d->app == ORC_APP_PPR);
    uint64_t tot_steps = 0, tot_visits = 0, tot_dangling = 0;
    const int64_t W = (int64_t)d->walkers;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nt) reduction(+ : tot_steps, tot_visits, tot_dangling)
#else
    (void)threads;
#endif
    for (int64_t w = 0; w < W; ++w) {
        uint64_t wid = d->wid_base + (uint64_t)w;
        uint32_t r[4];
        uint32_t v;
        if (d->start_mode == ORC_START_SOURCE) {
            v = d->source;
        } else if (d->start_mode == ORC_START_ROUNDROBIN) {
            v = (uint32_t)(wid % V);
        } else {
            draw(d->seed, d->stream, wid, 0xFFFFFFFFu, r);
            v = orc_mulhi64(r[0], r[1], V);
        }
        if (count_app && counts) {
            __atomic_fetch_add(&counts[v], 1, __ATOMIC_RELAXED);
        }
        tot_visits += count_app;
        uint32_t *path = paths ? paths + (uint64_t)w * d->walk_len : NULL;
        if (path) path[0] = v;
        uint32_t i = 0;
        for (;; ++i) {
            if (!pt && i + 1 >= d->walk_len) break;
            draw(d->seed, d->stream, wid, i, r);
            if (pt) {
                int coin = (d->semantics == ORC_SEM_TEXTBOOK) || i >= 1;
                if (coin && r[2] >= thr) break;
            }
            uint32_t beg = row_ptr[v], deg = row_ptr[v + 1] - beg;
            if (deg == 0) { tot_dangling++; break; }
            uint32_t j = orc_mulhi64(r[0], r[1], deg);
            if (alias_thr && r[3] >= alias_thr[beg + j]) j = alias_idx[beg + j];
            v = col[beg + j];
            if (count_app && counts) __atomic_fetch_add(&counts[v], 1, __ATOMIC_RELAXED);
            tot_visits += count_app;
            if (path) path[i + 1] = v;
        }
        tot_steps += i;
        if (path) {
            for (uint32_t k = i + 1; k < d->walk_len; ++k) path[k] = ORC_INVALID;
        }
        if (ends) ends[w] = v;
    }
    if (st) {
        st->steps = tot_steps;
        st->visits = tot_visits;
        st->dangling = tot_dangling;
    }
    return 0;
}

static int cmp_u32(const void *a, const void *b)
{
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

static int n2v_has(const uint32_t *a, uint32_t n, uint32_t x)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == x;
}

static uint64_t n2v_thr(double w, double M)
{
    double t = floor(w / M * 4294967296.0);
    return t >= 4294967296.0 ? 4294967296ull : (uint64_t)t;
}

typedef struct {
    const uint32_t *pn;
    uint32_t pdeg, prev;
    uint64_t w_ret, w_in, w_out;
} n2v_ctx;

static uint64_t n2v_w(const n2v_ctx *c, uint32_t x)
{
    if (x == c->prev) return c->w_ret;
    return n2v_has(c->pn, c->pdeg, x) ? c->w_in : c->w_out;
}

int orc_node2vec_replay(const uint32_t *row_ptr, const uint32_t *col, uint32_t V, const orc_walk_desc *d,
                        double p, double q, uint32_t *paths, orc_walk_stats *st, int threads)
{
    if (st) memset(st, 0, sizeof(*st));
    if (V == 0 || d->walkers == 0) return 0;
    if (d->walk_len == 0 || d->walk_len > (1u << 20)) return -1;
    if (!(p >= 1.0 / 65536 && p <= 65536) || !(q >= 1.0 / 65536 && q <= 65536)) return -1;
    if (d->start_mode == ORC_START_SOURCE && d->source >= V) return -1;
    const uint64_t E = row_ptr[V];
    uint32_t *scol = (uint32_t *)malloc((E ? E : 1) * sizeof(uint32_t));
    if (!scol) return -1;
    memcpy(scol, col, E * sizeof(uint32_t));
    uint32_t maxdeg = 0;
    /* simple (no row repeats a neighbour) and symmetric (u in x's row for every
     * edge u -> x): then the previous vertex is in the current row exactly once */
    int64_t bad = 0;
#ifdef _OPENMP
    const int nt0 = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nt0) reduction(max : maxdeg)
#endif
    for (int64_t v = 0; v < (int64_t)V; ++v) {
        qsort(scol + row_ptr[v], row_ptr[v + 1] - row_ptr[v], 4, cmp_u32);
        if (row_ptr[v + 1] - row_ptr[v] > maxdeg) maxdeg = row_ptr[v + 1] - row_ptr[v];
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nt0) reduction(+ : bad)
#endif
    for (int64_t u = 0; u < (int64_t)V; ++u)
        for (uint32_t k = row_ptr[u]; k < row_ptr[u + 1]; ++k) {
            const uint32_t x = scol[k];
            if ((k > row_ptr[u] && x == scol[k - 1]) ||
                !n2v_has(scol + row_ptr[x], row_ptr[x + 1] - row_ptr[x], (uint32_t)u)) {
                ++bad;
                break;
            }
        }
    const int simple_sym = bad == 0;
    /* Weights floor(2^32 w / M).  Unfolded: M = max(1/p, 1, 1/q), one proposal
     * slot per neighbour, each weight an acceptance threshold <= 2^32.
     * Folded (1/p the largest weight on a simple symmetric graph): M =
     * max(1, 1/q); prev's own slot accepts always and its excess weight over
     * 2^32 is spread over `fold` extra slots that accept prev below w_fold
     * (outlier folding, Yang et al., KnightKing, SOSP 2019). */
    const double wr = 1.0 / p, wo = 1.0 / q;
    const double Mf = wo > 1.0 ? wo : 1.0;
    const uint64_t t_ret = (uint64_t)floor(wr / Mf * 4294967296.0);
    const int folded = simple_sym && wr > Mf && t_ret > 4294967296ull && maxdeg < (1u << 31);
    const double M = folded ? Mf : (wr > Mf ? wr : Mf);
    const uint64_t w_in = n2v_thr(1.0, M), w_out = n2v_thr(wo, M);
    uint64_t w_ret, w_fold = 0;
    uint32_t fold = 0;
    if (folded) {
        const uint64_t excess = t_ret - 4294967296ull;
        fold = (uint32_t)((excess + 4294967295ull) >> 32);
        w_fold = excess / fold;
        w_ret = 4294967296ull + (uint64_t)fold * w_fold;
    } else {
        w_ret = n2v_thr(wr, M);
    }
    uint64_t tot_steps = 0, tot_dangling = 0, tot_prop = 0, tot_exact = 0;
    const int64_t W = (int64_t)d->walkers;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt) \
    reduction(+ : tot_steps, tot_dangling, tot_prop, tot_exact)
#else
    (void)threads;
#endif
    for (int64_t w = 0; w < W; ++w) {
        const uint64_t wid = d->wid_base + (uint64_t)w;
        uint32_t r[4];
        uint32_t v;
        if (d->start_mode == ORC_START_SOURCE) {
            v = d->source;
        } else if (d->start_mode == ORC_START_ROUNDROBIN) {
            v = (uint32_t)(wid % V);
        } else {
            draw(d->seed, d->stream, wid, 0xFFFFFFFFu, r);
            v = orc_mulhi64(r[0], r[1], V);
        }
        uint32_t *path = paths ? paths + (uint64_t)w * d->walk_len : NULL;
        if (path) path[0] = v;
        uint32_t prev = ORC_INVALID, i = 0;
        for (; i + 1 < d->walk_len; ++i) {
            const uint32_t beg = row_ptr[v], deg = row_ptr[v + 1] - beg;
            if (deg == 0) { tot_dangling++; break; }
            draw(d->seed, d->stream, wid, i, r);
            uint32_t x = 0;
            if (i == 0) {
                x = col[beg + orc_mulhi64(r[0], r[1], deg)];
            } else {
                n2v_ctx c = {scol + row_ptr[prev], row_ptr[prev + 1] - row_ptr[prev], prev, w_ret, w_in, w_out};
                int found = 0;
                for (uint32_t a = 0; a < 64 && !found; ++a) {
                    if (a) draw(d->seed, d->stream, wid, i + (a << 20), r);
                    ++tot_prop;
                    const uint32_t s = orc_mulhi64(r[0], r[1], deg + fold);
                    if (s >= deg) {  /* an extra slot of the return edge */
                        if ((uint64_t)r[2] < w_fold) { x = prev; found = 1; }
                        continue;
                    }
                    const uint32_t y = col[beg + s];
                    const uint64_t wy = (fold && y == prev) ? 4294967296ull : n2v_w(&c, y);
                    if ((uint64_t)r[2] < wy) { x = y; found = 1; }
                }
                if (!found) {
                    ++tot_exact;
                    uint64_t total = 0;
                    for (uint32_t k = 0; k < deg; ++k) total += n2v_w(&c, col[beg + k]);
                    draw(d->seed, d->stream, wid, i + (64u << 20), r);
                    const uint64_t u = ((uint64_t)r[1] << 32) | r[0];
                    const uint64_t pick = (uint64_t)(((unsigned __int128)u * total) >> 64);
                    uint64_t cum = 0;
                    x = col[beg + deg - 1];
                    for (uint32_t k = 0; k < deg; ++k) {
                        cum += n2v_w(&c, col[beg + k]);
                        if (pick < cum) { x = col[beg + k]; break; }
                    }
                }
            }
            prev = v;
            v = x;
            if (path) path[i + 1] = v;
        }
        tot_steps += i;
        if (path) for (uint32_t k = i + 1; k < d->walk_len; ++k) path[k] = ORC_INVALID;
    }
    free(scol);
    if (st) {
        st->steps = tot_steps;
        st->dangling = tot_dangling;
        st->proposals = tot_prop;
        st->exact_draws = tot_exact;
    }
    return 0;
}

int orc_build_alias(const uint32_t *row_ptr, uint32_t V, const float *w, uint32_t *thr, uint32_t *alias)
{
    uint64_t E = row_ptr[V];
    for (uint64_t e = 0; e < E; ++e)
        if (!(w[e] > 0.0f) || !isfinite(w[e])) return -1;
    uint32_t maxdeg = 0;
    for (uint32_t v = 0; v < V; ++v)
        if (row_ptr[v + 1] - row_ptr[v] > maxdeg) maxdeg = row_ptr[v + 1] - row_ptr[v];
    uint64_t *mass = (uint64_t *)malloc(sizeof(uint64_t) * (maxdeg + 1));
    uint32_t *sm = (uint32_t *)malloc(sizeof(uint32_t) * (maxdeg + 1));
    uint32_t *lg = (uint32_t *)malloc(sizeof(uint32_t) * (maxdeg + 1));
    for (uint32_t v = 0; v < V; ++v) {
        uint32_t b = row_ptr[v], n = row_ptr[v + 1] - b;
        if (n == 0) continue;
        double wmax = 0.0;
        for (uint32_t k = 0; k < n; ++k)
            if ((double)w[b + k] > wmax) wmax = (double)w[b + k];
        uint64_t Q = 0;
        for (uint32_t k = 0; k < n; ++k) {
            double qd = floor((double)w[b + k] / wmax * 16777216.0 + 0.5);
            uint64_t q = qd < 1.0 ? 1 : (qd > 16777216.0 ? 16777216 : (uint64_t)qd);
            mass[k] = q;
            Q += q;
        }
        uint32_t ns = 0, nl = 0;
        for (uint32_t k = 0; k < n; ++k) {
            mass[k] *= n;
            if (mass[k] < Q) sm[ns++] = k; else lg[nl++] = k;
        }
        while (ns > 0 && nl > 0) {
            uint32_t s = sm[--ns];
            uint32_t l = lg[--nl];
            thr[b + s] = (uint32_t)(((unsigned __int128)mass[s] << 32) / Q);
            alias[b + s] = l;
            mass[l] -= Q - mass[s];
            if (mass[l] < Q) sm[ns++] = l; else lg[nl++] = l;
        }
        for (uint32_t i = 0; i < nl; ++i) { thr[b + lg[i]] = 0xFFFFFFFFu; alias[b + lg[i]] = lg[i]; }
        for (uint32_t i = 0; i < ns; ++i) { thr[b + sm[i]] = 0xFFFFFFFFu; alias[b + sm[i]] = sm[i]; }
    }
    free(mass);
    free(sm);
    free(lg);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Reference process with the reference's own RNG                       */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t s; char pad[56]; } xorshift_t;

/* include/random.hpp:161-168: output from the pre-update state */
static inline uint32_t xs_gen(xorshift_t *g, uint32_t bound)
{
    uint32_t ret = (uint32_t)(g->s * UINT64_C(0x2545F4914F6CDD1D));
    g->s ^= g->s >> 12;
    g->s ^= g->s << 25;
    g->s ^= g->s >> 27;
    return ret % bound;
}

/* include/random.hpp:169-176: low 16 bits / 65535 (can be exactly 1.0) */
static inline float xs_gen_float(xorshift_t *g, float upper)
{
    uint32_t ret = (uint32_t)((g->s * UINT64_C(0x2545F4914F6CDD1D)) & 0xFFFF);
    g->s ^= g->s >> 12;
    g->s ^= g->s << 25;
    g->s ^= g->s >> 27;
    return (float)ret / (float)65535 * upper;
}

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += UINT64_C(0x9E3779B97F4A7C15));
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

static void xs_seed(xorshift_t *g, uint64_t seed, uint64_t index)
{
    uint64_t x = seed ^ (index * UINT64_C(0xD1B54A32D192ED03));
    g->s = splitmix64(&x);
    if (g->s == 0) g->s = 1; /* 0 is a fixed point of xorshift */
}

typedef struct { double cdf[128]; double tot; } geo_table;

/* src/core/pagerank.cpp:22-36 */
static void geo_init(geo_table *t, double tp)
{
    for (int si = 0; si < 128; ++si) t->cdf[si] = pow(1.0 - tp, si) * tp;
    for (int si = 0; si < 127; ++si) t->cdf[si + 1] += t->cdf[si];
    t->tot = t->cdf[127];
}

/* src/core/pagerank.cpp:41-68 */
static inline uint32_t geo_draw(const geo_table *t, xorshift_t *g)
{
    uint32_t stepid = 1;
    for (;;) {
        double u = xs_gen_float(g, 1.0f);
        if (u > t->tot) {
            stepid += 128;
            continue;
        }
        for (int si = 0; si < 128; ++si) {
            if (t->cdf[si] >= u) {
                stepid += (uint32_t)si;
                break;
            }
        }
        return stepid;
    }
}

void orc_ref_step_histogram(double tp, uint64_t seed, uint64_t n, uint64_t *hist, uint32_t hist_len)
{
    geo_table t;
    geo_init(&t, tp);
    xorshift_t g;
    xs_seed(&g, seed, 0);
    memset(hist, 0, sizeof(uint64_t) * hist_len);
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t L = geo_draw(&t, &g);
        if (L >= hist_len) L = hist_len - 1;
        hist[L]++;
    }
}

void orc_xorshift_gen(uint64_t state, uint32_t bound, uint64_t n, uint32_t *out)
{
    xorshift_t g;
    g.s = state;
    for (uint64_t i = 0; i < n; ++i) out[i] = xs_gen(&g, bound);
}

void orc_xorshift_gen_float(uint64_t state, float upper, uint64_t n, float *out)
{
    xorshift_t g;
    g.s = state;
    for (uint64_t i = 0; i < n; ++i) out[i] = xs_gen_float(&g, upper);
}

int orc_ref_walk(const uint32_t *row_ptr, const uint32_t *col, uint32_t V,
                 int app, uint32_t source, int start_uniform, double tp, uint32_t walk_len,
                 uint64_t walkers, uint64_t seed, int threads,
                 uint64_t *counts, uint32_t *ends, orc_walk_stats *st)
{
    if (V == 0) return -1;
    if (!start_uniform && source >= V) return -1;
    geo_table t;
    geo_init(&t, tp);
    const int count_app = (app == ORC_APP_PAGERANK || app == ORC_APP_PPR);
    uint64_t tot_steps = 0, tot_visits = 0, tot_dangling = 0;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    int nt = 1;
    (void)threads;
#endif
    xorshift_t *gens = (xorshift_t *)aligned_alloc(64, sizeof(xorshift_t) * (size_t)nt);
    for (int k = 0; k < nt; ++k) xs_seed(&gens[k], seed, (uint64_t)k);
    const int64_t W = (int64_t)walkers;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt) reduction(+ : tot_steps, tot_visits, tot_dangling)
#endif
    {
#ifdef _OPENMP
        xorshift_t *g = &gens[omp_get_thread_num()];
#pragma omp for schedule(static)
#else
        xorshift_t *g = &gens[0];
#endif
        for (int64_t w = 0; w < W; ++w) {
            /* walker_init then walking_step, src/core/solver.hpp:767-781 */
            uint32_t v = start_uniform ? xs_gen(g, V) : source;
            uint32_t L = (app == ORC_APP_DEEPWALK) ? walk_len - 1 : geo_draw(&t, g);
            if (count_app) {
                __atomic_fetch_add(&counts[v], 1, __ATOMIC_RELAXED);
                tot_visits++;
            }
            uint32_t i = 0;
            for (; i < L; ++i) {
                uint32_t beg = row_ptr[v], deg = row_ptr[v + 1] - beg;
                if (deg == 0) { tot_dangling++; break; }
                v = col[beg + xs_gen(g, deg)];
                if (count_app) {
                    __atomic_fetch_add(&counts[v], 1, __ATOMIC_RELAXED);
                    tot_visits++;
                }
            }
            tot_steps += i;
            if (ends) ends[w] = v;
        }
    }
    free(gens);
    if (st) {
        st->steps = tot_steps;
        st->visits = tot_visits;
        st->dangling = tot_dangling;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* FP64 power iteration                                                 */
/* ------------------------------------------------------------------ */
/* y <- y P (row-stochastic uniform transitions; dangling rows lose mass) */
static void push(const uint32_t *row_ptr, const uint32_t *col, uint32_t V,
                 const double *y, double *out)
{
    memset(out, 0, sizeof(double) * V);
    for (uint32_t u = 0; u < V; ++u) {
        uint32_t beg = row_ptr[u], deg = row_ptr[u + 1] - beg;
        if (deg == 0 || y[u] == 0.0) continue;
        double m = y[u] / (double)deg;
        for (uint32_t k = 0; k < deg; ++k) out[col[beg + k]] += m;
    }
}

int orc_power_iteration(const uint32_t *row_ptr, const uint32_t *col, uint32_t V,
                        uint32_t start, double tp, int semantics, int mode,
                        double tol, int max_iter, double *out)
{
    double *y = (double *)calloc(V, sizeof(double));
    double *z = (double *)calloc(V, sizeof(double));
    if (!y || !z) { free(y); free(z); return -1; }
    if (start == ORC_INVALID) {
        for (uint32_t v = 0; v < V; ++v) y[v] = 1.0 / (double)V;
    } else {
        y[start] = 1.0;
    }
    memset(out, 0, sizeof(double) * V);
    int it = 0;
    if (mode == 0) {
        /* E[count] per walker = e_s + sum_{i>=1} P(L >= i) e_s P^i, with
         * P(L >= i) = (1-tp)^(i-1) (reference) or (1-tp)^i (textbook). */
        for (uint32_t v = 0; v < V; ++v) out[v] = y[v];
        double w = (semantics == ORC_SEM_TEXTBOOK) ? (1.0 - tp) : 1.0;
        for (it = 1; it <= max_iter; ++it) {
            push(row_ptr, col, V, y, z);
            double mass = 0.0;
            for (uint32_t v = 0; v < V; ++v) {
                out[v] += w * z[v];
                mass += z[v];
            }
            double *tmp = y; y = z; z = tmp;
            if (w * mass < tol) break;
            w *= (1.0 - tp);
        }
    } else {
        /* E[end] = sum_{L} P(L) e_s P^L; reference: P(L = k+1) = tp (1-tp)^k */
        double w = (semantics == ORC_SEM_TEXTBOOK) ? tp : 0.0;
        if (semantics == ORC_SEM_TEXTBOOK) {
            for (uint32_t v = 0; v < V; ++v) out[v] = w * y[v];
        }
        double pl = (semantics == ORC_SEM_TEXTBOOK) ? tp * (1.0 - tp) : tp;
        for (it = 1; it <= max_iter; ++it) {
            push(row_ptr, col, V, y, z);
            double mass = 0.0;
            for (uint32_t v = 0; v < V; ++v) {
                out[v] += pl * z[v];
                mass += z[v];
            }
            double *tmp = y; y = z; z = tmp;
            /* remaining weight sum_{j>it} P(L=j) = P(L > it) */
            if (mass * (pl / tp) * (1.0 - tp) < tol) break;
            pl *= (1.0 - tp);
        }
    }
    free(y);
    free(z);
    return it;
}

/* ------------------------------------------------------------------ */
/* The reference's own schedule: partitions, walker pools, migration,  */
/* buffered hub sampling                                               */
/* ------------------------------------------------------------------ */
/* Graph side (src/core/graph.hpp:400-628): vertices of degree >= bound are
 * hubs; they are shuffled and cut into hub partitions of about
 * L2 * 0.9 / 80 vertices (an even number of partitions, half of them a
 * multiple of the thread count); the other vertices are shuffled and cut in
 * order into low partitions of at most L2 * 0.9 bytes (8 B per vertex + 4 B
 * per edge, dp_partition).  The relabel puts the hubs first, then the low
 * partitions, each padded to the largest one's vertex count, so a vertex's
 * partition is arithmetic (vertex_partition_id_fuc, :159-169).  The
 * reference first splits the low vertices into two max-cut groups; this
 * restatement keeps one group and takes the first half of the low
 * partitions as its "left" ones.
 * Solver side (src/core/solver.hpp:758-1028, 194-498): every thread keeps a
 * pool of 12 B walker records per partition; an iteration lets the threads
 * claim partitions in the reference's map order, each under its lock, and
 * advance every walker found in any thread's pool for it until the walker
 * leaves the partition (it moves to the claiming thread's pool of its new
 * partition) or ends.
 * Sampling (src/core/sampler.hpp:43-167, 214-217): hub partitions draw
 * from per-vertex buffers of pre-drawn neighbours, max(8, W L deg / E)
 * entries (+1 if a power of two above 8; W = walkers still walking, L =
 * floor(1 / tp) + 1), refilled when empty; low partitions pick directly.
 * Counts (type 0) are plain increments: a partition's vertices are only
 * touched by the thread that holds its lock. */
typedef struct { uint32_t wid, vid, sid; } pw_walker;
typedef struct { pw_walker *a; uint64_t n, cap; omp_lock_t lock; } pw_pool;
typedef struct { uint32_t head, end; } pw_buf;

typedef struct {
    uint32_t H, vph, np_hub, hub_left, vpd, nlow, np;
    uint64_t *rp;
    uint32_t *cl, *newid, *oldid, *pend;
} pw_graph;

static inline uint32_t pw_part(const pw_graph *g, uint32_t v)
{
    return v < g->H ? v / g->vph : g->np_hub + (v - g->H) / g->vpd;
}

static void pw_shuffle(uint32_t *a, uint32_t n, xorshift_t *g)
{
    for (uint32_t i = n; i > 1; --i) {
        uint32_t j = xs_gen(g, i);
        uint32_t t = a[i - 1];
        a[i - 1] = a[j];
        a[j] = t;
    }
}

static void pw_push(pw_pool *p, pw_walker w)
{
    if (p->n == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 4196; /* the reference's container size */
        p->a = (pw_walker *)realloc(p->a, p->cap * sizeof(pw_walker));
    }
    p->a[p->n++] = w;
}

static uint32_t pw_buf_len(uint32_t deg, uint64_t walkers, uint32_t walk_len, uint64_t E)
{
    uint64_t len = E ? walkers * (uint64_t)walk_len * deg / E : 0;
    if (len < 8) len = 8;
    if (len > 8 && (len & (len - 1)) == 0) len++;
    return (uint32_t)len;
}

int orc_ref_partitioned_walk(const uint32_t *row_ptr, const uint32_t *col, uint32_t V, int app, uint32_t source,
                             int start_uniform, double tp, uint64_t walkers, uint64_t seed, int threads,
                             uint64_t l2_bytes, uint32_t degree_bound, uint64_t *counts, uint32_t *ends,
                             orc_walk_stats *st, double *walk_seconds, uint32_t *partitions)
{
    if (V == 0 || (app != ORC_APP_PAGERANK && app != ORC_APP_PPR && app != ORC_APP_RWR)) return -1;
    if (!start_uniform && source >= V) return -1;
    if (walkers > 0xFFFFFFFFull || !(tp > 0.0 && tp <= 1.0)) return -1;
    const uint64_t E = row_ptr[V];
    int nt = threads > 0 ? threads : omp_get_max_threads();
    if (l2_bytes == 0) l2_bytes = 1u << 20;
    if (degree_bound == 0) degree_bound = 32;
    xorshift_t gg;
    xs_seed(&gg, seed, 0x9A47u);
    /* ---- partitions and relabel ---- */
    pw_graph G;
    memset(&G, 0, sizeof(G));
    uint32_t *hubs = (uint32_t *)malloc(sizeof(uint32_t) * V), *low = (uint32_t *)malloc(sizeof(uint32_t) * V);
    uint32_t nh = 0, nl = 0;
    for (uint32_t v = 0; v < V; ++v) {
        if (row_ptr[v + 1] - row_ptr[v] >= degree_bound) hubs[nh++] = v;
        else low[nl++] = v;
    }
    pw_shuffle(hubs, nh, &gg);
    pw_shuffle(low, nl, &gg);
    G.H = nh;
    if (nh) {
        uint64_t vph = l2_bytes * 9 / 10 / (8 + 8 + 64) - 1;
        if (vph == 0) vph = 1;
        uint32_t p = (uint32_t)(nh / vph + 1);
        if (p & 1) ++p;
        uint32_t left = p / 2;
        if (left % (uint32_t)nt) left = (left / (uint32_t)nt + 1) * (uint32_t)nt;
        G.np_hub = 2 * left;
        G.hub_left = left;
        G.vph = nh / G.np_hub + 1;
    }
    /* low partitions: consecutive runs of the shuffled order, at most subgraph bytes */
    const uint64_t subgraph = l2_bytes * 9 / 10;
    uint32_t *lcut = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)nl + 2));
    uint32_t nlow = 0;
    lcut[0] = 0;
    {
        uint64_t cur = 0;
        for (uint32_t i = 0; i < nl; ++i) {
            cur += 8 + 4ull * (row_ptr[low[i] + 1] - row_ptr[low[i]]);
            if (cur >= subgraph) {
                lcut[++nlow] = i + 1;
                cur = 0;
            }
        }
        if (lcut[nlow] != nl) lcut[++nlow] = nl;
    }
    G.nlow = nlow;
    G.vpd = 1;
    for (uint32_t c = 0; c < nlow; ++c)
        if (lcut[c + 1] - lcut[c] > G.vpd) G.vpd = lcut[c + 1] - lcut[c];
    G.np = G.np_hub + nlow;
    const uint64_t Vp = (uint64_t)nh + (uint64_t)nlow * G.vpd;
    G.newid = (uint32_t *)malloc(sizeof(uint32_t) * V);
    G.oldid = (uint32_t *)malloc(sizeof(uint32_t) * Vp);
    G.pend = (uint32_t *)malloc(sizeof(uint32_t) * (G.np ? G.np : 1));
    for (uint64_t u = 0; u < Vp; ++u) G.oldid[u] = ORC_INVALID;
    for (uint32_t i = 0; i < nh; ++i) {
        G.newid[hubs[i]] = i;
        G.oldid[i] = hubs[i];
    }
    for (uint32_t p = 0; p < G.np_hub; ++p) {
        uint64_t e = (uint64_t)(p + 1) * G.vph;
        G.pend[p] = (uint32_t)(e < nh ? e : nh);
    }
    for (uint32_t c = 0; c < nlow; ++c) {
        const uint32_t base = nh + c * G.vpd;
        for (uint32_t i = lcut[c]; i < lcut[c + 1]; ++i) {
            G.newid[low[i]] = base + (i - lcut[c]);
            G.oldid[base + (i - lcut[c])] = low[i];
        }
        G.pend[G.np_hub + c] = base + (lcut[c + 1] - lcut[c]);
    }
    free(hubs);
    free(low);
    free(lcut);
    G.rp = (uint64_t *)malloc(sizeof(uint64_t) * (Vp + 1));
    G.cl = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
    G.rp[0] = 0;
    for (uint64_t u = 0; u < Vp; ++u) {
        const uint32_t o = G.oldid[u];
        G.rp[u + 1] = G.rp[u] + (o == ORC_INVALID ? 0 : row_ptr[o + 1] - row_ptr[o]);
    }
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4096)
    for (int64_t u = 0; u < (int64_t)Vp; ++u) {
        const uint32_t o = G.oldid[u];
        if (o == ORC_INVALID) continue;
        for (uint32_t k = row_ptr[o]; k < row_ptr[o + 1]; ++k) G.cl[G.rp[u] + (k - row_ptr[o])] = G.newid[col[k]];
    }
    /* ---- hub edge buffers ---- */
    const uint32_t walk_len = (uint32_t)(1.0 / tp) + 1;
    pw_buf *hb = (pw_buf *)malloc(sizeof(pw_buf) * (nh ? nh : 1));
    uint64_t units_n = 0;
    for (uint32_t h = 0; h < nh; ++h) {
        units_n += pw_buf_len((uint32_t)(G.rp[h + 1] - G.rp[h]), walkers, walk_len, E);
        hb[h].end = (uint32_t)units_n;
        hb[h].head = hb[h].end;
    }
    uint32_t *units = (uint32_t *)malloc(sizeof(uint32_t) * (units_n ? units_n : 1));
    /* ---- pools, locks, map ---- */
    pw_pool *pools = (pw_pool *)calloc((size_t)nt * G.np, sizeof(pw_pool));
    for (size_t i = 0; i < (size_t)nt * G.np; ++i) omp_init_lock(&pools[i].lock);
    omp_lock_t *plock = (omp_lock_t *)malloc(sizeof(omp_lock_t) * G.np);
    for (uint32_t p = 0; p < G.np; ++p) omp_init_lock(&plock[p]);
    const uint32_t low_left = nlow / 2;
    uint32_t nmap = 0;
    uint32_t *map = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)G.np + G.np_hub + 1));
    for (uint32_t p = 0; p < G.hub_left; ++p) map[nmap++] = p;
    for (uint32_t p = 0; p < low_left; ++p) map[nmap++] = G.np_hub + p;
    for (uint32_t p = 0; p < G.hub_left; ++p) map[nmap++] = p;
    for (uint32_t p = G.hub_left; p < G.np_hub; ++p) map[nmap++] = p;
    for (uint32_t p = low_left; p < nlow; ++p) map[nmap++] = G.np_hub + p;
    for (uint32_t p = G.hub_left; p < G.np_hub; ++p) map[nmap++] = p;
    uint64_t *cnt = NULL;
    if (app != ORC_APP_RWR) cnt = (uint64_t *)calloc(Vp ? Vp : 1, sizeof(uint64_t));
    geo_table t;
    geo_init(&t, tp);
    xorshift_t *gens = (xorshift_t *)aligned_alloc(64, sizeof(xorshift_t) * (size_t)nt);
    for (int k = 0; k < nt; ++k) xs_seed(&gens[k], seed, (uint64_t)k + 1);
    const int type0 = app != ORC_APP_RWR;
    uint64_t tot_steps = 0, tot_visits = 0, tot_dangling = 0;
    const double t0 = omp_get_wtime();
    /* ---- init_start_walks ---- */
#pragma omp parallel num_threads(nt)
    {
        const int tid = omp_get_thread_num();
        xorshift_t *g = &gens[tid];
#pragma omp for schedule(static)
        for (int64_t w = 0; w < (int64_t)walkers; ++w) {
            const uint32_t v = G.newid[start_uniform ? xs_gen(g, V) : source];
            pw_walker r = {(uint32_t)w, v, geo_draw(&t, g)};
            pw_push(&pools[(size_t)tid * G.np + pw_part(&G, v)], r);
        }
    }
    /* ---- iterations ---- */
    uint64_t finished = 0;
    while (finished < walkers) {
        const uint64_t still = walkers - finished;
        uint32_t next = 0;
#pragma omp parallel num_threads(nt) reduction(+ : tot_steps, tot_visits, tot_dangling, finished)
        {
            const int tid = omp_get_thread_num();
            xorshift_t *g = &gens[tid];
            pw_pool taken;
            memset(&taken, 0, sizeof(taken));
            for (;;) {
                const uint32_t pos = __atomic_fetch_add(&next, 1, __ATOMIC_RELAXED);
                if (pos >= nmap) break;
                const uint32_t p = map[pos];
                omp_set_lock(&plock[p]);
                for (int s = 0; s < nt; ++s) {
                    pw_pool *src = &pools[(size_t)s * G.np + p];
                    omp_set_lock(&src->lock);
                    pw_walker *a = src->a;
                    const uint64_t n = src->n;
                    const uint64_t cap = src->cap;
                    src->a = taken.a;
                    src->n = 0;
                    src->cap = taken.cap;
                    omp_unset_lock(&src->lock);
                    taken.a = a;
                    taken.cap = cap;
                    for (uint64_t k = 0; k < n; ++k) {
                        pw_walker w = a[k];
                        for (;;) {
                            if (type0) {
                                cnt[w.vid]++;
                                tot_visits++;
                            }
                            if (w.sid == 0) {
                                finished++;
                                if (ends) ends[w.wid] = G.oldid[w.vid];
                                break;
                            }
                            const uint64_t beg = G.rp[w.vid];
                            const uint32_t deg = (uint32_t)(G.rp[w.vid + 1] - beg);
                            if (deg == 0) {
                                tot_dangling++;
                                finished++;
                                if (ends) ends[w.wid] = G.oldid[w.vid];
                                break;
                            }
                            uint32_t nx;
                            if (w.vid < G.H) {
                                pw_buf *b = &hb[w.vid];
                                if (b->head == b->end) {
                                    uint32_t len = pw_buf_len(deg, still, walk_len, E);
                                    const uint32_t room = b->end - (w.vid ? hb[w.vid - 1].end : 0);
                                    if (len > room) len = room;
                                    for (uint32_t e = b->end - len; e < b->end; ++e) units[e] = G.cl[beg + xs_gen(g, deg)];
                                    b->head = b->end - len;
                                }
                                nx = units[b->head++];
                            } else {
                                nx = G.cl[beg + xs_gen(g, deg)];
                            }
                            w.vid = nx;
                            w.sid--;
                            tot_steps++;
                            const uint32_t q = pw_part(&G, nx);
                            if (q != p) {
                                pw_pool *dst = &pools[(size_t)tid * G.np + q];
                                omp_set_lock(&dst->lock);
                                pw_push(dst, w);
                                omp_unset_lock(&dst->lock);
                                break;
                            }
                        }
                    }
                }
                omp_unset_lock(&plock[p]);
            }
            free(taken.a);
        }
    }
    const double t1 = omp_get_wtime();
    if (walk_seconds) *walk_seconds = t1 - t0;
    if (partitions) {
        partitions[0] = G.np_hub;
        partitions[1] = nlow;
    }
    if (counts && cnt)
        for (uint64_t u = 0; u < Vp; ++u)
            if (G.oldid[u] != ORC_INVALID) counts[G.oldid[u]] += cnt[u];
    if (st) {
        st->steps = tot_steps;
        st->visits = tot_visits;
        st->dangling = tot_dangling;
        st->proposals = 0;
        st->exact_draws = 0;
    }
    for (size_t i = 0; i < (size_t)nt * G.np; ++i) {
        omp_destroy_lock(&pools[i].lock);
        free(pools[i].a);
    }
    for (uint32_t p = 0; p < G.np; ++p) omp_destroy_lock(&plock[p]);
    free(pools);
    free(plock);
    free(map);
    free(cnt);
    free(gens);
    free(units);
    free(hb);
    free(G.rp);
    free(G.cl);
    free(G.newid);
    free(G.oldid);
    free(G.pend);
    return 0;
}
