// Host-side graph ingest and CSR construction.
//
// Load semantics follow the reference loader:
//   text:   one "src dst" pair of unsigned decimals per line, lines that
//           start with '#' skipped (read_text_graph, include/io.hpp:76-90)
//   binary: a raw array of {u32 src, u32 dst} records with no header
//           (read_binary_graph, include/io.hpp:39-53)
//   ids:    compacted in order of first appearance, scanning each edge's
//           src then dst (src/core/graph.hpp:293-319); with as_undirected
//           the reversed copies follow all forward edges (:289-317)
//   CSR:    each vertex's neighbours in edge order (src/core/graph.hpp:331-366)
// Deviation: a line with fewer than two numbers is skipped; the reference's
// sscanf would push the previous line's pair again (undefined on line 1).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_map>

#include "../../include/ptw.h"
#include "engine.hpp"
#include "philox.hpp"

namespace ptw {

unsigned host_threads()
{
    unsigned n = std::thread::hardware_concurrency();
    if (const char *e = getenv("OMP_NUM_THREADS")) {
        int v = atoi(e);
        if (v > 0) n = (unsigned)v;
    }
    if (n == 0) n = 1;
    if (n > 64) n = 64;
    return n;
}

void parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)> &body, unsigned threads,
                  uint64_t min_per_thread)
{
    if (threads == 0) threads = host_threads();
    if (min_per_thread == 0) min_per_thread = 1;
    const uint64_t useful = (n + min_per_thread - 1) / min_per_thread;
    if (useful < threads) threads = (unsigned)useful;
    if (threads <= 1) {
        if (n) body(0, n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(threads);
    for (unsigned t = 0; t < threads; ++t) {
        uint64_t lo = n * t / threads, hi = n * (t + 1) / threads;
        pool.emplace_back([&body, lo, hi] { if (hi > lo) body(lo, hi); });
    }
    for (auto &th : pool) th.join();
}

namespace {

struct Mapped {
    const char *data = nullptr;
    size_t size = 0;
    int fd = -1;
    ~Mapped()
    {
        if (data && size) munmap((void *)data, size);
        if (fd >= 0) close(fd);
    }
};

int map_file(const char *path, Mapped &m, std::string &err)
{
    m.fd = open(path, O_RDONLY);
    if (m.fd < 0) {
        err = std::string("cannot open graph file '") + path + "': " + strerror(errno);
        return PTW_ERR_IO;
    }
    struct stat st;
    if (fstat(m.fd, &st) != 0) {
        err = std::string("cannot stat '") + path + "'";
        return PTW_ERR_IO;
    }
    m.size = (size_t)st.st_size;
    if (m.size == 0) return PTW_OK;
    void *p = mmap(nullptr, m.size, PROT_READ, MAP_PRIVATE, m.fd, 0);
    if (p == MAP_FAILED) {
        err = std::string("cannot map '") + path + "'";
        m.data = nullptr;
        m.size = 0;
        return PTW_ERR_IO;
    }
    madvise(p, m.size, MADV_SEQUENTIAL);
    m.data = (const char *)p;
    return PTW_OK;
}

// Parse "%u %u" from [p, end) of one line.  Returns false if the line does
// not hold two numbers.
inline bool parse_pair(const char *p, const char *end, uint32_t &a, uint32_t &b)
{
    uint64_t v[2];
    for (int k = 0; k < 2; ++k) {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
        if (p < end && *p == '+') ++p;
        if (p >= end || *p < '0' || *p > '9') return false;
        uint64_t x = 0;
        while (p < end && *p >= '0' && *p <= '9') {
            x = x * 10 + (uint64_t)(*p - '0');
            if (x > 0xFFFFFFFFull) x = 0xFFFFFFFFull;  // %u saturates differently; ids are u32
            ++p;
        }
        v[k] = x;
    }
    a = (uint32_t)v[0];
    b = (uint32_t)v[1];
    return true;
}

int read_text(const char *path, std::vector<uint32_t> &src, std::vector<uint32_t> &dst,
              std::string &err)
{
    Mapped m;
    int rc = map_file(path, m, err);
    if (rc) return rc;
    const char *data = m.data;
    const size_t size = m.size;
    unsigned T = host_threads();
    if (size < (1u << 20)) T = 1;
    // chunk boundaries at line starts
    std::vector<size_t> cut(T + 1, size);
    cut[0] = 0;
    for (unsigned t = 1; t < T; ++t) {
        size_t pos = size * t / T;
        while (pos < size && data[pos - 1] != '\n') ++pos;
        cut[t] = std::max(pos, cut[t - 1]);
    }
    std::vector<std::vector<uint32_t>> ls(T), ld(T);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < T; ++t) {
        pool.emplace_back([&, t] {
            const char *p = data + cut[t];
            const char *end = data + cut[t + 1];
            auto &S = ls[t];
            auto &D = ld[t];
            S.reserve((size_t)(end - p) / 12 + 16);
            D.reserve((size_t)(end - p) / 12 + 16);
            while (p < end) {
                const char *nl = (const char *)memchr(p, '\n', (size_t)(end - p));
                const char *le = nl ? nl : end;
                uint32_t a, b;
                if (le > p && *p != '#' && parse_pair(p, le, a, b)) {
                    S.push_back(a);
                    D.push_back(b);
                }
                p = nl ? nl + 1 : end;
            }
        });
    }
    for (auto &th : pool) th.join();
    size_t total = 0;
    for (auto &v : ls) total += v.size();
    src.resize(total);
    dst.resize(total);
    size_t off = 0;
    for (unsigned t = 0; t < T; ++t) {
        std::copy(ls[t].begin(), ls[t].end(), src.begin() + off);
        std::copy(ld[t].begin(), ld[t].end(), dst.begin() + off);
        off += ls[t].size();
        std::vector<uint32_t>().swap(ls[t]);
        std::vector<uint32_t>().swap(ld[t]);
    }
    return PTW_OK;
}

int read_binary(const char *path, std::vector<uint32_t> &src, std::vector<uint32_t> &dst,
                std::string &err)
{
    Mapped m;
    int rc = map_file(path, m, err);
    if (rc) return rc;
    // read_binary_graph truncates a trailing partial record (total / sizeof(Edge))
    const size_t n = m.size / 8;
    src.resize(n);
    dst.resize(n);
    const uint32_t *w = (const uint32_t *)m.data;
    parallel_for(n, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            uint32_t s, d;
            memcpy(&s, w + 2 * i, 4);
            memcpy(&d, w + 2 * i + 1, 4);
            src[i] = s;
            dst[i] = d;
        }
    });
    return PTW_OK;
}

// CSR from an edge list in compact ids, adjacency in edge order.
void csr_from_edges(uint32_t V, const std::vector<uint32_t> &src, const std::vector<uint32_t> &dst,
                    bool as_undirected, HostCsr &g)
{
    const uint64_t m = src.size();
    g.V = V;
    g.row_ptr.assign((size_t)V + 1, 0);
    std::vector<uint64_t> deg(V, 0);
    for (uint64_t e = 0; e < m; ++e) {
        deg[src[e]]++;
        if (as_undirected) deg[dst[e]]++;
    }
    uint64_t acc = 0;
    for (uint32_t v = 0; v < V; ++v) {
        g.row_ptr[v] = (uint32_t)acc;
        acc += deg[v];
    }
    g.row_ptr[V] = (uint32_t)acc;
    g.col.resize(acc);
    std::vector<uint32_t> fill(g.row_ptr.begin(), g.row_ptr.end() - 1);
    for (uint64_t e = 0; e < m; ++e) g.col[fill[src[e]]++] = dst[e];
    if (as_undirected)
        for (uint64_t e = 0; e < m; ++e) g.col[fill[dst[e]]++] = src[e];
}

}  // namespace

int load_edge_list(const char *path, int format, bool as_undirected, HostCsr &g, std::string &err)
{
    if (!path) {
        err = "graph path is NULL";
        return PTW_ERR_INVALID_ARG;
    }
    if (format != PTW_FORMAT_TEXT && format != PTW_FORMAT_BINARY) {
        err = "unknown graph format " + std::to_string(format);
        return PTW_ERR_INVALID_ARG;
    }
    std::vector<uint32_t> src, dst;
    int rc = format == PTW_FORMAT_TEXT ? read_text(path, src, dst, err) : read_binary(path, src, dst, err);
    if (rc) return rc;
    const uint64_t m = src.size();
    const uint64_t total = as_undirected ? 2 * m : m;
    if (total >= 0xFFFFFFFFull) {
        err = "graph has " + std::to_string(total) + " edges; u32 offsets hold at most 2^32 - 2";
        return PTW_ERR_UNSUPPORTED;
    }
    // first-appearance compaction (src/core/graph.hpp:293-319)
    uint32_t max_name = 0;
    for (uint64_t e = 0; e < m; ++e) max_name = std::max(max_name, std::max(src[e], dst[e]));
    std::vector<uint32_t> names;  // the inverse: file id of each compact id
    uint32_t V = 0;
    // A table indexed by file id, as the reference keeps; a hash map when the
    // ids are sparse (the table would be far larger than the edge list).
    auto compact = [&](auto &name2id) {
        for (uint64_t e = 0; e < m; ++e) {
            uint32_t &a = name2id[src[e]];
            if (a == PTW_INVALID_VERTEX) {
                a = V++;
                names.push_back(src[e]);
            }
            uint32_t &b = name2id[dst[e]];
            if (b == PTW_INVALID_VERTEX) {
                b = V++;
                names.push_back(dst[e]);
            }
            src[e] = a;
            dst[e] = b;
        }
    };
    if ((uint64_t)max_name + 1 <= 16 * m + (1u << 20)) {
        std::vector<uint32_t> name2id(m ? (size_t)max_name + 1 : 0, PTW_INVALID_VERTEX);
        compact(name2id);
    } else {
        struct Slot {
            uint32_t v = PTW_INVALID_VERTEX;
        };
        std::unordered_map<uint32_t, Slot> map;
        map.reserve(2 * m);
        struct Ref {
            std::unordered_map<uint32_t, Slot> &m;
            uint32_t &operator[](uint32_t k) { return m[k].v; }
        } name2id{map};
        compact(name2id);
    }
    csr_from_edges(V, src, dst, as_undirected, g);
    g.names = std::move(names);
    return PTW_OK;
}

int write_edge_list(const HostCsr &g, const char *path, int format, std::string &err)
{
    FILE *f = fopen(path, "wb");
    if (!f) {
        err = std::string("cannot open '") + path + "' for writing: " + strerror(errno);
        return PTW_ERR_IO;
    }
    const size_t buf_cap = 1 << 22;
    std::vector<char> buf(buf_cap + 64);
    size_t len = 0;
    bool ok = true;
    for (uint32_t u = 0; u < g.V && ok; ++u) {
        for (uint32_t k = g.row_ptr[u]; k < g.row_ptr[u + 1]; ++k) {
            if (format == PTW_FORMAT_TEXT) {
                len += (size_t)snprintf(buf.data() + len, 64, "%u %u\n", u, g.col[k]);
            } else {
                uint32_t rec[2] = {u, g.col[k]};
                memcpy(buf.data() + len, rec, 8);
                len += 8;
            }
            if (len >= buf_cap) {
                ok = fwrite(buf.data(), 1, len, f) == len;
                len = 0;
            }
        }
    }
    if (ok && len) ok = fwrite(buf.data(), 1, len, f) == len;
    if (fclose(f) != 0) ok = false;
    if (!ok) {
        err = std::string("short write to '") + path + "'";
        return PTW_ERR_IO;
    }
    return PTW_OK;
}

int validate_csr(const uint32_t *row_ptr, const uint32_t *col, uint32_t V, uint64_t E, std::string &err)
{
    if (!row_ptr || (E && !col)) {
        err = "CSR pointers are NULL";
        return PTW_ERR_INVALID_ARG;
    }
    if (V == 0) {
        err = "graph has no vertices";
        return PTW_ERR_INVALID_ARG;
    }
    if (E >= 0xFFFFFFFFull) {
        err = "u32 offsets hold at most 2^32 - 2 edges";
        return PTW_ERR_UNSUPPORTED;
    }
    if (row_ptr[0] != 0 || row_ptr[V] != E) {
        err = "row_ptr[0] must be 0 and row_ptr[V] must equal E";
        return PTW_ERR_INVALID_ARG;
    }
    std::atomic<int> bad{0};
    parallel_for(V, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t v = lo; v < hi; ++v)
            if (row_ptr[v] > row_ptr[v + 1]) bad = 1;
    });
    if (bad) {
        err = "row_ptr is not monotone";
        return PTW_ERR_INVALID_ARG;
    }
    parallel_for(E, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t e = lo; e < hi; ++e)
            if (col[e] >= V) bad = 2;
    });
    if (bad) {
        err = "col holds a vertex id >= V";
        return PTW_ERR_INVALID_ARG;
    }
    return PTW_OK;
}

// R-MAT edge e: `scale` levels, each choosing a quadrant from one u32 of a
// Philox block keyed by the generator seed and counted by (e, level / 4).
int build_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
               bool keep_isolated, HostCsr &g, std::string &err)
{
    if (scale == 0 || scale > 31 || edge_factor == 0) {
        err = "R-MAT needs 1 <= scale <= 31 and edge_factor >= 1";
        return PTW_ERR_INVALID_ARG;
    }
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) {
        err = "R-MAT probabilities must be >= 0 with a + b + c <= 1";
        return PTW_ERR_INVALID_ARG;
    }
    const uint64_t N = 1ull << scale;
    const uint64_t M = (uint64_t)edge_factor * N;
    if (2 * M >= 0xFFFFFFFFull) {
        err = "R-MAT edge count exceeds u32 offsets";
        return PTW_ERR_UNSUPPORTED;
    }
    auto q = [](double p) -> uint64_t {
        double t = std::floor(p * 4294967296.0);
        return t > 4294967296.0 ? 4294967296ull : (uint64_t)t;
    };
    const uint64_t ta = q(a), tb = q(a + b), tc = q(a + b + c);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    auto gen = [&](uint64_t e, uint32_t &s, uint32_t &d) {
        uint32_t su = 0, du = 0;
        for (uint32_t lvl = 0; lvl < scale; lvl += 4) {
            u32x4 r = philox4x32_10((uint32_t)e, (uint32_t)(e >> 32), lvl / 4, 0x524D4154u, k0, k1);
            uint32_t w[4] = {r.x, r.y, r.z, r.w};
            for (uint32_t j = 0; j < 4 && lvl + j < scale; ++j) {
                uint64_t x = w[j];
                uint32_t sb = 0, db = 0;
                if (x < ta) {
                } else if (x < tb) {
                    db = 1;
                } else if (x < tc) {
                    sb = 1;
                } else {
                    sb = 1;
                    db = 1;
                }
                su = (su << 1) | sb;
                du = (du << 1) | db;
            }
        }
        s = su;
        d = du;
    };
    // pass 1: degree counts of the symmetrized multigraph without self-loops
    std::vector<std::atomic<uint32_t>> cnt(N);
    parallel_for(N, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t v = lo; v < hi; ++v) cnt[v].store(0, std::memory_order_relaxed);
    });
    parallel_for(M, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t e = lo; e < hi; ++e) {
            uint32_t s, d;
            gen(e, s, d);
            if (s == d) continue;
            cnt[s].fetch_add(1, std::memory_order_relaxed);
            cnt[d].fetch_add(1, std::memory_order_relaxed);
        }
    });
    std::vector<uint64_t> off(N + 1, 0);
    for (uint64_t v = 0; v < N; ++v) off[v + 1] = off[v] + cnt[v].load(std::memory_order_relaxed);
    std::vector<uint32_t> buf(off[N]);
    parallel_for(N, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t v = lo; v < hi; ++v) cnt[v].store(0, std::memory_order_relaxed);
    });
    // pass 2: scatter (order within a bucket is racy, fixed by the sort below)
    parallel_for(M, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t e = lo; e < hi; ++e) {
            uint32_t s, d;
            gen(e, s, d);
            if (s == d) continue;
            buf[off[s] + cnt[s].fetch_add(1, std::memory_order_relaxed)] = d;
            buf[off[d] + cnt[d].fetch_add(1, std::memory_order_relaxed)] = s;
        }
    });
    // sort + dedup each bucket
    std::vector<uint32_t> udeg(N);
    parallel_for(N, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t v = lo; v < hi; ++v) {
            uint32_t *p = buf.data() + off[v];
            uint32_t n = (uint32_t)(off[v + 1] - off[v]);
            std::sort(p, p + n);
            udeg[v] = (uint32_t)(std::unique(p, p + n) - p);
        }
    });
    // compaction of isolated ids (order preserving)
    std::vector<uint32_t> newid(N);
    uint32_t V = 0;
    for (uint64_t v = 0; v < N; ++v) {
        if (keep_isolated || udeg[v] > 0)
            newid[v] = V++;
        else
            newid[v] = PTW_INVALID_VERTEX;
    }
    g.names.clear();
    if (V < N) {  // the generator's id of each kept vertex
        g.names.resize(V);
        for (uint64_t v = 0; v < N; ++v)
            if (newid[v] != PTW_INVALID_VERTEX) g.names[newid[v]] = (uint32_t)v;
    }
    g.V = V;
    g.row_ptr.assign((size_t)V + 1, 0);
    {
        uint64_t acc = 0;
        for (uint64_t v = 0; v < N; ++v) {
            if (newid[v] == PTW_INVALID_VERTEX) continue;
            g.row_ptr[newid[v]] = (uint32_t)acc;
            acc += udeg[v];
        }
        g.row_ptr[V] = (uint32_t)acc;
        g.col.resize(acc);
    }
    parallel_for(N, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t v = lo; v < hi; ++v) {
            uint32_t id = newid[v];
            if (id == PTW_INVALID_VERTEX) continue;
            const uint32_t *p = buf.data() + off[v];
            uint32_t *o = g.col.data() + g.row_ptr[id];
            for (uint32_t k = 0; k < udeg[v]; ++k) o[k] = newid[p[k]];
        }
    });
    return PTW_OK;
}

void random_weights(uint64_t E, uint64_t seed, std::vector<float> &w)
{
    w.resize(E);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    parallel_for(E, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t e = lo; e < hi; ++e) {
            u32x4 r = philox4x32_10((uint32_t)e, (uint32_t)(e >> 32), 0, 0x57474854u, k0, k1);
            w[e] = (float)((r.x >> 8) + 1) * (1.0f / 16777216.0f);  // exact in f32
        }
    });
}

int build_alias_tables(const HostCsr &g, const std::vector<float> &w, std::vector<uint32_t> &thr,
                       std::vector<uint32_t> &alias, std::string &err)
{
    const uint64_t E = g.E();
    if (w.size() != E) {
        err = "weight count differs from the edge count";
        return PTW_ERR_INVALID_ARG;
    }
    std::atomic<int> bad{0};
    parallel_for(E, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t e = lo; e < hi; ++e)
            if (!(w[e] > 0.0f) || !std::isfinite(w[e])) bad = 1;
    });
    if (bad) {
        err = "edge weights must be finite and > 0";
        return PTW_ERR_INVALID_ARG;
    }
    thr.assign(E, 0);
    alias.assign(E, 0);
    parallel_for(g.V, [&](uint64_t lo, uint64_t hi) {
        std::vector<uint64_t> m;
        std::vector<uint32_t> small, large;
        for (uint64_t v = lo; v < hi; ++v) {
            const uint32_t beg = g.row_ptr[v], deg = g.row_ptr[v + 1] - beg;
            if (deg == 0) continue;
            double wmax = 0.0;
            for (uint32_t k = 0; k < deg; ++k) wmax = std::max(wmax, (double)w[beg + k]);
            m.resize(deg);
            uint64_t Q = 0;
            for (uint32_t k = 0; k < deg; ++k) {
                double qd = std::floor((double)w[beg + k] / wmax * 16777216.0 + 0.5);
                uint64_t q = qd < 1.0 ? 1 : (qd > 16777216.0 ? 16777216 : (uint64_t)qd);
                m[k] = q;
                Q += q;
            }
            small.clear();
            large.clear();
            for (uint32_t k = 0; k < deg; ++k) {
                m[k] *= deg;
                (m[k] < Q ? small : large).push_back(k);
            }
            uint32_t *T = thr.data() + beg, *A = alias.data() + beg;
            while (!small.empty() && !large.empty()) {
                const uint32_t s = small.back();
                small.pop_back();
                const uint32_t l = large.back();
                large.pop_back();
                T[s] = (uint32_t)(((unsigned __int128)m[s] << 32) / Q);
                A[s] = l;
                m[l] -= Q - m[s];
                (m[l] < Q ? small : large).push_back(l);
            }
            for (uint32_t k : large) {
                T[k] = 0xFFFFFFFFu;
                A[k] = k;
            }
            for (uint32_t k : small) {  // unreachable with exact sums; kept total
                T[k] = 0xFFFFFFFFu;
                A[k] = k;
            }
        }
    });
    return PTW_OK;
}

void relabel_by_degree(const HostCsr &g, std::vector<uint32_t> &perm, std::vector<uint32_t> &inv,
                       std::vector<uint32_t> &d_row_ptr, std::vector<uint32_t> &d_col)
{
    const uint32_t V = g.V;
    const uint32_t *rp = g.row_ptr.data();
    // Device order: degree descending, ties by graph id ascending -- a
    // counting sort on the degree (stable: vertices are placed in id order).
    uint32_t maxdeg = 0;
    for (uint32_t u = 0; u < V; ++u) maxdeg = std::max(maxdeg, rp[u + 1] - rp[u]);
    std::vector<uint64_t> cnt((size_t)maxdeg + 1, 0), first((size_t)maxdeg + 1, 0);
    for (uint32_t u = 0; u < V; ++u) ++cnt[rp[u + 1] - rp[u]];
    uint64_t pos = 0;
    for (uint64_t d = (uint64_t)maxdeg + 1; d-- > 0;) {  // positions of degree d: [first[d], first[d] + cnt[d])
        first[d] = pos;
        pos += cnt[d];
    }
    inv.resize(V);
    {
        std::vector<uint64_t> next(first);
        for (uint32_t u = 0; u < V; ++u) inv[next[rp[u + 1] - rp[u]]++] = u;
    }
    perm.resize(V);
    parallel_for(V, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t d = lo; d < hi; ++d) perm[inv[d]] = (uint32_t)d;
    });
    // row offsets in device order, block by block: no vertex is looked up
    d_row_ptr.assign((size_t)V + 1, 0);
    uint64_t acc = 0;
    for (uint64_t d = (uint64_t)maxdeg + 1; d-- > 0;) {
        for (uint64_t k = first[d]; k < first[d] + cnt[d]; ++k) {
            d_row_ptr[k] = (uint32_t)acc;
            acc += d;
        }
    }
    d_row_ptr[V] = (uint32_t)acc;
    d_col.resize(acc);
    parallel_for(V, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t d = lo; d < hi; ++d) {
            uint32_t u = inv[d];
            const uint32_t *s = g.col.data() + rp[u];
            uint32_t *o = d_col.data() + d_row_ptr[d];
            uint32_t n = rp[u + 1] - rp[u];
            for (uint32_t k = 0; k < n; ++k) o[k] = perm[s[k]];
        }
    });
}

}  // namespace ptw
