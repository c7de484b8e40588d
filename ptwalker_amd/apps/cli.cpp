// Shared CLI for the four walk binaries.  Flags, defaults and error exits
// follow the reference's option classes (include/option.hpp); the stdout
// lines follow the reference's printf calls so log scrapers keep working:
//   "Partition time %lf s"                             src/core/graph.hpp:485
//   "Root vertex %u"                                   src/core/ppr.cpp:56
//   "start to walk %lf"                                src/core/solver.hpp:1057
//   "init walks start %lf"                             src/core/solver.hpp:859
//   "***********iter %d, finished_walkers %u totstep %lu time %lf"   :926-927
//   "update %lf message %lf"                           src/core/solver.hpp:1025
//   "total time (including init start walker) %lf, finished step %lu" :1073
//   "ppr value %u"                                     src/core/pagerank.cpp:110
// The GPU has no NUMA threads, so "thread-%d bound to socket-%d" is not
// printed; "finished step" reports the true step count (the reference's is
// always 0 because its per-thread step counters are never incremented).
#include "cli.hpp"

#include <cerrno>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/ptw.h"

namespace ptw_cli {

namespace {

struct Opts {
    // graph
    std::string format, graph;
    bool have_format = false, have_graph = false;
    // numa (accepted; the GPU engine has one device)
    int threads = 0, sockets = 1;
    bool have_threads = false, have_sockets = false;
    std::string socket_mapping;
    uint64_t mem_gib = 0;
    // walk
    bool have_epoch = false, have_walkers = false, have_len = false, have_sp = false;
    long long epoch = 0;
    uint64_t walkers = 0;
    long long walk_len = 0;
    double sp = 0.0;
    bool have_p = false, have_q = false;
    double n2v_p = 1.0, n2v_q = 1.0;
    long long walker_size = 4196, degree_bod = 32;
    // engine extras
    uint64_t seed = 1;
    bool have_source = false;
    uint32_t source = 0;
    int device = 0;
    std::string out;
    bool json = false;
    int semantics = PTW_SEM_REFERENCE;
    int dangling = PTW_DANGLING_ERROR;
    bool undirected = false;
    std::string start = "";
    // ids of --source, --queries and every output: the input's own (file) ids,
    // or the compact graph ids the reference assigns on load
    bool ids_original = true;
    std::string queries;  // batch mode (PPR, RWR): one query per line
    uint32_t topk = 100;
};

void usage(FILE *f, const char *prog, Parser parser)
{
    fprintf(f, "  %s {OPTIONS}\n\n", prog);
    fprintf(f, "  OPTIONS:\n\n");
    fprintf(f, "      -h, --help                        Display this help menu\n");
    fprintf(f, "      -t[threads]                       [optional] number of threads this program will use\n");
    fprintf(f, "      -s[sockets]                       [optional] number of sockets\n");
    fprintf(f, "      --socket-mapping=[socket-mapping] [optional] example: --socket-mapping=0,1,2,3\n");
    fprintf(f, "      --mem=[mem]                       [optional] Maximum memory this program will use (in GiB)\n");
    fprintf(f, "      -f[format]                        graph format: binary | text\n");
    fprintf(f, "      -g[graph]                         graph path\n");
    fprintf(f, "      -e[epoch]                         walk epoch number\n");
    fprintf(f, "      -w[walker]                        walker number\n");
    fprintf(f, "      -l[length]                        walk length\n");
    if (parser == Parser::RWPT) {
        fprintf(f, "      --sp=[startprb]                   the rastart probability\n");
        fprintf(f, "      --ws=[walkersize]                 how many walkers each walekr array stores\n");
        fprintf(f, "      --db=[degree_boundary]            the boundary to partition high-low vertex\n");
    } else {
        fprintf(f, "      --sp=[startprb]                   [optional, RWR] termination probability (default 0.5)\n");
    }
    if (parser == Parser::NODE2VEC) {
        fprintf(f, "      -p[p]                             node2vec parameter p\n");
        fprintf(f, "      -q[q]                             node2vec parameter q\n");
    }
    fprintf(f, "      --seed=[seed]                     [optional] Philox seed (default 1)\n");
    fprintf(f, "      --source=[vertex]                 [optional] start vertex (PPR, RWR)\n");
    fprintf(f, "      --start=[uniform|roundrobin]      [optional] start rule (deepwalk, pagerank)\n");
    fprintf(f, "      --device=[id]                     [optional] HIP device\n");
    fprintf(f, "      --out=[path]                      [optional] write the result vector\n");
    fprintf(f, "      --json                            [optional] print a JSON result record\n");
    fprintf(f, "      --semantics=[reference|textbook]  [optional] step budget 1+Geom(tp) or Geom(tp)\n");
    fprintf(f, "      --dangling=[error|stop]           [optional] walker at a vertex without out-edges\n");
    fprintf(f, "      --undirected                      [optional] add reversed edges on load\n");
    fprintf(f, "      --out-ids=[original|graph]        [optional] ids of --source, --queries and outputs: the\n"
               "                                        file's own (default) or the compact load order\n");
    if (parser == Parser::RWPT || parser == Parser::WALK) {
        fprintf(f, "      --queries=[path]                  [optional, PPR, RWR] batch mode: lines \"source [walkers\n"
                   "                                        [tp]]\" (defaults -w, --sp); --out gets each query's top-k\n");
        fprintf(f, "      --topk=[k]                        [optional] top-k per query in batch mode (default 100)\n");
    }
}

[[noreturn]] void parse_error(const char *prog, Parser parser, const std::string &msg)
{
    fprintf(stderr, "%s\n", msg.c_str());
    usage(stderr, prog, parser);
    exit(1);
}

[[noreturn]] void check_failed(const std::string &msg)
{
    fprintf(stderr, "Check failed: %s\n", msg.c_str());
    exit(1);
}

bool to_u64(const std::string &s, uint64_t &v)
{
    if (s.empty()) return false;
    char *end = nullptr;
    errno = 0;
    unsigned long long x = strtoull(s.c_str(), &end, 10);
    if (errno || *end || s[0] == '-') return false;
    v = x;
    return true;
}

bool to_ll(const std::string &s, long long &v)
{
    if (s.empty()) return false;
    char *end = nullptr;
    errno = 0;
    long long x = strtoll(s.c_str(), &end, 10);
    if (errno || *end) return false;
    v = x;
    return true;
}

bool to_double(const std::string &s, double &v)
{
    if (s.empty()) return false;
    char *end = nullptr;
    errno = 0;
    double x = strtod(s.c_str(), &end);
    if (errno || *end) return false;
    v = x;
    return true;
}

Opts parse(const char *prog, Parser parser, int argc, char **argv)
{
    Opts o;
    // short flags taking a value, long flags taking a value, long switches
    const char *short_val = parser == Parser::NODE2VEC ? "fgtsewlpq" : "fgtsewl";
    const std::vector<std::string> long_val = {"socket-mapping", "mem", "sp", "ws", "db", "seed", "source",
                                               "device", "out", "semantics", "dangling", "start", "out-ids",
                                               "queries", "topk"};
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        std::string key, val;
        bool has_val = false;
        if (a == "-h" || a == "--help") {
            usage(stdout, prog, parser);
            exit(0);
        }
        if (a.size() >= 2 && a[0] == '-' && a[1] != '-') {
            char c = a[1];
            if (!strchr(short_val, c)) parse_error(prog, parser, "Flag could not be matched: " + a.substr(1));
            key = std::string(1, c);
            if (a.size() > 2) {
                val = a.substr(2);
            } else {
                if (i + 1 >= argc) parse_error(prog, parser, "Flag '" + key + "' requires an argument but received none");
                val = argv[++i];
            }
            has_val = true;
        } else if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
            std::string body = a.substr(2);
            size_t eq = body.find('=');
            key = body.substr(0, eq);
            if (key == "json" || key == "undirected") {
                if (eq != std::string::npos) parse_error(prog, parser, "Flag '" + key + "' does not take an argument");
                if (key == "json") o.json = true;
                else o.undirected = true;
                continue;
            }
            bool known = false;
            for (auto &k : long_val) known |= (k == key);
            if (!known || (parser != Parser::RWPT && (key == "ws" || key == "db")))
                parse_error(prog, parser, "Flag could not be matched: " + key);
            if (eq != std::string::npos) {
                val = body.substr(eq + 1);
            } else {
                if (i + 1 >= argc) parse_error(prog, parser, "Flag '" + key + "' requires an argument but received none");
                val = argv[++i];
            }
            has_val = true;
        } else {
            parse_error(prog, parser, "Passed in argument, but no positional arguments were ready to receive it: " + a);
        }
        (void)has_val;
        auto bad = [&]() { parse_error(prog, parser, "Argument '" + key + "' received invalid value type '" + val + "'"); };
        long long ll;
        uint64_t u;
        double dv;
        if (key == "f") { o.format = val; o.have_format = true; }
        else if (key == "g") { o.graph = val; o.have_graph = true; }
        else if (key == "t") { if (!to_ll(val, ll)) bad(); o.threads = (int)ll; o.have_threads = true; }
        else if (key == "s") { if (!to_ll(val, ll)) bad(); o.sockets = (int)ll; o.have_sockets = true; }
        else if (key == "e") { if (!to_ll(val, ll)) bad(); o.epoch = ll; o.have_epoch = true; }
        else if (key == "w") { if (!to_u64(val, u)) bad(); o.walkers = u; o.have_walkers = true; }
        else if (key == "l") { if (!to_ll(val, ll)) bad(); o.walk_len = ll; o.have_len = true; }
        else if (key == "p") { if (!to_double(val, dv)) bad(); o.n2v_p = dv; o.have_p = true; }
        else if (key == "q") { if (!to_double(val, dv)) bad(); o.n2v_q = dv; o.have_q = true; }
        else if (key == "socket-mapping") { o.socket_mapping = val; }
        else if (key == "mem") { if (!to_u64(val, u)) bad(); o.mem_gib = u; }
        else if (key == "sp") { if (!to_double(val, dv)) bad(); o.sp = dv; o.have_sp = true; }
        else if (key == "ws") { if (!to_ll(val, ll)) bad(); o.walker_size = ll; }
        else if (key == "db") { if (!to_ll(val, ll)) bad(); o.degree_bod = ll; }
        else if (key == "seed") { if (!to_u64(val, u)) bad(); o.seed = u; }
        else if (key == "source") { if (!to_u64(val, u) || u > 0xFFFFFFFEull) bad(); o.source = (uint32_t)u; o.have_source = true; }
        else if (key == "device") { if (!to_ll(val, ll)) bad(); o.device = (int)ll; }
        else if (key == "out") { o.out = val; }
        else if (key == "semantics") {
            if (val == "reference") o.semantics = PTW_SEM_REFERENCE;
            else if (val == "textbook") o.semantics = PTW_SEM_TEXTBOOK;
            else bad();
        } else if (key == "dangling") {
            if (val == "error") o.dangling = PTW_DANGLING_ERROR;
            else if (val == "stop") o.dangling = PTW_DANGLING_STOP;
            else bad();
        } else if (key == "start") {
            if (val != "uniform" && val != "roundrobin") bad();
            o.start = val;
        } else if (key == "out-ids") {
            if (val == "original") o.ids_original = true;
            else if (val == "graph") o.ids_original = false;
            else bad();
        } else if (key == "queries") {
            if (parser == Parser::NODE2VEC) parse_error(prog, parser, "Flag could not be matched: " + key);
            o.queries = val;
        } else if (key == "topk") {
            if (!to_u64(val, u) || u == 0 || u > 1024) bad();
            o.topk = (uint32_t)u;
        }
    }
    // validation in the order of the reference's parse() methods
    if (o.have_sockets && o.sockets < 1) check_failed("sockets >= 1");
    if (o.have_threads && o.threads < 1) check_failed("threads >= 1");
    if (!o.have_format) check_failed("graph_format_flag");
    if (o.format != "binary" && o.format != "text") {
        fprintf(stderr, "[error] Unknown graph format: %s\n", o.format.c_str());
        exit(1);
    }
    if (!o.have_graph) check_failed("graph_path_flag");
    if (!(o.have_epoch || o.have_walkers)) check_failed("epoch_num_flag || walker_num_flag");
    if (o.have_epoch && o.have_walkers) check_failed("!(epoch_num_flag && walker_num_flag)");
    if (parser == Parser::RWPT) {
        if (!o.have_sp) check_failed("rastart_probability_flag");
        if (!o.have_len) o.walk_len = (long long)(1.0 / o.sp) + 1;
    } else {
        if (!o.have_len) check_failed("walk_len_flag");
    }
    if (parser == Parser::NODE2VEC) {
        if (!o.have_p) check_failed("p_flag");
        if (!o.have_q) check_failed("q_flag");
    }
    return o;
}

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

const char *app_name(int app)
{
    switch (app) {
    case PTW_APP_PAGERANK: return "pagerank";
    case PTW_APP_PPR: return "PPR";
    case PTW_APP_RWR: return "RWR";
    case PTW_APP_NODE2VEC: return "node2vec";
    default: return "deepwalk";
    }
}

}  // namespace

namespace {

// Graph id <-> the ids of the command line and the outputs.
struct IdMap {
    bool original = true;
    std::vector<uint32_t> names;  // graph id -> file id
    std::unordered_map<uint32_t, uint32_t> index;  // file id -> graph id, built on first lookup
    uint32_t out(uint32_t v) const { return original && v != PTW_INVALID_VERTEX ? names[v] : v; }
    // false if the id names no vertex of the graph
    bool in(uint32_t id, uint32_t &v)
    {
        if (!original) {
            v = id;
            return id < names.size();
        }
        if (index.empty()) {
            index.reserve(names.size());
            for (uint32_t g = 0; g < names.size(); ++g) index.emplace(names[g], g);
        }
        auto it = index.find(id);
        if (it == index.end()) return false;
        v = it->second;
        return true;
    }
};

// Batch mode: every line of the query file is "source [walkers [tp]]" ('#'
// starts a comment); all queries run in one batched launch and --out gets the
// top-k of each (ptw_batch_write_topk).
int run_batch(ptw_ctx *ctx, int app, const Opts &o, const ptw_graph_info &gi, IdMap &ids)
{
    FILE *qf = fopen(o.queries.c_str(), "r");
    if (!qf) {
        fprintf(stderr, "[error] cannot open %s\n", o.queries.c_str());
        return 1;
    }
    std::vector<ptw_query> qs;
    char line[512];
    unsigned long lineno = 0;
    const double tp_default = o.have_sp ? o.sp : (app == PTW_APP_RWR ? 0.5 : 0.15);
    while (fgets(line, sizeof line, qf)) {
        ++lineno;
        char *c = line;
        while (*c == ' ' || *c == '\t') ++c;
        if (*c == '#' || *c == '\n' || *c == '\r' || *c == 0) continue;
        unsigned long long src = 0, walkers = o.walkers;
        double tp = tp_default;
        const int got = sscanf(c, "%llu %llu %lf", &src, &walkers, &tp);
        uint32_t v = 0;
        if (got < 1 || src > 0xFFFFFFFFull || !ids.in((uint32_t)src, v)) {
            fprintf(stderr, "[error] %s:%lu: no vertex %llu in the graph\n", o.queries.c_str(), lineno, src);
            fclose(qf);
            return 1;
        }
        ptw_query q;
        memset(&q, 0, sizeof(q));
        q.source = v;
        q.stream = (uint32_t)qs.size();
        q.walkers = walkers;
        q.tp = tp;
        q.kind = app == PTW_APP_RWR ? PTW_QUERY_ENDPOINTS : PTW_QUERY_VISITS;
        qs.push_back(q);
    }
    fclose(qf);
    ptw_batch_desc bd;
    memset(&bd, 0, sizeof(bd));
    bd.queries = qs.data();
    bd.num_queries = (uint32_t)qs.size();
    bd.topk = o.topk;
    bd.seed = o.seed;
    bd.semantics = o.semantics;
    std::vector<uint32_t> tv((size_t)qs.size() * o.topk + 1);
    std::vector<uint64_t> tc(tv.size()), steps(qs.size() + 1);
    ptw_batch_out bo;
    memset(&bo, 0, sizeof(bo));
    bo.topk_vertex = tv.data();
    bo.topk_count = tc.data();
    bo.steps = steps.data();
    printf("batch queries %zu topk %u\n", qs.size(), o.topk);
    if (ptw_batch_ppr(ctx, &bd, &bo)) {
        fprintf(stderr, "[error] batch: %s\n", ptw_last_error(ctx));
        return 1;
    }
    printf("update %lf message %lf\n", bo.kernel_ms * 1e-3, 0.0);
    printf("total time (including init start walker) %lf, finished step %lu\n", bo.total_ms * 1e-3,
           (unsigned long)bo.total_steps);
    if (!o.out.empty() &&
        ptw_batch_write_topk(ctx, o.out.c_str(), &bd, &bo, ids.original ? PTW_IDS_ORIGINAL : PTW_IDS_GRAPH)) {
        fprintf(stderr, "[error] writing %s: %s\n", o.out.c_str(), ptw_last_error(ctx));
        return 1;
    }
    if (o.json) {
        printf("{\"app\": \"%s\", \"queries\": %zu, \"vertices\": %u, \"edges\": %" PRIu64 ", \"steps\": %" PRIu64
               ", \"seed\": %" PRIu64 ", \"kernel_ms\": %.3f, \"steps_per_s\": %.6g, \"queries_per_s\": %.6g}\n",
               app_name(app), qs.size(), gi.num_vertices, gi.num_edges, bo.total_steps, o.seed, bo.kernel_ms,
               bo.kernel_ms > 0 ? bo.total_steps / (bo.kernel_ms * 1e-3) : 0.0,
               bo.kernel_ms > 0 ? qs.size() / (bo.kernel_ms * 1e-3) : 0.0);
    }
    return 0;
}

}  // namespace

int run(int app, Parser parser, const char *prog, int argc, char **argv)
{
    Opts o = parse(prog, parser, argc, argv);
    if (!o.queries.empty() && app != PTW_APP_PPR && app != PTW_APP_RWR)
        parse_error(prog, parser, "--queries is for the PPR and RWR binaries");
    ptw_ctx *ctx = nullptr;
    int rc = ptw_create(&ctx, o.device);
    if (rc) {
        fprintf(stderr, "[error] cannot open HIP device %d (code %d)\n", o.device, rc);
        return 1;
    }
    auto die = [&](const char *what) {
        fprintf(stderr, "[error] %s: %s\n", what, ptw_last_error(ctx));
        ptw_destroy(ctx);
        return 1;
    };
    fprintf(stderr, "Load graph %s\n", o.graph.c_str());
    const double tl = now_s();
    // node2vec loads its graph as undirected (src/core/node2vec.cpp:18)
    rc = ptw_graph_load(ctx, o.graph.c_str(), o.format == "text" ? PTW_FORMAT_TEXT : PTW_FORMAT_BINARY,
                        (o.undirected || app == PTW_APP_NODE2VEC) ? 1 : 0, 0);
    if (rc) return die("graph load");
    ptw_graph_info gi;
    ptw_graph_info_get(ctx, &gi);
    fprintf(stderr, "Vertices number: %u\n", gi.num_vertices);
    fprintf(stderr, "Edges number: %" PRIu64 "\n", gi.num_edges);
    printf("Partition time %lf s\n", now_s() - tl);
    IdMap ids;
    ids.original = o.ids_original;
    ids.names.resize(gi.num_vertices);
    if (ptw_graph_export_names(ctx, ids.names.data())) return die("graph names");

    if (!o.queries.empty()) {
        rc = run_batch(ctx, app, o, gi, ids);
        ptw_destroy(ctx);
        return rc;
    }

    ptw_walk_desc d;
    memset(&d, 0, sizeof(d));
    d.app = app;
    d.semantics = o.semantics;
    d.dangling_policy = o.dangling;
    d.seed = o.seed;
    d.walkers = o.have_walkers ? o.walkers : (uint64_t)o.epoch * gi.num_vertices;
    d.tp = o.have_sp ? o.sp : 0.5;  // RWR.cpp:122 hard-codes 0.5
    d.walk_len = (uint32_t)(o.walk_len > 0 ? o.walk_len : 1);
    d.start_mode = PTW_START_UNIFORM;
    if (o.start == "roundrobin") d.start_mode = PTW_START_ROUNDROBIN;
    uint32_t source = 0;
    if (o.have_source && !ids.in(o.source, source)) {
        fprintf(stderr, "[error] --source %u is not a vertex of the graph\n", o.source);
        ptw_destroy(ctx);
        return 1;
    }
    if (app == PTW_APP_PPR) {
        // the reference intends the highest-degree vertex (src/core/ppr.cpp:36-45)
        d.start_mode = PTW_START_SOURCE;
        d.source = o.have_source ? source : gi.max_degree_vertex;
        printf("Root vertex %u\n", ids.out(d.source));
    } else if (o.have_source) {
        d.start_mode = PTW_START_SOURCE;
        d.source = source;
    }
    if ((app == PTW_APP_DEEPWALK || app == PTW_APP_NODE2VEC) && o.walk_len < 1) check_failed("walk length >= 1");
    d.p = o.n2v_p;
    d.q = o.n2v_q;
    fprintf(stderr, "walker size %lld\n", o.walker_size);

    std::vector<uint64_t> counts;
    std::vector<uint32_t> ends, starts, paths;
    ptw_walk_out out;
    memset(&out, 0, sizeof(out));
    if (app == PTW_APP_PAGERANK || app == PTW_APP_PPR) {
        counts.assign(gi.num_vertices, 0);
        out.counts = counts.data();
    } else if (app == PTW_APP_RWR) {
        ends.assign(d.walkers, 0);
        out.ends = ends.data();
        if (!o.out.empty()) {  // the reference's {start_v, end_v} pairs (src/core/RWR.cpp:9-13)
            starts.assign(d.walkers, 0);
            out.starts = starts.data();
        }
    } else {
        paths.assign(d.walkers * d.walk_len, 0);
        out.paths = paths.data();
    }
    const double tw = now_s();
    printf("start to walk %lf\n", 0.0);
    printf("init walks start %lf\n", now_s() - tw);
    printf("***********iter %d, finished_walkers %u totstep %lu time %lf\n", 0, 0u, 0ul, now_s() - tw);
    rc = ptw_walk(ctx, &d, &out);
    if (rc) return die("walk");
    const double total = now_s() - tw;
    printf("update %lf message %lf\n", out.kernel_ms * 1e-3, 0.0);
    printf("total time (including init start walker) %lf, finished step %lu\n", total, (unsigned long)out.steps);
    if (app == PTW_APP_PAGERANK || app == PTW_APP_PPR)
        printf("ppr value %u\n", (unsigned)counts[0]);
    else if (app == PTW_APP_DEEPWALK)
        printf("ppr value %u\n", 0u);  // src/core/deepwalk.cpp:61 prints an unused counter

    if (!o.out.empty()) {
        FILE *f = fopen(o.out.c_str(), "wb");
        if (!f) {
            fprintf(stderr, "[error] cannot open %s\n", o.out.c_str());
            ptw_destroy(ctx);
            return 1;
        }
        bool ok = true;
        if (!counts.empty()) {  // "vertex count", vertices in load order
            for (uint32_t v = 0; v < gi.num_vertices && ok; ++v)
                ok = fprintf(f, "%u %" PRIu64 "\n", ids.out(v), counts[v]) > 0;
        } else if (!ends.empty()) {  // "start end", one line per walker
            for (uint64_t w = 0; w < d.walkers && ok; ++w)
                ok = fprintf(f, "%u %u\n", ids.out(starts[w]), ids.out(ends[w])) > 0;
        } else {  // u32 paths, walk_len per walker
            if (ids.original)
                for (auto &x : paths) x = ids.out(x);
            ok = fwrite(paths.data(), 4, paths.size(), f) == paths.size();
        }
        if (fclose(f) != 0 || !ok) {
            fprintf(stderr, "[error] short write to %s\n", o.out.c_str());
            ptw_destroy(ctx);
            return 1;
        }
    }
    if (o.json) {
        printf("{\"app\": \"%s\", \"vertices\": %u, \"edges\": %" PRIu64 ", \"walkers\": %" PRIu64
               ", \"steps\": %" PRIu64 ", \"seed\": %" PRIu64 ", \"kernel_ms\": %.3f, \"walk_s\": %.6f"
               ", \"steps_per_s\": %.6g}\n",
               app_name(app), gi.num_vertices, gi.num_edges, d.walkers, out.steps, d.seed, out.kernel_ms, total,
               out.kernel_ms > 0 ? out.steps / (out.kernel_ms * 1e-3) : 0.0);
    }
    ptw_destroy(ctx);
    return 0;
}

}  // namespace ptw_cli
