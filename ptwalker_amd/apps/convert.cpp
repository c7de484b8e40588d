// Edge-list converters, as the reference's tools are run:
//
//   format_knk   -i <text edge list> -o <binary edge list>
//       Ids compacted in order of first appearance (src, then dst, edge by
//       edge), edges kept in file order, written as u32 pairs; a side file
//       <output>.info.txt holds the header lines.
//       (src/tools/format_knk.cpp:38-57, write_binary_graph in
//       include/io.hpp:56-74)
//
//   format_yahoo -i <adjacency file> [-i <file> ...] [-b <binary out>] [-t <text out>]
//       Input files hold records "src degree dst_1 ... dst_degree" separated
//       by any white space; edges (src, dst_k) are written in input order,
//       ids unchanged, as u32 pairs (header in <out>.info.txt) and/or as text
//       lines "src dst" after the header.
//       (src/tools/format_yahoo.cpp:47-82, write_text_graph in
//       include/io.hpp:92-100)
//
// Text lines follow the library's loader: '#' lines and lines without two
// unsigned numbers are skipped.  Errors print a message and exit 1.
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

struct Edge {
    uint32_t src, dst;
};

[[noreturn]] void die(const std::string &msg)
{
    fprintf(stderr, "%s\n", msg.c_str());
    exit(1);
}

bool parse_u32(const char *&p, const char *end, uint32_t &out)
{
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f' || *p == '\n')) ++p;
    if (p < end && *p == '+') ++p;
    if (p >= end || *p < '0' || *p > '9') return false;
    uint64_t x = 0;
    while (p < end && *p >= '0' && *p <= '9') {
        x = x * 10 + (uint64_t)(*p - '0');
        if (x > 0xFFFFFFFFull) x = 0xFFFFFFFFull;
        ++p;
    }
    out = (uint32_t)x;
    return true;
}

std::string read_all(const std::string &path)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) die("Check failed: cannot open '" + path + "': " + strerror(errno));
    std::string s;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    fclose(f);
    return s;
}

std::vector<Edge> read_text_edges(const std::string &path)
{
    const std::string s = read_all(path);
    std::vector<Edge> edges;
    const char *p = s.data(), *end = s.data() + s.size();
    while (p < end) {
        const char *nl = (const char *)memchr(p, '\n', (size_t)(end - p));
        const char *le = nl ? nl : end;
        if (le > p && *p != '#') {
            const char *q = p;
            uint32_t a, b;
            // the two numbers must lie on this line
            if (parse_u32(q, le, a) && parse_u32(q, le, b)) edges.push_back(Edge{a, b});
        }
        p = nl ? nl + 1 : end;
    }
    return edges;
}

void write_info(const std::string &out, const std::string &header)
{
    const std::string info = out + ".info.txt";
    FILE *f = fopen(info.c_str(), "w");
    if (!f || fputs(header.c_str(), f) < 0 || fclose(f) != 0) die("Check failed: cannot write '" + info + "'");
}

void write_binary(const std::string &out, const std::vector<Edge> &edges, const std::string &header)
{
    write_info(out, header);
    FILE *f = fopen(out.c_str(), "wb");
    if (!f) die("Check failed: cannot open '" + out + "'");
    if (!edges.empty() && fwrite(edges.data(), sizeof(Edge), edges.size(), f) != edges.size())
        die("Check failed: short write to '" + out + "'");
    if (fclose(f) != 0) die("Check failed: cannot close '" + out + "'");
}

void write_text(const std::string &out, const std::vector<Edge> &edges, const std::string &header)
{
    FILE *f = fopen(out.c_str(), "w");
    if (!f) die("Check failed: cannot open '" + out + "'");
    fputs(header.c_str(), f);
    for (const Edge &e : edges) fprintf(f, "%u %u\n", e.src, e.dst);
    if (fclose(f) != 0) die("Check failed: cannot close '" + out + "'");
}

void usage(const char *prog, bool yahoo)
{
    printf("  %s {OPTIONS}\n\n  OPTIONS:\n\n", prog);
    printf("      -h, --help                        Display this help menu\n");
    if (yahoo) {
        printf("      -i[input...]                      input paths\n");
        printf("      -b[binary output]                 binary output path\n");
        printf("      -t[text output]                   text output path\n");
    } else {
        printf("      -i[input]                         input path\n");
        printf("      -o[output]                        output path\n");
    }
}

// Flags as "-x value", "-xvalue" or "--x=value".
bool flag_value(int &i, int argc, char **argv, char name, std::string &out)
{
    const char *a = argv[i];
    if (a[0] == '-' && a[1] == name) {
        if (a[2]) {
            out = a + 2;
            return true;
        }
        if (i + 1 >= argc) die(std::string("Flag '") + name + "' requires an argument but received none");
        out = argv[++i];
        return true;
    }
    return false;
}

int format_knk(int argc, char **argv)
{
    std::string in, out;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "-h") || !strcmp(argv[i], "--help")) {
            usage(argv[0], false);
            return 0;
        }
        if (flag_value(i, argc, argv, 'i', in) || flag_value(i, argc, argv, 'o', out)) continue;
        die(std::string("Flag could not be matched: ") + argv[i]);
    }
    if (in.empty()) die("Check failed: input path (-i) is required");
    if (out.empty()) die("Check failed: output path (-o) is required");
    std::vector<Edge> edges = read_text_edges(in);
    std::unordered_map<uint32_t, uint32_t> name2id;
    name2id.reserve(edges.size() + 1);
    auto id_of = [&](uint32_t name) {
        auto it = name2id.find(name);
        if (it != name2id.end()) return it->second;
        const uint32_t id = (uint32_t)name2id.size();
        name2id.emplace(name, id);
        return id;
    };
    for (Edge &e : edges) {
        e.src = id_of(e.src);
        e.dst = id_of(e.dst);
    }
    // the header lines of the reference tool, its spelling included
    std::string header = "# Converted from: " + in + "\n# vertex number: " + std::to_string(name2id.size()) +
                         "\n# edegs: " + std::to_string(edges.size()) + "\n";
    write_binary(out, edges, header);
    return 0;
}

int format_yahoo(int argc, char **argv)
{
    std::vector<std::string> inputs;
    std::string bin, text, v;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "-h") || !strcmp(argv[i], "--help")) {
            usage(argv[0], true);
            return 0;
        }
        if (flag_value(i, argc, argv, 'i', v)) {
            inputs.push_back(v);
            continue;
        }
        if (flag_value(i, argc, argv, 'b', bin) || flag_value(i, argc, argv, 't', text)) continue;
        die(std::string("Flag could not be matched: ") + argv[i]);
    }
    if (inputs.empty()) die("Check failed: at least one input path (-i) is required");
    if (bin.empty() && text.empty()) die("Check failed: binary_output_path_flag || text_output_path_flag");
    std::vector<Edge> edges;
    for (const std::string &path : inputs) {
        const std::string s = read_all(path);
        const char *p = s.data(), *end = s.data() + s.size();
        uint32_t src, degree;
        while (parse_u32(p, end, src) && parse_u32(p, end, degree)) {
            for (uint32_t k = 0; k < degree; ++k) {
                uint32_t dst;
                if (!parse_u32(p, end, dst))
                    die("Check failed: 1 == fscanf(f, \"%u\", &dst) (" + path + ": vertex " + std::to_string(src) +
                        " lists fewer than " + std::to_string(degree) + " neighbours)");
                edges.push_back(Edge{src, dst});
            }
        }
    }
    std::string header = "# Converted from:";
    for (const std::string &path : inputs) header += " " + path;
    header += "\n# Edges: " + std::to_string(edges.size()) + "\n";
    if (!bin.empty()) write_binary(bin, edges, header);
    if (!text.empty()) write_text(text, edges, header);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    const char *slash = strrchr(argv[0], '/');
    const std::string name = slash ? slash + 1 : argv[0];
    if (name == "format_yahoo") return format_yahoo(argc, argv);
    if (name == "format_knk") return format_knk(argc, argv);
    fprintf(stderr, "run this program as format_knk or format_yahoo\n");
    return 1;
}
