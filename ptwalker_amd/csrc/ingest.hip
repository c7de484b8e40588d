// Graph ingest on the device: the edge-list file is copied to HBM, split
// into lines, parsed, compacted to first-appearance ids and turned into a
// CSR there; only the finished CSR comes back to the host.
//
// The result is the host loader's, bit for bit (load_edge_list in
// graph_host.cpp, which follows Graph::load, src/core/graph.hpp:293-366, and
// read_text_graph / read_binary_graph, include/io.hpp:39-53,76-90):
//   text   lines end at '\n'; an empty line or one starting with '#' is
//          skipped; a line holds an edge if it starts with two unsigned
//          decimal numbers (blanks, an optional '+' before each; values
//          saturate at 2^32 - 1), anything after them is ignored
//   binary u32 (src, dst) pairs; a trailing partial pair is dropped
//   ids    assigned in order of first appearance over src0, dst0, src1, ...
//   CSR    adjacency in edge order; with as_undirected every reversed edge
//          follows all forward edges
//
// Steps:
//   1. newlines per 64 KiB tile, an exclusive scan, and the newline
//      positions (u64: files may exceed 4 GiB)
//   2. one thread per line parses it into a (src, dst) pair and a flag;
//      an order-preserving select keeps the edges
//   3. a stable radix sort of (raw id, position) over all 2m endpoints: the
//      first element of each run holds the id's first position; sorting the
//      runs by that position gives the new ids, scattered back to positions
//   4. a stable radix sort of the edges by new source id gives the adjacency
//      in edge order, and row_ptr follows from the sorted sources
//   5. (with the relabel) the degree-descending device order: a stable sort
//      of the vertices by complemented degree (ties keep graph-id order), a
//      scan of the degrees in that order, and one pass over the edges that
//      writes each neighbour's device id to its place in the device CSR
// The scans, sorts and select are prims.hip's.  The graph-id CSR and the
// relabel come back to the host as its mirror; the device CSR stays.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ptw.h"
#include "engine.hpp"
#include "prims.hpp"

namespace ptw {

namespace {

constexpr uint64_t kTile = 65536;  // bytes per newline-count tile
constexpr int kTileThreads = 256;  // each thread scans kTile / kTileThreads = 256 bytes

// High bit of each byte of w that is a newline (exact, carry-free: see
// newlines_in).
__device__ __forceinline__ uint32_t newline_mask(uint32_t w)
{
    const uint32_t y = w ^ 0x0A0A0A0Au;
    const uint32_t t = (y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
    return ~(t | y | 0x7F7F7F7Fu);
}

__device__ __forceinline__ uint32_t newlines_in(uint4 x)
{
    uint32_t n = 0;
    const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        // bytes equal to 0x0A are the zero bytes of y = w ^ 0x0A0A0A0A; the
        // carry-free form flags exactly those (the shorter borrow form can
        // also flag a 0x01 byte above a zero byte)
        const uint32_t y = w[k] ^ 0x0A0A0A0Au;
        const uint32_t t = (y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
        n += __popc(~(t | y | 0x7F7F7F7Fu));
    }
    return n;
}

// Bytes are read as uint4 from a buffer padded with zeros to a multiple of
// kTile, so every tile is whole.
__global__ __launch_bounds__(kTileThreads) void count_newlines_kernel(const uint4 *data, uint64_t tiles,
                                                                     uint64_t *tile_count)
{
    __shared__ uint32_t s_sum;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        if (threadIdx.x == 0) s_sum = 0;
        __syncthreads();
        const uint4 *p = data + t * (kTile / 16) + (uint64_t)threadIdx.x * (kTile / 16 / kTileThreads);
        uint32_t n = 0;
        for (uint32_t k = 0; k < kTile / 16 / kTileThreads; ++k) n += newlines_in(p[k]);
        atomicAdd(&s_sum, n);
        __syncthreads();
        if (threadIdx.x == 0) tile_count[t] = s_sum;
        __syncthreads();
    }
}

// Positions of the newlines, in file order: tile base + the thread's
// exclusive prefix inside the tile + order within the thread's bytes.
__global__ __launch_bounds__(kTileThreads) void newline_positions_kernel(const uint4 *data, uint64_t tiles,
                                                                        const uint64_t *tile_base, uint64_t *pos)
{
    __shared__ uint32_t s_count[kTileThreads];
    constexpr uint32_t kWords = kTile / 16 / kTileThreads;  // 16 B words per thread
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t first = t * kTile + (uint64_t)threadIdx.x * (kTile / kTileThreads);
        const uint4 *p = data + first / 16;
        uint4 w[kWords];  // the thread's bytes stay in registers for the second pass
        uint32_t n = 0;
#pragma unroll
        for (uint32_t k = 0; k < kWords; ++k) {
            w[k] = p[k];
            n += newlines_in(w[k]);
        }
        s_count[threadIdx.x] = n;
        __syncthreads();
        for (uint32_t off = 1; off < kTileThreads; off <<= 1) {  // inclusive scan
            const uint32_t x = threadIdx.x >= off ? s_count[threadIdx.x - off] : 0u;
            __syncthreads();
            s_count[threadIdx.x] += x;
            __syncthreads();
        }
        uint64_t out = tile_base[t] + s_count[threadIdx.x] - n;
#pragma unroll
        for (uint32_t k = 0; k < kWords; ++k) {
            const uint32_t c[4] = {w[k].x, w[k].y, w[k].z, w[k].w};
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                uint32_t m = newline_mask(c[j]);
                while (m) {  // byte b of the word is a newline when bit 8b + 7 is set
                    const uint32_t b = (uint32_t)(__ffs(m) - 1) >> 3;
                    pos[out++] = first + 16 * k + 4 * j + b;
                    m &= m - 1;
                }
            }
        }
        __syncthreads();
    }
}

__device__ __forceinline__ bool is_blank(unsigned char c)
{
    return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f';
}

__device__ bool parse_pair_dev(const unsigned char *p, const unsigned char *end, uint32_t &a, uint32_t &b)
{
    uint64_t v[2];
    for (int k = 0; k < 2; ++k) {
        while (p < end && is_blank(*p)) ++p;
        if (p < end && *p == '+') ++p;
        if (p >= end || *p < '0' || *p > '9') return false;
        uint64_t x = 0;
        while (p < end && *p >= '0' && *p <= '9') {
            x = x * 10 + (uint64_t)(*p - '0');
            if (x > 0xFFFFFFFFull) x = 0xFFFFFFFFull;
            ++p;
        }
        v[k] = x;
    }
    a = (uint32_t)v[0];
    b = (uint32_t)v[1];
    return true;
}

// Line i runs from just after newline i - 1 (or the file start) to newline
// i (or the file end).
__global__ void parse_lines_kernel(const unsigned char *file, uint64_t size, const uint64_t *nl, uint64_t n_nl,
                                   uint64_t lines, uint2 *pairs, uint8_t *flags)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < lines; i += stride) {
        const uint64_t s = i == 0 ? 0 : nl[i - 1] + 1;
        const uint64_t e = i < n_nl ? nl[i] : size;
        uint32_t a = 0, b = 0;
        const bool ok = e > s && file[s] != '#' && parse_pair_dev(file + s, file + e, a, b);
        pairs[i] = uint2{a, b};
        flags[i] = ok ? 1 : 0;
    }
}

__global__ void endpoints_kernel(const uint2 *pairs, uint64_t m, uint32_t *keys, uint32_t *vals)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const uint2 e = pairs[i];
        keys[2 * i] = e.x;
        keys[2 * i + 1] = e.y;
        vals[2 * i] = (uint32_t)(2 * i);
        vals[2 * i + 1] = (uint32_t)(2 * i + 1);
    }
}

__global__ void run_heads_kernel(const uint32_t *skeys, uint64_t n, uint32_t *head)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
        head[k] = (k == 0 || skeys[k] != skeys[k - 1]) ? 1u : 0u;
}

// run_of = exclusive scan of head: a head's run index; first_pos[run] is the
// head's position (the smallest of its run: the sort is stable and positions
// were ascending) and run_key[run] the raw id the run holds.
__global__ void first_positions_kernel(const uint32_t *head, const uint32_t *run_excl, const uint32_t *skeys,
                                       const uint32_t *svals, uint64_t n, uint32_t *first_pos, uint32_t *run_idx,
                                       uint32_t *run_key)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        if (head[k]) {
            const uint32_t r = run_excl[k];
            first_pos[r] = svals[k];
            run_idx[r] = r;
            run_key[r] = skeys[k];
        }
    }
}

// order[rank] is the run with the rank-th first appearance: that run's new id
// is rank, and the raw id it holds is the new id's name.
__global__ void rank_runs_kernel(const uint32_t *order, const uint32_t *run_key, uint32_t runs, uint32_t *new_id,
                                 uint32_t *names)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < runs; r += stride) {
        new_id[order[r]] = (uint32_t)r;
        names[r] = run_key[order[r]];
    }
}

// Every endpoint position gets its new id: the run of sorted element k is
// head-inclusive-scan - 1, i.e. run_excl[k] + head[k] - 1.
__global__ void scatter_ids_kernel(const uint32_t *head, const uint32_t *run_excl, const uint32_t *svals,
                                   const uint32_t *new_id, uint64_t n, uint32_t *mapped)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
        mapped[svals[k]] = new_id[run_excl[k] + head[k] - 1];
}

__global__ void edges_kernel(const uint32_t *mapped, uint64_t m, int undirected, uint32_t *esrc, uint32_t *edst)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const uint32_t s = mapped[2 * i], d = mapped[2 * i + 1];
        esrc[i] = s;
        edst[i] = d;
        if (undirected) {
            esrc[m + i] = d;
            edst[m + i] = s;
        }
    }
}

// Row offsets from the sources sorted ascending: row[v] is the first index
// whose source is >= v, row[V] = E.  Index i writes row[v] for the vertices
// v in (ssrc[i - 1], ssrc[i]] (from 0 at i = 0, up to V at i = E), so every
// vertex is written once, those without out-edges included.
__global__ void row_offsets_kernel(const uint32_t *ssrc, uint64_t E, uint32_t V, uint32_t *row)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= E; i += stride) {
        const uint64_t lo = i == 0 ? 0 : (uint64_t)ssrc[i - 1] + 1;
        const uint64_t hi = i == E ? V : ssrc[i];
        for (uint64_t v = lo; v <= hi; ++v) row[v] = (uint32_t)i;
    }
}

// Relabel, step 5: key = ~degree (ascending = degree descending), value =
// graph id.
__global__ void degree_keys_kernel(const uint32_t *row, uint32_t V, uint32_t *key, uint32_t *val)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        key[v] = ~(row[v + 1] - row[v]);
        val[v] = (uint32_t)v;
    }
}

// inv[d] is the graph id at device position d: perm is its inverse and
// ddeg[d] its degree (scanned into the device row offsets).
__global__ void permutation_kernel(const uint32_t *inv, const uint32_t *row, uint32_t V, uint32_t *perm,
                                   uint32_t *ddeg)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; d < V; d += stride) {
        const uint32_t u = inv[d];
        perm[u] = (uint32_t)d;
        ddeg[d] = row[u + 1] - row[u];
    }
}

__global__ void set_word_kernel(uint32_t *p, uint32_t v) { *p = v; }

// Edge i of source u (graph ids, sorted by source) is entry i - row[u] of
// u's row; in the device CSR that row starts at drow[perm[u]] and holds
// device ids.
__global__ void device_col_kernel(const uint32_t *ssrc, const uint32_t *col, const uint32_t *row,
                                  const uint32_t *perm, const uint32_t *drow, uint64_t E, uint32_t *dcol)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += stride) {
        const uint32_t u = ssrc[i];
        dcol[(uint64_t)drow[perm[u]] + (i - row[u])] = perm[col[i]];
    }
}

uint32_t grid_for(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 65536));
}

// Device buffers released on every exit path.
struct Buffers {
    std::vector<void *> p;
    ~Buffers()
    {
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
    template <typename T>
    hipError_t get(T **out, uint64_t count)
    {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<uint64_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) p.push_back(q);
        *out = (T *)q;
        return e;
    }
    // hands a buffer to the caller: it is no longer freed here
    void release(void *q)
    {
        for (void *&x : p)
            if (x == q) x = nullptr;
    }
};

struct Pinned {
    void *p = nullptr;
    ~Pinned()
    {
        if (p) (void)hipHostFree(p);
    }
};

struct Fd {
    int fd = -1;
    ~Fd()
    {
        if (fd >= 0) close(fd);
    }
};

#define ING_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            err = std::string("device ingest: ") + #expr + ": " + hipGetErrorString(e_); \
            return e_ == hipErrorOutOfMemory ? PTW_ERR_OOM : PTW_ERR_HIP;          \
        }                                                                          \
    } while (0)

// Temporary storage of the primitives, grown and kept for the whole load.
struct Temp {
    Buffers &B;
    hipStream_t s;
    uint8_t *p = nullptr;
    size_t cap = 0;
    hipError_t need(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        // an earlier primitive may still be using the buffer
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        if (p) {
            (void)hipFree(p);
            B.release(p);
            p = nullptr;
            cap = 0;
        }
        if ((e = B.get(&p, bytes)) == hipSuccess) cap = bytes;
        return e;
    }
};

// Stable radix sort of (key, value) pairs over `bits` key bits.
int sort_pairs(Temp &T, const uint32_t *kin, uint32_t *kout, const uint32_t *vin, uint32_t *vout, uint64_t n, int bits,
               hipStream_t s, std::string &err)
{
    ING_TRY(T.need(prims::sort_temp_bytes(n)));
    ING_TRY(prims::sort_pairs(kin, kout, vin, vout, n, bits, T.p, s));
    return PTW_OK;
}

template <typename X, typename U>
int exclusive_sum(Temp &T, const X *in, U *out, uint64_t n, hipStream_t s, std::string &err)
{
    ING_TRY(T.need(prims::scan_temp_bytes(n, sizeof(U))));
    ING_TRY(prims::exclusive_scan(in, out, n, T.p, s));
    return PTW_OK;
}

int bits_for(uint64_t max_value)
{
    int b = 1;
    while (b < 32 && (max_value >> b) != 0) ++b;
    return b;
}

}  // namespace

int device_load_edge_list(hipStream_t s, const char *path, int format, bool as_undirected, bool relabel,
                          HostCsr &g, DeviceBuiltGraph &built, std::string &err)
{
    if (!path) {
        err = "graph path is NULL";
        return PTW_ERR_INVALID_ARG;
    }
    if (format != PTW_FORMAT_TEXT && format != PTW_FORMAT_BINARY) {
        err = "unknown graph format " + std::to_string(format);
        return PTW_ERR_INVALID_ARG;
    }
    Fd f;
    f.fd = open(path, O_RDONLY);
    if (f.fd < 0) {
        err = std::string("cannot open graph file '") + path + "': " + strerror(errno);
        return PTW_ERR_IO;
    }
    struct stat st;
    if (fstat(f.fd, &st) != 0) {
        err = std::string("cannot stat '") + path + "'";
        return PTW_ERR_IO;
    }
    const uint64_t size = (uint64_t)st.st_size;
    Buffers B;
    Temp T{B, s};
    int rc;
    // ---- 1. the file, padded with zeros to whole tiles ----
    const uint64_t tiles = (size + kTile - 1) / kTile;
    const uint64_t padded = std::max<uint64_t>(tiles, 1) * kTile;
    unsigned char *file = nullptr;
    ING_TRY(B.get(&file, padded));
    ING_TRY(hipMemsetAsync(file + (padded - kTile), 0, kTile, s));
    {
        const uint64_t chunk = 64ull << 20;
        Pinned stage;
        ING_TRY(hipHostMalloc(&stage.p, 2 * chunk, hipHostMallocDefault));
        hipEvent_t done[2];
        ING_TRY(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
        ING_TRY(hipEventCreateWithFlags(&done[1], hipEventDisableTiming));
        struct Ev {
            hipEvent_t *e;
            ~Ev()
            {
                (void)hipEventDestroy(e[0]);
                (void)hipEventDestroy(e[1]);
            }
        } ev{done};
        ING_TRY(hipEventRecord(done[0], s));
        ING_TRY(hipEventRecord(done[1], s));
        uint64_t off = 0;
        for (int slot = 0; off < size; slot ^= 1) {
            ING_TRY(hipEventSynchronize(done[slot]));  // the slot's previous copy has left
            char *buf = (char *)stage.p + slot * chunk;
            const uint64_t want = std::min(chunk, size - off);
            // one thread copies a page-cache file at a few GB/s: the chunk is
            // read as 4 MiB pieces on the host threads
            const uint64_t piece = 4ull << 20;
            std::atomic<bool> short_read{false};
            parallel_for(
                (want + piece - 1) / piece,
                [&](uint64_t lo, uint64_t hi) {
                    for (uint64_t k = lo; k < hi && !short_read.load(std::memory_order_relaxed); ++k) {
                        const uint64_t a = k * piece, b = std::min(want, a + piece);
                        for (uint64_t got = a; got < b;) {
                            const ssize_t r = pread(f.fd, buf + got, b - got, (off_t)(off + got));
                            if (r <= 0) {
                                short_read.store(true, std::memory_order_relaxed);
                                return;
                            }
                            got += (uint64_t)r;
                        }
                    }
                },
                0, 1);
            if (short_read.load()) {
                err = std::string("short read from '") + path + "'";
                (void)hipStreamSynchronize(s);
                return PTW_ERR_IO;
            }
            ING_TRY(hipMemcpyAsync(file + off, buf, want, hipMemcpyHostToDevice, s));
            ING_TRY(hipEventRecord(done[slot], s));
            off += want;
        }
        ING_TRY(hipStreamSynchronize(s));
        // the tail of the last tile beyond the file must read as zeros
        if (size % kTile) ING_TRY(hipMemsetAsync(file + size, 0, padded - size, s));
    }
    // ---- 2. edges ----
    uint2 *pairs = nullptr;
    uint64_t m = 0;
    if (format == PTW_FORMAT_BINARY) {
        m = size / 8;  // a trailing partial record is dropped
        pairs = reinterpret_cast<uint2 *>(file);
    } else {
        uint64_t *tile_count = nullptr;  // u64 so that the scan accumulates in 64 bits
        uint64_t *tile_base = nullptr;
        ING_TRY(B.get(&tile_count, tiles));
        ING_TRY(B.get(&tile_base, tiles));
        if (tiles) {
            hipLaunchKernelGGL(count_newlines_kernel, dim3(grid_for(tiles * 256)), dim3(kTileThreads), 0, s,
                               reinterpret_cast<const uint4 *>(file), tiles, tile_count);
            ING_TRY(hipGetLastError());
            if ((rc = exclusive_sum(T, tile_count, tile_base, tiles, s, err))) return rc;
        }
        uint64_t last_base = 0, last_count = 0;
        if (tiles) {
            ING_TRY(hipMemcpyAsync(&last_base, tile_base + tiles - 1, 8, hipMemcpyDeviceToHost, s));
            ING_TRY(hipMemcpyAsync(&last_count, tile_count + tiles - 1, 8, hipMemcpyDeviceToHost, s));
        }
        unsigned char last_byte = '\n';
        if (size) ING_TRY(hipMemcpyAsync(&last_byte, file + size - 1, 1, hipMemcpyDeviceToHost, s));
        ING_TRY(hipStreamSynchronize(s));
        const uint64_t n_nl = last_base + last_count;
        const uint64_t lines = n_nl + (size && last_byte != '\n' ? 1 : 0);
        uint64_t *nl = nullptr;
        ING_TRY(B.get(&nl, n_nl));
        if (n_nl) {
            hipLaunchKernelGGL(newline_positions_kernel, dim3(grid_for(tiles * 256)), dim3(kTileThreads), 0, s,
                               reinterpret_cast<const uint4 *>(file), tiles, tile_base, nl);
            ING_TRY(hipGetLastError());
        }
        uint2 *all = nullptr;
        uint8_t *flags = nullptr;
        uint64_t *n_sel = nullptr;
        ING_TRY(B.get(&all, lines));
        ING_TRY(B.get(&flags, lines));
        ING_TRY(B.get(&pairs, lines));
        ING_TRY(B.get(&n_sel, 1));
        if (lines) {
            hipLaunchKernelGGL(parse_lines_kernel, dim3(grid_for(lines)), dim3(256), 0, s, file, size, nl, n_nl,
                               lines, all, flags);
            ING_TRY(hipGetLastError());
            ING_TRY(T.need(prims::select_temp_bytes(lines)));
            ING_TRY(prims::select_flagged(all, flags, pairs, lines, n_sel, T.p, s));
            ING_TRY(hipMemcpyAsync(&m, n_sel, 8, hipMemcpyDeviceToHost, s));
            ING_TRY(hipStreamSynchronize(s));
        }
    }
    const uint64_t total = as_undirected ? 2 * m : m;
    if (total >= 0xFFFFFFFFull || 2 * m >= 0xFFFFFFFFull) {
        err = "graph has " + std::to_string(total) + " edges; u32 offsets hold at most 2^32 - 2";
        return PTW_ERR_UNSUPPORTED;
    }
    if (m == 0) {
        g = HostCsr{};
        return PTW_OK;
    }
    // ---- 3. first-appearance ids ----
    const uint64_t n = 2 * m;
    uint32_t *keys = nullptr, *vals = nullptr, *skeys = nullptr, *svals = nullptr;
    ING_TRY(B.get(&keys, n));
    ING_TRY(B.get(&vals, n));
    ING_TRY(B.get(&skeys, n));
    ING_TRY(B.get(&svals, n));
    // the id sort needs only the bits of the largest raw id
    uint32_t *dmax = nullptr, max_id = 0;
    ING_TRY(B.get(&dmax, 1));
    ING_TRY(prims::max_u32(pairs, m, dmax, s));
    ING_TRY(hipMemcpyAsync(&max_id, dmax, 4, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(endpoints_kernel, dim3(grid_for(m)), dim3(256), 0, s, pairs, m, keys, vals);
    ING_TRY(hipGetLastError());
    ING_TRY(hipStreamSynchronize(s));
    // one allocation covers every later primitive: the sorts are the largest
    ING_TRY(T.need(std::max(prims::sort_temp_bytes(n), prims::scan_temp_bytes(n, 4))));
    if ((rc = sort_pairs(T, keys, skeys, vals, svals, n, bits_for(max_id), s, err))) return rc;
    uint32_t *head = keys, *run_excl = vals;  // reuse
    hipLaunchKernelGGL(run_heads_kernel, dim3(grid_for(n)), dim3(256), 0, s, skeys, n, head);
    ING_TRY(hipGetLastError());
    if ((rc = exclusive_sum(T, head, run_excl, n, s, err))) return rc;
    uint32_t last_excl = 0, last_head = 0;
    ING_TRY(hipMemcpyAsync(&last_excl, run_excl + n - 1, 4, hipMemcpyDeviceToHost, s));
    ING_TRY(hipMemcpyAsync(&last_head, head + n - 1, 4, hipMemcpyDeviceToHost, s));
    ING_TRY(hipStreamSynchronize(s));
    const uint32_t V = last_excl + last_head;
    uint32_t *first_pos = nullptr, *run_idx = nullptr, *sfirst = nullptr, *order = nullptr, *new_id = nullptr;
    uint32_t *run_key = nullptr, *names = nullptr;
    ING_TRY(B.get(&first_pos, V));
    ING_TRY(B.get(&run_idx, V));
    ING_TRY(B.get(&sfirst, V));
    ING_TRY(B.get(&order, V));
    ING_TRY(B.get(&new_id, V));
    ING_TRY(B.get(&run_key, V));
    ING_TRY(B.get(&names, V));
    hipLaunchKernelGGL(first_positions_kernel, dim3(grid_for(n)), dim3(256), 0, s, head, run_excl, skeys, svals, n,
                       first_pos, run_idx, run_key);
    ING_TRY(hipGetLastError());
    if ((rc = sort_pairs(T, first_pos, sfirst, run_idx, order, V, bits_for(n - 1), s, err))) return rc;
    hipLaunchKernelGGL(rank_runs_kernel, dim3(grid_for(V)), dim3(256), 0, s, order, run_key, V, new_id, names);
    ING_TRY(hipGetLastError());
    uint32_t *mapped = skeys;  // reuse: position -> new id
    hipLaunchKernelGGL(scatter_ids_kernel, dim3(grid_for(n)), dim3(256), 0, s, head, run_excl, svals, new_id, n,
                       mapped);
    ING_TRY(hipGetLastError());
    // ---- 4. CSR ----
    const uint64_t E = total;
    uint32_t *esrc = nullptr, *edst = nullptr, *ssrc = nullptr, *col = nullptr, *row = nullptr;
    ING_TRY(B.get(&esrc, E));
    ING_TRY(B.get(&edst, E));
    ING_TRY(B.get(&ssrc, E));
    ING_TRY(B.get(&col, E));
    ING_TRY(B.get(&row, (uint64_t)V + 1));
    hipLaunchKernelGGL(edges_kernel, dim3(grid_for(m)), dim3(256), 0, s, mapped, m, as_undirected ? 1 : 0, esrc,
                       edst);
    ING_TRY(hipGetLastError());
    if ((rc = sort_pairs(T, esrc, ssrc, edst, col, E, bits_for(V - 1), s, err))) return rc;
    hipLaunchKernelGGL(row_offsets_kernel, dim3(grid_for(E + 1)), dim3(256), 0, s, ssrc, E, V, row);
    ING_TRY(hipGetLastError());
    // ---- 5. device order ----
    uint32_t *perm = nullptr, *inv = nullptr, *drow = nullptr, *dcol = nullptr;
    if (relabel) {
        // the endpoint arrays (2m >= V entries) are free after the CSR sort
        uint32_t *dkey = keys, *dval = vals, *skey = svals;
        ING_TRY(B.get(&inv, V));
        ING_TRY(B.get(&perm, V));
        ING_TRY(B.get(&drow, (uint64_t)V + 1));
        ING_TRY(B.get(&dcol, E));
        hipLaunchKernelGGL(degree_keys_kernel, dim3(grid_for(V)), dim3(256), 0, s, row, V, dkey, dval);
        ING_TRY(hipGetLastError());
        if ((rc = sort_pairs(T, dkey, skey, dval, inv, V, 32, s, err))) return rc;
        uint32_t *ddeg = dkey;
        hipLaunchKernelGGL(permutation_kernel, dim3(grid_for(V)), dim3(256), 0, s, inv, row, V, perm, ddeg);
        ING_TRY(hipGetLastError());
        if ((rc = exclusive_sum(T, ddeg, drow, V, s, err))) return rc;
        hipLaunchKernelGGL(set_word_kernel, dim3(1), dim3(1), 0, s, drow + V, (uint32_t)E);
        ING_TRY(hipGetLastError());
        hipLaunchKernelGGL(device_col_kernel, dim3(grid_for(E)), dim3(256), 0, s, ssrc, col, row, perm, drow, E,
                           dcol);
        ING_TRY(hipGetLastError());
    }
    // ---- the host mirror ----
    g.V = V;
    g.row_ptr.resize((size_t)V + 1);
    g.col.resize(E);
    g.names.resize(V);
    ING_TRY(hipMemcpyAsync(g.row_ptr.data(), row, ((uint64_t)V + 1) * 4, hipMemcpyDeviceToHost, s));
    ING_TRY(hipMemcpyAsync(g.col.data(), col, E * 4, hipMemcpyDeviceToHost, s));
    ING_TRY(hipMemcpyAsync(g.names.data(), names, (uint64_t)V * 4, hipMemcpyDeviceToHost, s));
    built = DeviceBuiltGraph{};
    if (relabel) {
        built.h_perm.resize(V);
        built.h_inv.resize(V);
        built.h_drow.resize((size_t)V + 1);
        ING_TRY(hipMemcpyAsync(built.h_perm.data(), perm, (uint64_t)V * 4, hipMemcpyDeviceToHost, s));
        ING_TRY(hipMemcpyAsync(built.h_inv.data(), inv, (uint64_t)V * 4, hipMemcpyDeviceToHost, s));
        ING_TRY(hipMemcpyAsync(built.h_drow.data(), drow, ((uint64_t)V + 1) * 4, hipMemcpyDeviceToHost, s));
    }
    ING_TRY(hipStreamSynchronize(s));
    // the device graph leaves with the caller
    built.relabelled = relabel;
    built.row_ptr = relabel ? drow : row;
    built.col = relabel ? dcol : col;
    built.perm = perm;
    built.inv = inv;
    for (void *q : {(void *)built.row_ptr, (void *)built.col, (void *)perm, (void *)inv})
        if (q) B.release(q);
    return PTW_OK;
}

}  // namespace ptw
