// Internal declarations shared by the host graph builder, the C ABI and the
// kernel launchers.  Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <functional>
#include <string>
#include <vector>

namespace ptw {

// CSR in graph ids, host resident.  Adjacency order is significant: the
// neighbour pick indexes into it.
struct HostCsr {
    uint32_t V = 0;
    std::vector<uint32_t> row_ptr;  // V + 1
    std::vector<uint32_t> col;      // E
    // the id each vertex had before compaction (the file's id for a loaded
    // edge list, the generator's for R-MAT); empty: graph ids are those ids
    std::vector<uint32_t> names;
    uint64_t E() const { return col.size(); }
    uint32_t name(uint32_t v) const { return names.empty() ? v : names[v]; }
};

// Device copy of the graph under the degree-descending relabel.
struct DeviceGraph {
    uint32_t V = 0;
    uint64_t E = 0;
    uint32_t *row_ptr = nullptr;  // V + 1, device ids
    uint32_t *col = nullptr;      // E, device ids, per-vertex order kept
    uint32_t *perm = nullptr;     // graph id -> device id (nullptr: identity)
    uint32_t *inv = nullptr;      // device id -> graph id (nullptr: identity)
    // E weighted-edge records, device order: {neighbour, alias threshold,
    // the alias slot's neighbour, 0}.  A weighted pick of slot j reads one
    // 16 B record instead of the alias entry and then the neighbour.
    uint4 *wedge = nullptr;
    // neighbour records, built on first use: E x {neighbour, its row start,
    // its degree, its graph id}, and 2E weighted ones for path walks (both
    // alias candidates of a slot: row start, degree, graph id; the
    // threshold in the first)
    uint4 *nrec = nullptr;
    uint4 *wnrec = nullptr;
    // transpose (in-edges, cut into pieces) and out-degrees for the FP64
    // solver, built on first use
    uint32_t *t_col = nullptr, *out_deg = nullptr, *piece_begin = nullptr, *vertex_piece = nullptr;
    // node2vec, built on first use: the set of edges (u << 32 | x, device
    // ids) in an open-addressing table of eset_mask + 1 >= 2E slots, and
    // whether the graph is simple and symmetric (no row repeats a neighbour,
    // every edge has its reverse), so a walker's previous vertex is in its
    // current row exactly once
    unsigned long long *eset = nullptr;
    uint64_t eset_mask = 0;
    bool simple_symmetric = false;
    uint64_t pieces = 0;
    uint32_t hubs = 0;            // device ids [0, hubs) are the top-degree vertices
    bool relabelled = false;
    uint64_t bytes = 0;
};

// Host-side parallel loop over [0, n) in contiguous chunks of at least
// min_per_thread items.
void parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)> &body,
                  unsigned threads = 0, uint64_t min_per_thread = 4096);
unsigned host_threads();

// Edge-list ingest with the reference's load semantics.  Returns 0 or a
// PTW_ERR_* code with `err` filled.
int load_edge_list(const char *path, int format, bool as_undirected, HostCsr &g, std::string &err);
int write_edge_list(const HostCsr &g, const char *path, int format, std::string &err);
int build_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
               bool keep_isolated, HostCsr &g, std::string &err);
int validate_csr(const uint32_t *row_ptr, const uint32_t *col, uint32_t V, uint64_t E,
                 std::string &err);

// Edge weights -> per-vertex integer alias tables (weighted neighbour choice).
// Weights are quantized per vertex to q = clamp(floor(w / w_max * 2^24 + 0.5),
// 1, 2^24); the table is built by Vose's method in exact integer arithmetic
// with LIFO small/large stacks filled in index order:
//   slot s (mass m_s = q_s * deg < Q = sum q): thr = floor(m_s 2^32 / Q), alias = l
//   slot left full:                             thr = 2^32 - 1,          alias = self
// A pick j then takes neighbour j if coin < thr[j], else alias[j].
int build_alias_tables(const HostCsr &g, const std::vector<float> &w, std::vector<uint32_t> &thr,
                       std::vector<uint32_t> &alias, std::string &err);
// U(0, 1] weights on 2^-24 steps from Philox(seed, edge index).
void random_weights(uint64_t E, uint64_t seed, std::vector<float> &w);

// Degree-descending relabel (stable by graph id).  perm[g] = d, inv[d] = g;
// d_row_ptr / d_col are the relabelled CSR.
void relabel_by_degree(const HostCsr &g, std::vector<uint32_t> &perm, std::vector<uint32_t> &inv,
                       std::vector<uint32_t> &d_row_ptr, std::vector<uint32_t> &d_col);

// ---- kernel launchers (walk.hip) -------------------------------------------

struct WalkLaunch {
    // graph
    const uint32_t *row_ptr;
    const uint32_t *col;
    const uint32_t *perm;  // nullptr: identity
    const uint32_t *inv;   // nullptr: identity
    const uint4 *wedge;    // weighted-edge records; nullptr: uniform neighbour choice
    // neighbour records (edge_records_kernel / weighted_records_kernel);
    // nullptr: the walk reads row_ptr, col (or wedge) and inv per step
    const uint4 *nrec;
    const uint4 *wnrec;    // weighted (path walks only; used instead of wedge)
    uint32_t V;
    uint32_t hubs;         // LDS-privatized counters for device ids < hubs
    // process
    int app;
    int semantics;
    int start_mode;
    uint32_t source_dev;   // device id of the source
    uint32_t walk_len;
    uint32_t thr;          // continue iff r.z < thr
    uint32_t seed_lo, seed_hi, stream;
    uint64_t wid_base;
    uint64_t walkers;
    // node2vec: sorted adjacency and the integer weights of a neighbour that
    // is the previous vertex, a neighbour of it, or neither.  w_in and w_out
    // are acceptance thresholds in [0, 2^32].  Unfolded, w_ret is one too.
    // Folded (n2v_fold > 0 extra proposal slots, each accepting the previous
    // vertex below w_fold), w_ret = 2^32 + n2v_fold * w_fold.
    const unsigned long long *eset;  // edge set: is x a neighbour of prev
    uint64_t eset_mask;
    uint64_t w_ret, w_in, w_out, w_fold;
    uint32_t n2v_fold;
    // outputs (device)
    unsigned long long *counts_dev;  // device-id counters (count apps)
    uint32_t *ends;                  // graph ids
    uint32_t *starts;                // graph ids (end apps; nullptr: not written)
    uint32_t *paths;                 // graph ids
    // visit log (count apps; nullptr: device atomics): vlog_cap entries (a
    // multiple of kVlogSeg), claimed in segments from scratch[3]; every
    // appended entry is counted in vlog_hist[id >> vlog_shift]
    uint32_t *vlog;
    uint64_t vlog_cap;
    uint32_t *vlog_hist;
    uint32_t vlog_shift;
    uint32_t vlog_buckets;
    // scratch (device): [0] queue head, [1] steps, [2] dangling, [3] log head
    unsigned long long *scratch;
};

// Visit log: entries a wavefront claims at once, the filler of unused
// entries, and the id buckets its partition uses (ids >> shift, buckets of
// 2^13..2^15 ids, at most kVlogMaxBuckets).
constexpr uint32_t kVlogSeg = 1024;
constexpr uint32_t kNoVisit = 0xFFFFFFFFu;
constexpr uint32_t kVlogMaxBuckets = 4096;
constexpr uint32_t kVlogMinShift = 13, kVlogMaxShift = 15;
// Bucket shift for V vertices: about 512 buckets, within the shift range;
// 0 if the graph needs more than kVlogMaxBuckets buckets.
inline uint32_t vlog_shift_for(uint32_t V)
{
    uint32_t lg = 0;
    while (lg < 32 && (1ull << lg) < (uint64_t)V) ++lg;
    uint32_t sh = lg > 9 ? lg - 9 : 0;
    sh = sh < kVlogMinShift ? kVlogMinShift : sh > kVlogMaxShift ? kVlogMaxShift : sh;
    const uint64_t nb = V ? (((uint64_t)V - 1) >> sh) + 1 : 1;
    return nb <= kVlogMaxBuckets ? sh : 0;
}
inline uint32_t vlog_buckets_for(uint32_t V, uint32_t shift) { return V ? ((V - 1) >> shift) + 1 : 1; }

struct LaunchShape {
    int block = 1024;        // threads per workgroup
    int blocks_per_cu = 0;   // 0: fill the device from the occupancy query
    int chunk = 1024;        // walker ids a wave takes from the queue at once
    int lds_hubs = 77824;    // u16 LDS counters for device ids below this (count apps)
    int cache_slots = -1;    // LDS visit-cache slots per workgroup (power of two, 0 = off, -1 = automatic)
    int batch_chunk = 16384;  // walker ids per chunk of a batched launch
    int topk_tile = 2048;    // top-k LDS candidate slots
    int dense_chunk = 0;     // walker ids per workgroup chunk of a dense batched query (0: automatic)
    int visit_log = 1;       // count apps: log the visits LDS does not count (0: device atomics)
    bool cache_auto = false; // cache_slots came from the automatic setting (a walk's shape only)
    int edge_records = 1;    // neighbour records: 1 every walk, -1 path walks only, 0 never
};

// The launch a walk actually ran with, after the automatic rules.
struct EffectiveShape {
    int block = 0, blocks = 0;
    uint32_t chunk = 0, hubs = 0, slots = 0;
    bool vlog = false;  // the visit log was used (the caller counts it)
};

// Visit-cache slots of the automatic setting (cache_slots = -1) where a cache
// is kept, and the hub counters given the cache's LDS where it is not.
constexpr int kAutoCacheSlots = 512;
constexpr int kMaxLdsHubs = 81408;
// Cache slots of a batched dense launch under a shape.
inline uint32_t batch_cache_slots(const LaunchShape &s)
{
    return (uint32_t)(s.cache_slots < 0 ? kAutoCacheSlots : s.cache_slots);
}

// The no-allocation walk's scratch (walk.hip): 8 control words, then device-id
// u64 visit counters for graphs of at most kStaticMaxVertices vertices.
constexpr uint64_t kStaticScratchWords = 4ull << 20;  // 32 MiB
constexpr uint64_t kStaticMaxVertices = kStaticScratchWords - 8;
constexpr int kMaxStaticDevices = 64;
// Device address of the static scratch on `device`, which must be the current
// device; looked up once per device and remembered.
hipError_t static_scratch_address(int device, unsigned long long **out);

hipError_t launch_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v, hipStream_t s);
// Inserts every CSR edge into the edge set (table of mask + 1 slots, filled
// with kEdgeSetEmpty beforehand, at least twice the edges).
constexpr unsigned long long kEdgeSetEmpty = ~0ull;
hipError_t launch_edge_set_build(const uint32_t *row_ptr, const uint32_t *col, uint32_t V, uint64_t E,
                                 unsigned long long *table, uint64_t mask, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s);
hipError_t launch_edge_records(const uint32_t *row_ptr, const uint32_t *col, const uint32_t *inv, uint64_t E,
                               uint4 *rec, hipStream_t s);
hipError_t launch_weighted_records(const uint32_t *row_ptr, const uint4 *wedge, const uint32_t *inv, uint64_t E,
                                   uint4 *rec, hipStream_t s);
// `shape.cache_slots` must be resolved (>= 0) by the caller; `eff`, if given,
// receives the launch after the small-walk rule of blocks_per_cu = 0.
hipError_t launch_walk(const WalkLaunch &p, const LaunchShape &shape, int num_cus, hipStream_t s,
                       EffectiveShape *eff = nullptr);

// A graph built on the device: its CSR in device order (degree descending
// when relabelled, else graph ids), the relabel on both sides.  The device
// buffers are owned by whoever receives it.
struct DeviceBuiltGraph {
    uint32_t *row_ptr = nullptr, *col = nullptr, *perm = nullptr, *inv = nullptr;
    bool relabelled = false;
    std::vector<uint32_t> h_perm, h_inv, h_drow;
};

// Graph ingest on the device with the host loader's semantics (ingest.hip):
// `g` receives the graph-id CSR and names, `built` the device graph (with
// the degree relabel of relabel_by_degree when `relabel`).
int device_load_edge_list(hipStream_t s, const char *path, int format, bool as_undirected, bool relabel,
                          HostCsr &g, DeviceBuiltGraph &built, std::string &err);
// ---- visit log (vlog.hip) ---------------------------------------------------
// Device buffers of the visit log of one walk: the log, its partition by id
// bucket, and the partition's plan.
struct VlogBuffers {
    uint32_t *log = nullptr;     // cap entries
    uint16_t *parted = nullptr;  // cap entries: offsets inside their bucket
    uint64_t cap = 0;
    // hist[NB] (zeroed before the walk), start[NB + 1], cursor[NB],
    // task_first[NB + 1], then the segment length word
    uint32_t *plan = nullptr;
    uint32_t plan_buckets = 0;  // NB the plan area was sized for
    uint32_t *hist() const { return plan; }
};
inline uint64_t vlog_plan_words(uint32_t nb) { return 4ull * nb + 4; }
// Largest log a walk allocates (entries per buffer); a walk whose log fills
// up counts its later visits with device atomics.
constexpr uint64_t kVlogMaxEntries = 1ull << 30;
// Zeroes the bucket counts the walk kernel adds to.
hipError_t launch_vlog_begin(const VlogBuffers &b, uint32_t nb, hipStream_t s);
// After the walk kernel: partitions the log's first min(*head, cap) entries
// (head: the scratch word the walk claimed segments from; cap: the capacity
// the walk was given, <= b.cap) by bucket, then adds each bucket's counts
// into counts_dev (device ids).
hipError_t launch_vlog_count(const VlogBuffers &b, const unsigned long long *head, uint64_t cap, uint32_t shift,
                             uint32_t nb, unsigned long long *counts_dev, int num_cus, hipStream_t s);

hipError_t launch_counts_to_graph_ids(const unsigned long long *counts_dev, const uint32_t *perm,
                                      uint32_t V, uint64_t *out, bool accumulate,
                                      uint32_t extra_vertex, uint64_t extra, hipStream_t s);
// ---- batched queries (batch.hip) -------------------------------------------

// SLICE_LOG: a dense query counted without a V-sized slice.  Its visits to
// the LDS hub range go to a hub slice of hub_len counters; every other visit
// is appended to the batch's visit log in segments of kBatchLogSeg entries,
// each segment holding visits of one query.  After the walk the log is
// partitioned by (query, id bucket) and every bucket is counted in LDS
// straight into the query's (vertex, count) pair list.
enum SliceKind : uint32_t { SLICE_DENSE = 0, SLICE_HASH = 1, SLICE_LOG = 2 };

constexpr uint32_t kBatchLogSeg = 1024;     // entries per wavefront segment
constexpr uint32_t kBatchLogTile = 32;      // segments per partition tile (half where the LDS lacks room)
constexpr uint32_t kBatchLogShift = 14;     // id bucket = device id >> kBatchLogShift (batch_log_fits)
constexpr uint32_t kBatchRunIds = 2048;     // ids per run of a logged query's pair list
constexpr uint32_t kBatchLogMaxBuckets = 4096;
constexpr uint32_t kBatchLogQWords = 8;     // per-query words of the log plan (BatchLaunch::qlog)
constexpr uint32_t kBatchHdrWords = 16;     // BatchLaunch::hdr
// Workgroup LDS of the dense kernel (160 KiB less its static variables).
constexpr uint32_t kDenseLdsBytes = 163840 - 256;
// Dense threshold (walkers) of a batch that gives none: logged dense queries
// cost no V-sized slice and a small one is counted a bucket per wavefront,
// so every query is dense (4096 walkers: 71.9 ms for 4096 queries on R-MAT
// scale 24, 1 walker: 69.5 ms; serving as fast, its latency lower); with
// slices, only the large ones.
constexpr uint64_t kDenseThresholdLogged = 1;
constexpr uint64_t kDenseThresholdSlices = 100000;
// A dense query is logged while its expected counts are at most this many
// per vertex.  Its log holds about 17.5 B per expected count where a slice
// holds 12 B per vertex, and one workgroup partitions and counts a logged
// query, where a slice is compacted by the whole device: beyond about one
// count per vertex the slice takes less memory, its V-sized scan is a small
// share of the query's walk, and a lone large query finishes sooner (serving
// at 3.75 counts per vertex: p50 3.5 ms with slices, 17.6 ms logged).
constexpr double kLogCountsPerVertex = 1.0;
// Whether dense queries of a graph of V vertices can be logged (bucket count).
inline bool batch_log_fits(uint32_t V)
{
    return V && ((uint64_t)(V - 1) >> kBatchLogShift) < kBatchLogMaxBuckets;
}

// Hub counters a dense launch with logged queries keeps for lds_hubs
// requested, beside nb bucket counters and the visit cache of its sliced
// queries (slots): a multiple of 64 that fits the LDS.
inline uint32_t batch_log_hubs(int lds_hubs, uint32_t nb, uint32_t slots)
{
    const uint64_t words = (kDenseLdsBytes - 4ull * nb - 8ull * slots) / 4;
    uint64_t h = std::min<uint64_t>((uint64_t)std::max(lds_hubs, 0), 2 * words);
    return (uint32_t)(h & ~63ull);
}
// Log segments a batch of `entries` expected log entries may claim: every
// segment is full except those a wavefront abandons when its workgroup
// changes query or ends (at most one per wavefront per dense chunk, plus
// one per resident wavefront, of which a CU holds at most 32).
inline uint64_t batch_log_segments(uint64_t entries, uint64_t dense_chunks, int block, int num_cus)
{
    return (entries + kBatchLogSeg - 1) / kBatchLogSeg + dense_chunks * (uint64_t)(block / 64) +
           32ull * (uint64_t)num_cus + 1;
}

// Dense slices are laid out every dense_pitch(V) counters: a multiple of 64,
// so every slice starts on a 256 B boundary and can be read with 16 B loads.
__host__ __device__ inline uint32_t dense_pitch(uint32_t V) { return (V + 63u) & ~63u; }

// Device ids below this (the highest-degree vertices after the relabel) of a
// hash query are counted in a dense "hub slice" of the dense pool (1 MiB),
// so most of its visits are plain atomic adds; only the rest probe its table.
constexpr uint32_t kHashHubLen = 262144;

// The sizes one batched query is planned with.  Visits are bounded by twice
// walkers * (1 + 1/tp) plus 64 for the table ("upper").  A query above the
// dense threshold gets a dense slice; one below it a hub slice of the first
// kHashHubLen device ids and an open-addressing table of the power of two
// >= 2 * upper slots (at most that of 2V) for the rest.  Either one
// gets a list of (vertex, count) pairs for the compaction after the walk,
// sized for 1.25x the expected visits (distinct vertices never exceed
// visits); a query whose list would overflow is read from its slice instead.
struct SliceSizes {
    bool dense;
    uint64_t slots;        // dense: u32 counters (the pitch); hash: uint2 slots
    uint32_t touched_cap;  // (vertex, count) pairs
    uint32_t hub_len;      // hash: u32 counters of the hub slice
    uint64_t bytes() const { return (dense ? 4 * slots : 8 * slots + 4ull * hub_len) + 8ull * touched_cap; }
};
// Counts per walker a query makes: every visit (1 + L; at most 1 + 1/tp on
// average, reference semantics) or only its last vertex (1).
inline double batch_counts_per_walker(double tp, bool ends_only) { return ends_only ? 1.0 : 1.0 + 1.0 / tp; }

inline SliceSizes batch_slice_sizes(uint32_t V, uint64_t walkers, double tp, uint64_t dense_threshold,
                                    bool ends_only = false)
{
    const double per_walker = batch_counts_per_walker(tp, ends_only);
    const uint64_t upper = (uint64_t)(2.0 * (double)walkers * per_walker) + 64;
    const uint64_t expect = (uint64_t)(1.25 * (double)walkers * per_walker) + 64;
    SliceSizes z;
    z.dense = walkers > dense_threshold;
    if (z.dense) {
        z.slots = dense_pitch(V);
    } else {
        uint64_t cap = 1;
        while (cap < 2 * upper && cap < 2ull * V) cap <<= 1;
        z.slots = cap;
    }
    z.hub_len = z.dense ? 0 : std::min<uint32_t>(dense_pitch(V), kHashHubLen);
    z.touched_cap = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(V, expect), z.dense ? V : z.slots + z.hub_len);
    return z;
}

// HBM one batched query takes under the planner's rule: a logged dense query
// holds its hub slice and, per expected count (1.25x), a log entry, a
// partition entry and a pair place.
inline uint64_t batch_query_bytes(uint32_t V, uint64_t walkers, double tp, bool ends_only, uint64_t dense_threshold,
                                  bool logged, uint32_t log_hub_len)
{
    const SliceSizes z = batch_slice_sizes(V, walkers, tp, dense_threshold, ends_only);
    const double counts = (double)walkers * batch_counts_per_walker(tp, ends_only);
    if (!(z.dense && logged && counts <= kLogCountsPerVertex * V)) return z.bytes();
    const uint64_t expect = (uint64_t)(1.25 * counts) + 64;
    return 4ull * log_hub_len + 14ull * expect;
}

// One query of a batch, as the device sees it.
struct BatchQuery {
    uint32_t source_dev;
    uint32_t thr;          // continue iff r.z < thr
    uint64_t walkers;
    uint64_t slice_off;    // dense: u32 element offset; hash: uint2 slot offset
    uint32_t kind;         // SliceKind
    uint32_t mask;         // hash: capacity - 1
    uint64_t touched_off;  // first pair of the (vertex, count) list
    uint32_t touched_cap;  // list capacity in pairs
    uint32_t stream;       // 4th Philox counter word
    uint64_t cand_off;     // first top-k candidate slot in the candidate pool
    uint32_t cand_cap;     // candidate slots (0 without top-k)
    uint32_t hub_len;      // hash: device ids below this count in a dense hub slice
    uint64_t hub_off;      // hash: u32 offset of the hub slice in the dense pool
    uint32_t ends_only;    // 1: count each walker's last vertex only (PTW_QUERY_ENDPOINTS)
    uint32_t pad;
    // SLICE_LOG: hub_len / hub_off give its hub slice; touched_off and
    // touched_cap are set on the device once the log's size is known
};


// Count bins of the top-k statistics: counts below 16 have a bin each; from
// 16 on, every power of two is cut into 8 bins by the 3 bits after the
// leading one (bin width at most 1/8 of the bin's lowest count).  240 bins
// cover all u32 counts; bin 0 (count 0) stays empty.
constexpr uint32_t kCountBins = 256;
__host__ __device__ inline uint32_t count_bin(uint32_t c)
{
    if (c < 16) return c;
    const uint32_t e = 31 - __builtin_clz(c);
    return 8 * (e - 2) + ((c >> (e - 3)) & 7);
}
__host__ __device__ inline uint64_t count_bin_lo(uint32_t b)
{
    if (b < 16) return b;
    return (uint64_t)(8 + b % 8) << (b / 8 - 1);
}
__host__ __device__ inline uint64_t count_bin_hi(uint32_t b)
{
    if (b < 16) return b;
    return ((uint64_t)(9 + b % 8) << (b / 8 - 1)) - 1;
}

// Per-query statistics of the visit counts, summed by many workgroups with
// atomics (integers only, so the result does not depend on their order):
// entries per count bin, the largest count, sum c and sum c ln c in 2^-20
// fixed point; ncand counts the top-k candidates gathered so far.
struct TopkStats {
    uint32_t hist[kCountBins];
    uint32_t maxc;
    uint32_t ncand;
    unsigned long long sum;
    unsigned long long clogc;
    uint32_t pad[2];  // 16 B multiple: hist rows stay 16 B aligned
};
static_assert(sizeof(TopkStats) % 16 == 0, "TopkStats rows must stay 16 B aligned");

struct BatchChunk {
    uint32_t q;
    uint32_t len;
    uint64_t start;  // first local walker index of the chunk
};

struct BatchLaunch {
    const uint32_t *row_ptr;
    const uint32_t *col;
    const uint4 *wedge;  // weighted-edge records, used when hdr[5] != 0
    const uint4 *nrec;   // neighbour records for unweighted batches (nullptr: row_ptr and col)
    uint32_t V;
    const BatchQuery *queries;
    const BatchChunk *chunks;
    // per-batch values live on the device so one captured graph serves every
    // batch: hdr[0] chunks, [1] queries, [2] top-k length, [3] seed,
    // [4] semantics, [5] weighted, [6] workgroup chunks of dense queries (they
    // come first in the chunk table; the wavefront chunks of hash queries
    // follow), [7] and [8] the visit log's (below); kBatchHdrWords words
    const unsigned long long *hdr;
    uint32_t *dense;         // dense slice pool (u32 counters, device ids)
    uint2 *hash;             // hash table pool {key, count}
    uint2 *touched;          // per query: (device id, count) of its nonzero entries
    uint32_t *touched_count; // per query
    uint32_t *overflow;      // per query: hash insertions that found no slot
    unsigned long long *steps;  // per query
    double *entropy;            // per query, written by the top-k kernel
    // [0]: next wavefront chunk, [1]: next workgroup chunk, [2]: next log
    // segment, [3]: next logged query to partition, [4]: to count
    unsigned long long *queue;
    TopkStats *tk_stats;        // per query
    unsigned long long *tk_cand;  // candidate pool: (count << 32 | ~device id)
    // visit log of SLICE_LOG queries (vlog == nullptr: the batch has none);
    // hdr[7] is the first pair of their lists in the touched pool, hdr[8] the
    // segments this batch may claim (at most log_segs), hdr[9] the number of
    // logged queries, counted in the order of log_order
    uint32_t *vlog;              // log_segs segments of kBatchLogSeg device ids
    uint64_t log_segs;
    uint32_t *seg_q;             // per segment: its query
    uint32_t *seg_n;             // per segment: entries written
    uint32_t *seg_list;          // segment ids grouped by query (after the walk)
    // per query, kBatchLogQWords words: [0] segments, [1] segment cursor,
    // [2] first of its segments in seg_list, [4] visits to the source,
    // [5] log entries
    uint32_t *qlog;
    const uint32_t *log_order;   // logged query indices, largest expected log first
    unsigned long long *q_base;  // per query: first entry of its partition
    uint32_t log_shift, log_nb;  // id bucket = device id >> log_shift; log_nb buckets
    // per (query, bucket): entries, first entry in the query's partition,
    // scatter cursor; per (query, run of kBatchRunIds ids): first pair of the
    // run in the query's pair list, run length
    uint32_t *part_cnt, *part_loc, *part_cur, *run_loc, *run_len;
    uint16_t *part16;            // partitioned entries: device id - bucket start
};

// Zeroes the per-query counters and the chunk queue of the batch in hdr.
hipError_t launch_batch_begin(const BatchLaunch &p, uint32_t max_queries, hipStream_t s);
// Dense queries (workgroup chunks, LDS hub counters + cache of `lds_hubs` /
// `cache_slots`), then hash queries (wavefront chunks), then the compaction
// of every query's slice into (vertex, count) pairs and count statistics.
// The persistent walk kernels fill walk_cus CUs; the compaction grid is
// sized for num_cus.
// With p.vlog, dense queries are SLICE_LOG: the dense kernel keeps lds_hubs
// hub counters (batch_log_hubs) and no cache, and the log is partitioned and
// counted before the compaction of the hash queries.
hipError_t launch_batch_walk(const BatchLaunch &p, int block, int num_cus, int walk_cus, int lds_hubs,
                             int cache_slots, hipStream_t s);
// Per-query entropy and top-k (count desc, device id asc) into out_v (graph
// ids) / out_c, from the count statistics the compaction summed: the
// candidate gather runs over segments of every query's entries on the whole
// grid; one workgroup per query sorts the candidates in a `tile`-slot LDS
// tile (or narrows the count range over the query's entries when they do
// not fit).
hipError_t launch_batch_topk(const BatchLaunch &p, const uint32_t *inv, uint32_t *out_v,
                             unsigned long long *out_c, int tile, int num_cus, hipStream_t s);
// Zero what a batch left in its slices, so the pools stay clean: a sparse
// dense slice entry by entry from its pair list, otherwise (and hash queries'
// hub slices and tables) by streaming zeros over every segment.
hipError_t launch_batch_reset(const BatchLaunch &p, int num_cus, hipStream_t s);
// out[i] = count of (graph id) vertices[i] in query qidx[i].
hipError_t launch_batch_gather(const BatchLaunch &p, uint64_t n, const uint32_t *qidx, const uint32_t *vertices,
                               const uint32_t *perm, uint64_t *out, hipStream_t s);
// Query q's full count vector in graph ids.
hipError_t launch_batch_export(const BatchLaunch &p, uint32_t q, const uint32_t *perm, const uint32_t *inv,
                               uint64_t *out, hipStream_t s);

// ---- FP64 power iteration (exact.hip) --------------------------------------
// Transpose CSR cut into pieces of at most kExactPiece in-edges.
constexpr uint32_t kExactPiece = 512;
struct ExactGraph {
    uint32_t V;
    const uint32_t *t_col;         // in-neighbours, grouped by destination
    const uint32_t *out_deg;       // out-degree of every vertex
    const uint32_t *piece_begin;   // pieces + 1 edge offsets into t_col
    const uint32_t *vertex_piece;  // V + 1 piece offsets per destination
    uint64_t pieces;
};
hipError_t launch_exact_init(uint32_t V, uint32_t B, const uint32_t *src_dev, double *y, double *x, hipStream_t s);
hipError_t launch_exact_step(const ExactGraph &g, uint32_t B, const double *y, double *y_next, double *x,
                             double *partial, const double *weight, int num_cus, hipStream_t s);
hipError_t launch_exact_export(uint32_t V, uint32_t B, const uint32_t *perm, const double *x, double *out,
                               hipStream_t s);

hipError_t launch_copy_bench(const uint4 *src, uint4 *dst, uint64_t n16, int num_cus, hipStream_t s);
// n device ids drawn in proportion to out-degree (a uniform edge's source).
hipError_t launch_stationary_starts(const uint32_t *row_ptr, uint32_t V, uint64_t E, uint64_t n, uint32_t seed_lo,
                                   uint32_t seed_hi, uint32_t *out, hipStream_t s);
// One independent walk step from each of n start records.
hipError_t launch_step_bench(const uint32_t *row_ptr, const uint32_t *col, const uint32_t *starts, uint64_t n,
                             uint32_t seed_lo, uint32_t seed_hi, unsigned long long *sink, int num_cus, hipStream_t s);
hipError_t launch_gather_bench(const uint32_t *table, uint64_t table_words, uint64_t gathers,
                               uint32_t seed_lo, uint32_t seed_hi, unsigned long long *sink,
                               int num_cus, hipStream_t s);

}  // namespace ptw
