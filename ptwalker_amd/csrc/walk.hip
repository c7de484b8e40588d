// Walk kernels for gfx950 (CDNA4, wave64).
//
// walk_kernel is persistent: the grid is sized to the device's resident
// capacity and every lane owns one walker at a time.  When a walker ends,
// its lane is refilled from the wave's private slice of walker ids; a wave
// takes a new slice of `chunk` ids from one global queue counter when its
// slice runs dry, so the queue sees one atomic per `chunk` walkers.  Lanes
// needing a walker are ranked with ballot + mbcnt (wave-level prefix sum),
// so the wave stays fully occupied until the queue drains; a wave exits
// only when the queue is empty and none of its lanes is active, which
// every wave reaches because every walk ends.
//
// Every random decision comes from one Philox4x32-10 block keyed by the
// seed and counted by (walker id, step, stream), so results do not depend
// on lane, slot, wave or launch order.
//
// Count apps privatize the `lds_hubs` highest-degree vertices (device ids
// below `hubs` after the degree-descending relabel) in LDS u16 counters, two
// per LDS word, flushed once per workgroup; other visits go to an LDS visit
// cache, then to the visit log or to device-scope u64 atomics.  The log
// takes the visits of a step from all lanes of a wavefront at once (ballot +
// mbcnt) and appends them as consecutive u32 stores into the wavefront's
// current segment of kVlogSeg entries, so a visit costs 4 B of a
// line-sized write instead of one 32 B memory-side atomic; the visit log
// kernels (vlog.hip) count the log after the walk.  Every logged visit is
// also counted in its id bucket in LDS, which plans the log's partition.
//
// The endpoint app stages ends in LDS: each wavefront keeps the ends of its
// last kEndSlots slices of walker ids (at most kEndSlice ids each) and
// writes a slice out with whole-line stores once all of its walkers have
// ended, where writing each end when it happens costs a partial-line write
// per walker.  A slice whose slot is needed again while some of its walkers
// still run (very long walks) writes the ends it has and lets the rest write
// their own (a slot generation tells them).
#include <algorithm>
#include <mutex>

#include "../../include/ptw.h"
#include "engine.hpp"
#include "philox.hpp"
#include "wave.hpp"

namespace ptw {

// Scratch of ptw_walk_static, which allocates nothing: kStaticScratchWords
// u64 words (32 MiB) on every GPU the library runs on, held until the process
// ends.  [0, 8) are the queue and total words of a walk; the rest are the
// device-id visit counters of count apps.  In host code this name is a
// host-side stand-in, not a device address: host code uses only the address
// static_scratch_address returns for the current device.
static __device__ __align__(256) unsigned long long scratch[kStaticScratchWords];

hipError_t static_scratch_address(int device, unsigned long long **out)
{
    static std::mutex mu;
    static unsigned long long *addr[kMaxStaticDevices] = {};
    if (device < 0 || device >= kMaxStaticDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        // each GPU holds its own copy of the array; the caller made `device` current
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = static_cast<unsigned long long *>(p);
    }
    *out = addr[device];
    return hipSuccess;
}

namespace {

enum Kind { KIND_COUNT = 0, KIND_END = 1, KIND_PATH = 2, KIND_N2V = 3 };

// ---- node2vec: second-order neighbour choice ---------------------------------

constexpr uint32_t kN2vAttempts = 64;  // proposals before the exact draw

// Slot of an edge key u << 32 | x: MurmurHash3's 64-bit finalizer (fmix64,
// A. Appleby), which spreads the clustered keys of one row over the table.
__host__ __device__ __forceinline__ uint64_t edge_hash(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xFF51AFD7ED558CCDull;
    k ^= k >> 33;
    k *= 0xC4CEB9FE1A85EC53ull;
    k ^= k >> 33;
    return k;
}

// Is x a neighbour of u: linear probing from the key's slot until the key or
// an empty slot (the table is at most half full, so one exists).  About one
// random line per test, where a binary search in u's sorted row takes
// log2(deg u) dependent reads.
__device__ __forceinline__ bool edge_has(const unsigned long long *T, uint64_t mask, uint32_t u, uint32_t x)
{
    const unsigned long long key = ((unsigned long long)u << 32) | x;
    for (uint64_t h = edge_hash(key) & mask;; h = (h + 1) & mask) {
        const unsigned long long k = T[h];
        if (k == key) return true;
        if (k == kEdgeSetEmpty) return false;
    }
}

// Integer weight of moving to x after coming from prev: the return, in and
// out weights.
__device__ __forceinline__ uint64_t n2v_weight(const WalkLaunch &p, uint32_t x, uint32_t prev)
{
    if (x == prev) return p.w_ret;
    return edge_has(p.eset, p.eset_mask, prev, x) ? p.w_in : p.w_out;
}

__device__ uint32_t n2v_exact(const WalkLaunch &p, uint32_t wlo, uint32_t whi, uint32_t step, uint32_t beg,
                              uint32_t deg, uint32_t prev);

// CSR slot k as a neighbour record {id, row start, degree, graph id}: read
// from the records, or only its id from `col` when the walk has none.
__device__ __forceinline__ uint4 n2v_slot(const WalkLaunch &p, uint64_t k)
{
    if (p.nrec) return p.nrec[k];
    return uint4{p.col[k], 0u, 0u, 0u};
}

// The next vertex from the row [beg, beg + deg) of the current vertex.
// Attempt a draws Philox(wid, step + a * 2^20) (attempt 0's block is `r`, the
// step's own), proposes a uniform slot and accepts it iff r.z < its weight.
// Unfolded, the slots are the deg neighbours and the weights are scaled so
// the largest is 2^32.  Folded (the return weight 1/p is the largest and the
// graph is simple and symmetric, so prev sits in the row exactly once), the
// envelope is the larger of the other two weights; the return edge keeps
// 2^32 in its own slot and its excess rides on n2v_fold extra slots, each
// accepting prev below w_fold (KnightKing's outlier folding: Yang et al.,
// SOSP 2019).  A p < 1 walk then needs about half the proposals.  After
// kN2vAttempts rejections one exact draw from the same integer weights
// follows, so the law is exactly w(x) / sum w whichever attempt decides.
// Returns the chosen neighbour's record (n2v_slot); the return edge's extra
// slots give prev_rec, the record of prev.
__device__ uint4 n2v_pick(const WalkLaunch &p, uint32_t wlo, uint32_t whi, uint32_t step, u32x4 r,
                          uint32_t beg, uint32_t deg, uint4 prev_rec)
{
    const uint32_t prev = prev_rec.x;
    // The same decisions with fewer searches: a coin below every weight
    // accepts whatever x is, a coin at or above both non-return weights
    // rejects any x but prev; only coins in between need x's class.
    const uint64_t w_io = p.w_in > p.w_out ? p.w_in : p.w_out;
    if (p.n2v_fold) {
        const uint64_t w_min = p.w_in < p.w_out ? p.w_in : p.w_out;
        for (uint32_t a = 0; a < kN2vAttempts; ++a) {
            if (a) r = philox4x32_10(wlo, whi, step + (a << 20), p.stream, p.seed_lo, p.seed_hi);
            const uint32_t s = mulhi64(r.x, r.y, deg + p.n2v_fold);
            const uint64_t z = r.z;
            if (s >= deg) {  // an extra slot of the return edge
                if (z < p.w_fold) return prev_rec;
                continue;
            }
            const uint4 e = n2v_slot(p, (uint64_t)beg + s);
            const uint32_t x = e.x;
            if (x == prev || z < w_min) return e;  // prev's own slot weighs 2^32
            if (z >= w_io) continue;
            if (z < (edge_has(p.eset, p.eset_mask, prev, x) ? p.w_in : p.w_out)) return e;
        }
        return n2v_slot(p, (uint64_t)beg + n2v_exact(p, wlo, whi, step, beg, deg, prev));
    }
    const uint64_t w_min = p.w_ret < p.w_in ? (p.w_ret < p.w_out ? p.w_ret : p.w_out) : (p.w_in < p.w_out ? p.w_in : p.w_out);
    for (uint32_t a = 0; a < kN2vAttempts; ++a) {
        if (a) r = philox4x32_10(wlo, whi, step + (a << 20), p.stream, p.seed_lo, p.seed_hi);
        const uint4 e = n2v_slot(p, (uint64_t)beg + mulhi64(r.x, r.y, deg));
        const uint32_t x = e.x;
        const uint64_t z = r.z;
        if (z < w_min) return e;
        if (x == prev) {
            if (z < p.w_ret) return e;
            continue;
        }
        if (z >= w_io) continue;
        if (z < (edge_has(p.eset, p.eset_mask, prev, x) ? p.w_in : p.w_out)) return e;
    }
    return n2v_slot(p, (uint64_t)beg + n2v_exact(p, wlo, whi, step, beg, deg, prev));
}

// The exact draw after kN2vAttempts rejections: a pick from the whole row by
// cumulative integer weight (the return edge at its full w_ret); returns the
// slot in the row.
__device__ uint32_t n2v_exact(const WalkLaunch &p, uint32_t wlo, uint32_t whi, uint32_t step, uint32_t beg,
                              uint32_t deg, uint32_t prev)
{
    uint64_t total = 0;  // < deg * 2^32 + w_ret < 2^64
    for (uint32_t k = 0; k < deg; ++k) total += n2v_weight(p, p.col[beg + k], prev);
    const u32x4 r = philox4x32_10(wlo, whi, step + (kN2vAttempts << 20), p.stream, p.seed_lo, p.seed_hi);
    const uint64_t pick = __umul64hi(((uint64_t)r.y << 32) | r.x, total);  // floor(u * total / 2^64)
    uint64_t cum = 0;
    for (uint32_t k = 0; k < deg; ++k) {
        cum += n2v_weight(p, p.col[beg + k], prev);
        if (pick < cum) return k;
    }
    return deg - 1;  // not reached: pick < total
}

// One thread per edge: its source is the last row starting at or before it
// (row_ptr[lo] <= e < row_ptr[lo + 1]), then a compare-and-swap insert
// (a repeated edge finds its own key and stops).
__global__ void edge_set_build_kernel(const uint32_t *row_ptr, const uint32_t *col, uint32_t V, uint64_t E,
                                      unsigned long long *T, uint64_t mask)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        uint32_t lo = 0, hi = V;  // row_ptr[lo] <= e < row_ptr[hi] = E
        while (hi - lo > 1) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (row_ptr[mid] <= e)
                lo = mid;
            else
                hi = mid;
        }
        const unsigned long long key = ((unsigned long long)lo << 32) | col[e];
        for (uint64_t h = edge_hash(key) & mask;; h = (h + 1) & mask) {
            const unsigned long long old = atomicCAS(&T[h], kEdgeSetEmpty, key);
            if (old == kEdgeSetEmpty || old == key) break;
        }
    }
}

// Neighbour records, one per CSR slot e (device order): {neighbour u, u's
// row start, u's degree, u's graph id}.
__global__ void edge_records_kernel(const uint32_t *row_ptr, const uint32_t *col, const uint32_t *inv, uint64_t E,
                                    uint4 *rec)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        const uint32_t u = col[e];
        const uint32_t b = row_ptr[u];
        rec[e] = uint4{u, b, row_ptr[u + 1] - b, inv ? inv[u] : u};
    }
}

// Weighted neighbour records, two per CSR slot e from its weighted-edge
// record {x, threshold, alias slot's neighbour z, 0}: rec[2e] = {x's row
// start, x's degree, x's graph id, threshold}, rec[2e + 1] = the same for z
// and 0.  Path walks only: no device ids.
__global__ void weighted_records_kernel(const uint32_t *row_ptr, const uint4 *wedge, const uint32_t *inv,
                                        uint64_t E, uint4 *rec)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        const uint4 w = wedge[e];
        const uint32_t bx = row_ptr[w.x], bz = row_ptr[w.z];
        rec[2 * e] = uint4{bx, row_ptr[w.x + 1] - bx, inv ? inv[w.x] : w.x, w.y};
        rec[2 * e + 1] = uint4{bz, row_ptr[w.z + 1] - bz, inv ? inv[w.z] : w.z, 0u};
    }
}

__global__ void fill_u64_kernel(unsigned long long *p, uint64_t n, unsigned long long v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Workgroup-private visit counters in LDS.
//   direct: u16 counters for device ids [0, hubs) (the top-degree vertices),
//           vertex x in half (x & 1) of word x >> 1: twice the hubs of u32
//           counters in the same LDS
//   cache:  open-addressing table of `slots` (power of two) {u32 key, u32 count}
//           that admits any vertex first come, first served -- the vertices a
//           workgroup visits first and most (a source's neighbourhood, the
//           hubs of a walk) end up resident; a vertex that finds no slot in
//           kProbes probes goes to a device-scope u64 atomic.
// The increment that takes a direct counter to 2^15 moves 2^15 to the device
// counter (cache counts: 2^31), so no LDS counter reaches its width.
struct LdsCounters {
    uint32_t *direct;  // (hubs + 1) / 2 words
    uint32_t *keys;
    uint32_t *vals;
    uint32_t hubs;
    uint32_t mask;  // slots - 1, or 0xFFFFFFFF when the cache is off
};

constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;
constexpr int kProbes = 8;

__device__ __forceinline__ void lds_add1(uint32_t *c, unsigned long long *dev_counter)
{
    const uint32_t old = atomicAdd(c, 1u);
    if (old == 0x7FFFFFFFu) {  // exactly one increment crosses 2^31
        atomicSub(c, 0x80000000u);
        atomicAdd(dev_counter, 0x80000000ull);
    }
}

// One visit to the u16 counter in bits [shift, shift + 16) of an LDS word.
// Exactly one increment reads 0x7FFF; it takes 2^15 back out at once.  In
// between, the other lanes of the workgroup can add at most about one visit
// each (every further visit of a lane waits for its next step's global
// loads), so the half stays below 2^15 + 1024 and never carries into the
// other half.
__device__ __forceinline__ void lds_add16(uint32_t *word, uint32_t shift, unsigned long long *dev_counter)
{
    const uint32_t old = atomicAdd(word, 1u << shift);
    if (((old >> shift) & 0xFFFFu) == 0x7FFFu) {
        atomicSub(word, 0x8000u << shift);
        atomicAdd(dev_counter, 0x8000ull);
    }
}

// Counts a visit in LDS if it has a counter or a cache slot there; false
// if it must be counted on the device.
__device__ __forceinline__ bool count_in_lds(const LdsCounters &L, uint32_t x, unsigned long long *counts)
{
    if (x < L.hubs) {
        lds_add16(&L.direct[x >> 1], (x & 1u) << 4, &counts[x]);
        return true;
    }
    if (L.mask != 0xFFFFFFFFu) {
        uint32_t h = (x * 0x9E3779B1u) & L.mask;
#pragma unroll 1
        for (int probe = 0; probe < kProbes; ++probe) {
            uint32_t k = L.keys[h];
            if (k == kEmptyKey) k = atomicCAS(&L.keys[h], kEmptyKey, x);
            if (k == kEmptyKey || k == x) {
                lds_add1(&L.vals[h], &counts[x]);
                return true;
            }
            h = (h + 1) & L.mask;
        }
    }
    return false;
}

// Path staging: entries per lane and line (128 B); a lane's LDS row (odd
// pitch), and a wavefront's LDS words: 64 rows, then 64 u64 line addresses
// and 64 lane ids of its ready lines.
constexpr uint32_t kPathChunk = 32;
constexpr uint32_t kPathRow = kPathChunk + 1;
constexpr uint32_t kPathWaveWords = 64 * kPathRow + 128 + 64;

// Endpoint staging: slots per wavefront, walker ids per slot, and the LDS
// words of one wavefront (the slots, then 5 words per slot: base lo, base
// hi, length, ended, generation).
constexpr uint32_t kEndSlots = 4;
constexpr uint32_t kEndSlice = 256;
constexpr uint32_t kEndMeta = 5;
constexpr uint32_t kEndWaveWords = kEndSlots * kEndSlice + 32;
static_assert(kEndSlots * kEndMeta <= 32, "slot words fit the wave's metadata area");
static_assert((kEndSlots & (kEndSlots - 1)) == 0, "slot ring indexed by a mask");

template <int KIND>
__global__ __launch_bounds__(1024) void walk_kernel(WalkLaunch p, uint32_t chunk, uint32_t lds_hubs,
                                                   uint32_t cache_slots)
{
    extern __shared__ uint32_t lds_words[];
    const uint32_t hub_words = (lds_hubs + 1) >> 1;
    LdsCounters L{lds_words, lds_words + hub_words, lds_words + hub_words + cache_slots, lds_hubs,
                  cache_slots ? cache_slots - 1 : 0xFFFFFFFFu};
    // visit log: per-bucket entries this workgroup appended
    const bool logging = KIND == KIND_COUNT && p.vlog != nullptr;
    uint32_t *bucket_lds = lds_words + hub_words + 2 * cache_slots;
    if (KIND == KIND_COUNT) {
        for (uint32_t i = threadIdx.x; i < hub_words; i += blockDim.x) L.direct[i] = 0;
        for (uint32_t i = threadIdx.x; i < cache_slots; i += blockDim.x) {
            L.keys[i] = kEmptyKey;
            L.vals[i] = 0;
        }
        if (logging)
            for (uint32_t i = threadIdx.x; i < p.vlog_buckets; i += blockDim.x) bucket_lds[i] = 0;
        __syncthreads();
    }
    const uint32_t lane = __lane_id();
    // endpoint staging (chunk <= kEndSlice for this kind)
    uint32_t *stage = nullptr, *meta = nullptr;
    uint32_t slice_no = 0, cur_slot = 0;  // wave-uniform
    uint64_t cur_base = 0;                // wave-uniform: first id of the current slice
    uint32_t my_slot = 0, my_gen = 0, my_off = 0;
    if (KIND == KIND_END) {
        stage = lds_words + (threadIdx.x >> 6) * kEndWaveWords;
        meta = stage + kEndSlots * kEndSlice;
        if (lane < kEndSlots * kEndMeta) meta[lane] = 0;  // every slot free (ended == length == 0)
    }
    uint64_t q_next = 0, q_end = 0;  // wave-uniform slice of local walker indices
    bool exhausted = false;
    bool active = false;
    uint64_t w = 0;      // local walker index in [0, walkers)
    uint32_t wlo = 0, whi = 0;  // global walker id words
    uint32_t v = 0, step = 0;
    uint32_t prev = 0;  // node2vec: the vertex before v
    // Neighbour records: a step reads its neighbour's record, which holds
    // the neighbour's row (start, degree) and graph id for the next step, so
    // a step is one dependent read instead of the row offsets, then the
    // neighbour, then its graph id.
    const bool use_rec = p.nrec != nullptr || p.wnrec != nullptr;
    uint32_t cur_beg = 0, cur_deg = 0, cur_g = 0;
    uint32_t prev_beg = 0, prev_deg = 0, prev_g = 0;  // node2vec with records: prev's row and graph id
    uint64_t my_steps = 0;
    uint32_t my_dangling = 0;
    // Single-source walks return to the source often: its visits are counted
    // in a register and reduced once per wave instead of hammering one address.
    const uint32_t hot = p.start_mode == PTW_START_SOURCE ? p.source_dev : PTW_INVALID_VERTEX;
    uint64_t my_hot = 0;

    // A visit that LDS does not count waits in `pend` for the wavefront's next
    // log append (at most one per lane between appends).
    uint32_t pend = kNoVisit;
    auto count_visit = [&](uint32_t x) {
        if (x == hot) {
            ++my_hot;
        } else if (!count_in_lds(L, x, p.counts_dev)) {
            if (logging)
                pend = x;
            else
                atomicAdd(&p.counts_dev[x], 1ull);
        }
    };
    // The wavefront's log segment [seg_pos, seg_end) (wave-uniform).  A
    // segment is claimed whole from the log head; once a claim finds the log
    // full, the wavefront's later visits go to device atomics.
    uint64_t seg_pos = 0, seg_end = 0;
    bool log_full = false;
    // Called by all lanes of the wavefront together.
    auto log_append = [&]() {
        if (!logging) return;
        const uint64_t m = __ballot(pend != kNoVisit);
        if (m == 0) return;
        const uint32_t n = (uint32_t)__popcll(m);
        const uint32_t r = lane_rank(m);
        const uint64_t room = seg_end - seg_pos;
        uint64_t slot = ~0ull;
        if (r < room) slot = seg_pos + r;
        if (n <= room) {
            seg_pos += n;
        } else {
            seg_pos = seg_end;
            if (!log_full) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&p.scratch[3], (unsigned long long)kVlogSeg);
                base = wave_broadcast_first(base);
                if (base + kVlogSeg <= p.vlog_cap) {
                    if (r >= room) slot = base + (r - room);
                    seg_pos = base + (n - room);
                    seg_end = base + kVlogSeg;
                } else {
                    log_full = true;
                }
            }
        }
        if (pend != kNoVisit) {
            if (slot != ~0ull) {
                p.vlog[slot] = pend;
                atomicAdd(&bucket_lds[pend >> p.vlog_shift], 1u);
            } else {
                atomicAdd(&p.counts_dev[pend], 1ull);
            }
        }
        pend = kNoVisit;
    };

    // Path rows go out through a per-lane LDS staging line of kPathChunk
    // entries (lane l of wavefront w stages entry k at [w * kPathWaveWords +
    // l * kPathRow + k]; the odd row pitch spreads lanes over the banks).
    // An entry that completes an aligned kPathChunk-entry chunk of the flat
    // path array marks the lane's line ready; after the step the wavefront
    // writes its ready lines together, eight lanes to a line, so one store
    // instruction writes eight whole 128 B lines (one request each) where a
    // lane writing its own line would issue eight 16 B requests per line.  A
    // row's unaligned head and tail go out from the lane itself (uint4 where
    // aligned).
    uint32_t *const pwave = lds_words + (threadIdx.x >> 6) * kPathWaveWords;
    uint32_t *const pstage = pwave + lane * kPathRow;  // this lane's entry k: pstage[k]
    unsigned long long *const fl_base = reinterpret_cast<unsigned long long *>(pwave + 64 * kPathRow);
    uint32_t *const fl_lane = pwave + 64 * kPathRow + 128;
    uint64_t pos = 0;  // flat index of the next path entry
    uint32_t nbuf = 0;  // staged entries of the current row not yet written
    bool line_ready = false;  // the staged line is a whole chunk of this row, for the wavefront's flush
    uint64_t line_base = 0;
    // entries [k0, k1) of the current chunk, at flat indices first + (k - k0)
    auto path_store_run = [&](uint64_t first, uint32_t k0, uint32_t k1) {
        uint32_t k = k0;
        for (; k < k1 && ((first + (k - k0)) & 3); ++k) p.paths[first + (k - k0)] = pstage[k];
        for (; k + 4 <= k1; k += 4)
            *reinterpret_cast<uint4 *>(p.paths + first + (k - k0)) =
                uint4{pstage[k], pstage[k + 1], pstage[k + 2], pstage[k + 3]};
        for (; k < k1; ++k) p.paths[first + (k - k0)] = pstage[k];
    };
    auto path_append = [&](uint32_t x) {
        const uint32_t k = (uint32_t)pos & (kPathChunk - 1);
        pstage[k] = x;
        ++nbuf;
        if (k == kPathChunk - 1) {
            const uint64_t base = pos - (kPathChunk - 1);
            if (nbuf >= kPathChunk) {
                line_ready = true;
                line_base = base;
            } else {
                // the row began inside this chunk: its entries are the last nbuf
                path_store_run(base + (kPathChunk - nbuf), kPathChunk - nbuf, kPathChunk);
            }
            nbuf = 0;
        }
        ++pos;
    };
    auto path_flush_tail = [&]() {
        // the last nbuf (< kPathChunk) entries: flat pos - nbuf .. pos - 1, staged at k1 - nbuf .. k1 - 1
        const uint32_t k1 = (uint32_t)pos & (kPathChunk - 1);
        path_store_run(pos - nbuf, k1 - nbuf, k1);
        nbuf = 0;
    };
    // The wavefront's ready lines, written together.  Every lane calls it.
    auto path_flush_lines = [&]() {
        const uint64_t m = __ballot(line_ready);
        if (m == 0) return;
        const uint32_t n = (uint32_t)__popcll(m);
        if (line_ready) {
            const uint32_t r = lane_rank(m);
            fl_base[r] = line_base;
            fl_lane[r] = lane;
        }
        wave_lds_sync();
        const uint32_t q = lane & 7;  // this lane's 16 B of each line
        for (uint32_t r = lane >> 3; r < n; r += 8) {
            const uint32_t *row = pwave + fl_lane[r] * kPathRow + 4 * q;
            *reinterpret_cast<uint4 *>(p.paths + fl_base[r] + 4 * q) = uint4{row[0], row[1], row[2], row[3]};
        }
        wave_lds_sync();  // the lines are read before their lanes stage again
        line_ready = false;
    };

    for (;;) {
        // ---- refill idle lanes (wave-uniform control flow) ----
        uint64_t idle = __ballot(!active);
        while (idle != 0 && !exhausted) {
            if (q_next == q_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&p.scratch[0], (unsigned long long)chunk);
                base = wave_broadcast_first(base);
                if (base >= p.walkers) {
                    exhausted = true;
                    break;
                }
                q_next = base;
                q_end = base + chunk < p.walkers ? base + chunk : p.walkers;
                if (KIND == KIND_END) {
                    const uint32_t s = slice_no & (kEndSlots - 1);
                    uint32_t *m = meta + s * kEndMeta;
                    uint32_t *buf = stage + s * kEndSlice;
                    const uint32_t old_len = m[2], old_done = m[3];
                    if (old_done != old_len) {
                        // the slot's older slice still has walkers running: its
                        // ended walkers go out now, the running ones write their own
                        const uint64_t ob = (uint64_t)m[0] | ((uint64_t)m[1] << 32);
                        for (uint32_t i = lane; i < old_len; i += 64) {
                            const uint32_t e = buf[i];
                            if (e != PTW_INVALID_VERTEX) p.ends[ob + i] = e;
                        }
                    }
                    const uint32_t len = (uint32_t)(q_end - q_next);
                    for (uint32_t i = lane; i < len; i += 64) buf[i] = PTW_INVALID_VERTEX;
                    const uint32_t gen = m[4] + 1;
                    if (lane == 0) {
                        m[0] = (uint32_t)q_next;
                        m[1] = (uint32_t)(q_next >> 32);
                        m[2] = len;
                        m[3] = 0;
                        m[4] = gen;
                    }
                    cur_slot = s;
                    cur_base = q_next;
                    ++slice_no;
                }
            }
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            const uint64_t avail = q_end - q_next;
            const uint32_t take = avail < n_idle ? (uint32_t)avail : n_idle;
            const uint32_t rank = lane_rank(idle);
            if (!active && rank < take) {
                w = q_next + rank;
                if (KIND == KIND_END) {
                    my_slot = cur_slot;
                    my_gen = meta[cur_slot * kEndMeta + 4];
                    my_off = (uint32_t)(w - cur_base);
                }
                const uint64_t gw = p.wid_base + w;
                wlo = (uint32_t)gw;
                whi = (uint32_t)(gw >> 32);
                uint32_t g;
                if (p.start_mode == PTW_START_SOURCE) {
                    v = p.source_dev;
                    if (KIND == KIND_END && p.starts) p.starts[w] = p.inv ? p.inv[v] : v;
                } else {
                    if (p.start_mode == PTW_START_ROUNDROBIN) {
                        g = (uint32_t)(gw % p.V);
                    } else {
                        u32x4 r = philox4x32_10(wlo, whi, kInitStep, p.stream, p.seed_lo, p.seed_hi);
                        g = mulhi64(r.x, r.y, p.V);
                    }
                    v = p.perm ? p.perm[g] : g;
                    if (KIND == KIND_COUNT) count_visit(v);
                    if (KIND == KIND_END && p.starts) p.starts[w] = g;
                }
                if (use_rec) {
                    cur_beg = p.row_ptr[v];
                    cur_deg = p.row_ptr[v + 1] - cur_beg;
                    cur_g = p.inv ? p.inv[v] : v;
                }
                if (KIND == KIND_PATH || KIND == KIND_N2V) {
                    pos = w * p.walk_len;
                    nbuf = 0;
                    path_append(use_rec ? cur_g : p.inv ? p.inv[v] : v);
                }
                step = 0;
                active = true;
            }
            if (KIND == KIND_COUNT) log_append();
            q_next += take;
            idle = __ballot(!active);
        }
        if (__ballot(active) == 0) break;

        // ---- one step of every active walker ----
        bool staged_end = false;  // endpoint app: this lane's walker ended into its slot
        if (active) {
            const u32x4 r = philox4x32_10(wlo, whi, step, p.stream, p.seed_lo, p.seed_hi);
            bool stop;
            if (KIND == KIND_PATH || KIND == KIND_N2V) {
                stop = step + 1 >= p.walk_len;
            } else {
                const bool coin = (p.semantics == PTW_SEM_TEXTBOOK) || step >= 1;
                stop = coin && r.z >= p.thr;
            }
            bool dangled = false;
            if (!stop) {
                uint32_t beg, deg;
                if (use_rec) {
                    beg = cur_beg;
                    deg = cur_deg;
                } else {
                    beg = p.row_ptr[v];
                    deg = p.row_ptr[v + 1] - beg;
                }
                if (deg == 0) {
                    dangled = true;
                    stop = true;
                } else {
                    uint32_t nx;
                    if (KIND == KIND_N2V && use_rec) {
                        // the pick returns the neighbour's record; v's becomes prev's
                        const uint4 e =
                            step > 0 ? n2v_pick(p, wlo, whi, step, r, beg, deg, uint4{prev, prev_beg, prev_deg, prev_g})
                                     : p.nrec[(uint64_t)beg + mulhi64(r.x, r.y, deg)];
                        prev_beg = beg;
                        prev_deg = deg;
                        prev_g = cur_g;
                        nx = e.x;
                        cur_beg = e.y;
                        cur_deg = e.z;
                        cur_g = e.w;
                    } else if (KIND == KIND_N2V && step > 0) {
                        nx = n2v_pick(p, wlo, whi, step, r, beg, deg, uint4{prev, 0u, 0u, 0u}).x;
                    } else {
                        const uint32_t j = mulhi64(r.x, r.y, deg);
                        if (KIND != KIND_N2V && p.wnrec) {
                            // weighted records of both alias candidates: the
                            // coin picks one, which carries its row and name
                            const uint4 a = p.wnrec[2 * ((uint64_t)beg + j)];
                            const uint4 b = p.wnrec[2 * ((uint64_t)beg + j) + 1];
                            const bool alias = r.w >= a.w;
                            cur_beg = alias ? b.x : a.x;
                            cur_deg = alias ? b.y : a.y;
                            cur_g = alias ? b.z : a.z;
                            nx = 0;  // path walks keep no device id on this path
                        } else if (KIND != KIND_N2V && p.nrec) {
                            const uint4 e = p.nrec[(uint64_t)beg + j];
                            nx = e.x;
                            cur_beg = e.y;
                            cur_deg = e.z;
                            cur_g = e.w;
                        } else if (p.wedge) {  // weighted: alias coin from the 4th word
                            const uint4 e = p.wedge[beg + j];
                            nx = r.w >= e.y ? e.z : e.x;
                        } else {
                            nx = p.col[beg + j];
                        }
                    }
                    if (KIND == KIND_N2V) prev = v;
                    v = nx;
                    ++step;
                    if (KIND == KIND_COUNT) count_visit(v);
                    if (KIND == KIND_PATH || KIND == KIND_N2V) path_append(use_rec ? cur_g : p.inv ? p.inv[v] : v);
                }
            }
            if (stop) {
                my_steps += step;
                my_dangling += dangled;
                if (KIND == KIND_END) {
                    const uint32_t e = use_rec ? cur_g : p.inv ? p.inv[v] : v;
                    if (meta[my_slot * kEndMeta + 4] == my_gen) {
                        stage[my_slot * kEndSlice + my_off] = e;
                        staged_end = true;
                    } else {
                        p.ends[w] = e;
                    }
                }
                if (KIND == KIND_PATH || KIND == KIND_N2V) {
                    path_flush_tail();
                    if (dangled) {
                        for (uint32_t k = step + 1; k < p.walk_len; ++k)
                            p.paths[w * p.walk_len + k] = PTW_INVALID_VERTEX;
                    }
                }
                active = false;
            }
        }
        if (KIND == KIND_COUNT) log_append();
        if (KIND == KIND_PATH || KIND == KIND_N2V) path_flush_lines();
        if (KIND == KIND_END && __ballot(staged_end) != 0) {
            // count the step's ends per slot; a slot whose walkers have all
            // ended goes out with whole-line stores and is free again
            for (uint32_t s = 0; s < kEndSlots; ++s) {
                const uint32_t c = (uint32_t)__popcll(__ballot(staged_end && my_slot == s));
                if (c == 0) continue;
                uint32_t *m = meta + s * kEndMeta;
                const uint32_t done = m[3] + c, len = m[2];
                if (lane == 0) m[3] = done;
                if (done == len) {
                    const uint64_t b = (uint64_t)m[0] | ((uint64_t)m[1] << 32);
                    const uint32_t *buf = stage + s * kEndSlice;
                    for (uint32_t i = lane; i < len; i += 64) p.ends[b + i] = buf[i];
                }
            }
        }
    }

    // ---- the log segment's unused tail, per-wave totals, then the LDS flush ----
    if (logging)
        for (uint64_t i = seg_pos + lane; i < seg_end; i += 64) p.vlog[i] = kNoVisit;
    uint64_t hot64 = my_hot;
    for (int off = 32; off > 0; off >>= 1) {
        my_steps += __shfl_xor(my_steps, off);
        my_dangling += __shfl_xor(my_dangling, off);
        hot64 += __shfl_xor(hot64, off);
    }
    if (lane == 0) {
        if (my_steps) atomicAdd(&p.scratch[1], (unsigned long long)my_steps);
        if (my_dangling) atomicAdd(&p.scratch[2], (unsigned long long)my_dangling);
        if (KIND == KIND_COUNT && hot64) atomicAdd(&p.counts_dev[hot], (unsigned long long)hot64);
    }
    if (KIND == KIND_COUNT) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < hub_words; i += blockDim.x) {
            const uint32_t c = L.direct[i];
            if (c & 0xFFFFu) atomicAdd(&p.counts_dev[2 * i], (unsigned long long)(c & 0xFFFFu));
            if (c >> 16) atomicAdd(&p.counts_dev[2 * i + 1], (unsigned long long)(c >> 16));
        }
        for (uint32_t i = threadIdx.x; i < cache_slots; i += blockDim.x) {
            const uint32_t k = L.keys[i], c = L.vals[i];
            if (k != kEmptyKey && c) atomicAdd(&p.counts_dev[k], (unsigned long long)c);
        }
        if (logging)
            for (uint32_t i = threadIdx.x; i < p.vlog_buckets; i += blockDim.x)
                if (bucket_lds[i]) atomicAdd(&p.vlog_hist[i], bucket_lds[i]);
    }
}

__global__ void counts_to_graph_ids_kernel(const unsigned long long *counts_dev, const uint32_t *perm,
                                           uint32_t V, uint64_t *out, int accumulate,
                                           uint32_t extra_vertex, uint64_t extra)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < V; u += stride) {
        uint64_t c = counts_dev[perm ? perm[u] : u];
        if (u == extra_vertex) c += extra;
        out[u] = accumulate ? out[u] + c : c;
    }
}

// Independent random u32 gathers: 8 loads in flight per lane per iteration.
__global__ __launch_bounds__(256) void gather_bench_kernel(const uint32_t *table, uint64_t words,
                                                          uint64_t gathers, uint32_t k0, uint32_t k1,
                                                          unsigned long long *sink)
{
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t rounds = (gathers + 8 * nthreads - 1) / (8 * nthreads);
    uint32_t acc = 0;
    const uint32_t wl = (uint32_t)(words > 0xFFFFFFFFull ? 0xFFFFFFFFull : words);
    for (uint64_t it = 0; it < rounds; ++it) {
        const u32x4 a = philox4x32_10((uint32_t)tid, (uint32_t)it, 0, 0x47415448u, k0, k1);
        const u32x4 b = philox4x32_10((uint32_t)tid, (uint32_t)it, 1, 0x47415448u, k0, k1);
        const uint32_t idx[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        uint32_t val[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint64_t g = (it * nthreads + tid) * 8 + j;
            const uint32_t i = (uint32_t)(((uint64_t)idx[j] * wl) >> 32);
            val[j] = g < gathers ? __builtin_nontemporal_load(&table[i]) : 0u;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) acc ^= val[j];
    }
    if (acc == 0x9E3779B9u) atomicAdd(sink, 1ull);  // keeps the loads live
}

// Vertices in proportion to degree: the source of a uniform random edge
// (Philox(i, seed) picks the edge; a binary search over row_ptr finds its row).
__global__ void stationary_starts_kernel(const uint32_t *row_ptr, uint32_t V, uint64_t E, uint64_t n, uint32_t k0,
                                         uint32_t k1, uint32_t *out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const u32x4 r = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 0, 0x53544154u, k0, k1);
        const uint64_t e = __umul64hi(((uint64_t)r.y << 32) | r.x, E);
        uint32_t lo = 0, hi = V;  // row_ptr[lo] <= e < row_ptr[hi]
        while (hi - lo > 1) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (row_ptr[mid] <= e)
                lo = mid;
            else
                hi = mid;
        }
        out[i] = lo;
    }
}

// One walk step from each record's vertex -- row_ptr[v], row_ptr[v + 1], then
// one uniformly picked neighbour -- with the records independent and read in
// order, four per lane in flight: the memory system's rate for a step's own
// accesses, without the dependence between a walker's steps.
__global__ __launch_bounds__(256) void step_bench_kernel(const uint32_t *row_ptr, const uint32_t *col,
                                                        const uint32_t *starts, uint64_t n, uint32_t k0, uint32_t k1,
                                                        unsigned long long *sink)
{
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
    uint32_t acc = 0;
    for (uint64_t base = tid; base < n; base += 4 * nthreads) {
        uint32_t v[4], beg[4], deg[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t i = base + (uint64_t)j * nthreads;
            v[j] = i < n ? starts[i] : 0u;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            beg[j] = row_ptr[v[j]];
            deg[j] = row_ptr[v[j] + 1] - beg[j];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t i = base + (uint64_t)j * nthreads;
            const u32x4 r = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 1, 0x53544550u, k0, k1);
            if (i < n && deg[j]) acc ^= col[beg[j] + mulhi64(r.x, r.y, deg[j])];
        }
    }
    if (acc == 0x9E3779B9u) atomicAdd(sink, 1ull);  // keeps the loads live
}

// Plain streaming copy, 16 B per lane per iteration: the HBM stream roofline
// and the byte-count calibration for the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void copy_bench_kernel(const uint4 *src, uint4 *dst, uint64_t n16)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

template <int KIND>
int occupancy_blocks(int block, size_t lds)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, walk_kernel<KIND>, block, lds) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return per_cu;
}

}  // namespace

hipError_t launch_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_edge_set_build(const uint32_t *row_ptr, const uint32_t *col, uint32_t V, uint64_t E,
                                 unsigned long long *table, uint64_t mask, hipStream_t s)
{
    if (E == 0) return hipSuccess;
    uint64_t blocks = (E + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(edge_set_build_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, row_ptr, col, V, E, table,
                       mask);
    return hipGetLastError();
}

hipError_t launch_edge_records(const uint32_t *row_ptr, const uint32_t *col, const uint32_t *inv, uint64_t E,
                               uint4 *rec, hipStream_t s)
{
    if (E == 0) return hipSuccess;
    uint64_t blocks = (E + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(edge_records_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, row_ptr, col, inv, E, rec);
    return hipGetLastError();
}

hipError_t launch_weighted_records(const uint32_t *row_ptr, const uint4 *wedge, const uint32_t *inv, uint64_t E,
                                   uint4 *rec, hipStream_t s)
{
    if (E == 0) return hipSuccess;
    uint64_t blocks = (E + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(weighted_records_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, row_ptr, wedge, inv, E,
                       rec);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_walk(const WalkLaunch &p_in, const LaunchShape &shape, int num_cus, hipStream_t s, EffectiveShape *eff)
{
    WalkLaunch p = p_in;
    const int kind = (p.app == PTW_APP_PAGERANK || p.app == PTW_APP_PPR) ? KIND_COUNT
                     : p.app == PTW_APP_RWR                            ? KIND_END
                     : p.app == PTW_APP_NODE2VEC                       ? KIND_N2V
                                                                       : KIND_PATH;
    int block = shape.block;
    uint32_t hubs = 0, slots = 0;
    if (kind == KIND_COUNT) {
        hubs = (uint32_t)shape.lds_hubs < p.hubs ? (uint32_t)shape.lds_hubs : p.hubs;
        slots = shape.cache_slots > 0 ? (uint32_t)shape.cache_slots : 0u;
    } else {
        p.vlog = nullptr;
    }
    if (kind == KIND_N2V) p.wnrec = nullptr;  // node2vec walks are unweighted
    if (kind != KIND_PATH) p.wnrec = nullptr;  // weighted records carry no device ids
    if (p.wnrec) p.nrec = nullptr;
    auto count_lds = [&]() {
        return (((size_t)hubs + 1) / 2 + 2ull * slots + (p.vlog ? p.vlog_buckets : 0u)) * sizeof(uint32_t);
    };
    auto path_lds = [&]() { return (size_t)(block / 64) * kPathWaveWords * sizeof(uint32_t); };
    size_t lds = kind == KIND_END    ? (size_t)(block / 64) * kEndWaveWords * sizeof(uint32_t)
                 : kind == KIND_COUNT ? count_lds()
                                      : path_lds();
    // Workgroups beyond the resident capacity are harmless (no workgroup waits
    // on another): they start late and find the queue drained.
    int per_cu = shape.blocks_per_cu;
    if (per_cu <= 0) {
        // The endpoint kernel's steps wait on one read each (neighbour
        // records) and its LDS holds more workgroups than 1024-thread ones
        // can use (SGPRs allow 28 waves per CU): 512-thread workgroups put
        // 24 waves on a CU instead of 16.
        if (kind == KIND_END && block > 512) {
            block = 512;
            lds = (size_t)(block / 64) * kEndWaveWords * sizeof(uint32_t);
        }
        per_cu = kind == KIND_COUNT ? occupancy_blocks<KIND_COUNT>(block, lds)
                 : kind == KIND_END ? occupancy_blocks<KIND_END>(block, lds)
                 : kind == KIND_N2V ? occupancy_blocks<KIND_N2V>(block, lds)
                                    : occupancy_blocks<KIND_PATH>(block, lds);
        // A small walk (under 8 walkers per resident lane) takes as long as its
        // longest walker chains plus each workgroup's LDS setup and flush: one
        // 512-thread workgroup per CU with at most 8192 hub counters measured
        // fastest.  A visit cache only pays for single-source walks, whose
        // source neighbours are hot; uniform starts spread their visits.
        if (p.walkers < 8ull * (uint64_t)per_cu * (uint64_t)num_cus * (uint64_t)block) {
            block = block < 512 ? block : 512;
            per_cu = 1;
            hubs = hubs < 8192u ? hubs : 8192u;
            slots = p.start_mode == PTW_START_SOURCE ? (slots < 1024u ? slots : 1024u) : 0u;
            p.vlog = nullptr;  // a few visits per vertex: the atomics are cheaper than the log's passes
        }
    }
    if (kind == KIND_COUNT && p.vlog) {
        // The log takes a source's hot neighbourhood as cheaply as the visit
        // cache did (it only had to keep those visits off hot device
        // counters), so an automatic cache gives its LDS to hub counters.
        if (shape.cache_auto && slots) {
            const uint32_t more = hubs + 4 * slots < (uint32_t)kMaxLdsHubs ? hubs + 4 * slots : (uint32_t)kMaxLdsHubs;
            hubs = more < p.hubs ? more : p.hubs;
            slots = 0;
        }
        // The log's bucket counters share the workgroup's LDS (160 KiB) with
        // the hub counters and the cache; the hub range gives way.
        constexpr uint32_t kLdsWords = 160u * 1024u / 4u;
        if (hubs >= p.V || p.vlog_buckets == 0 || p.vlog_buckets > kVlogMaxBuckets ||
            2ull * slots + p.vlog_buckets >= kLdsWords) {
            p.vlog = nullptr;  // every visit has an LDS counter, or no room
        } else {
            const uint32_t max_hub_words = kLdsWords - 2 * slots - p.vlog_buckets;
            if ((hubs + 1) / 2 > max_hub_words) hubs = 2 * max_hub_words;
        }
    }
    if (kind == KIND_COUNT) lds = count_lds();
    else if (kind == KIND_END) lds = (size_t)(block / 64) * kEndWaveWords * sizeof(uint32_t);
    else lds = path_lds();
    const int blocks = per_cu * num_cus;
    // Walker ids per queue take: the configured chunk, cut down for small
    // walks so that every resident wavefront gets about four takes.
    const uint64_t waves = (uint64_t)blocks * (uint64_t)(block / 64);
    uint64_t fair = p.walkers / (waves * 4);
    fair = (fair + 63) / 64 * 64;
    uint32_t chunk = (uint32_t)std::max<uint64_t>(64, std::min<uint64_t>(fair, (uint64_t)shape.chunk));
    if (kind == KIND_END && chunk > kEndSlice) chunk = kEndSlice;  // a slice fits its staging slot
    if (eff) {
        eff->block = block;
        eff->blocks = blocks;
        eff->chunk = chunk;
        eff->hubs = hubs;
        eff->slots = slots;
        eff->vlog = p.vlog != nullptr;
    }
    if (kind == KIND_COUNT)
        hipLaunchKernelGGL(walk_kernel<KIND_COUNT>, dim3(blocks), dim3(block), lds, s, p, chunk, hubs, slots);
    else if (kind == KIND_END)
        hipLaunchKernelGGL(walk_kernel<KIND_END>, dim3(blocks), dim3(block), lds, s, p, chunk, 0u, 0u);
    else if (kind == KIND_N2V)
        hipLaunchKernelGGL(walk_kernel<KIND_N2V>, dim3(blocks), dim3(block), lds, s, p, chunk, 0u, 0u);
    else
        hipLaunchKernelGGL(walk_kernel<KIND_PATH>, dim3(blocks), dim3(block), lds, s, p, chunk, 0u, 0u);
    return hipGetLastError();
}

hipError_t launch_counts_to_graph_ids(const unsigned long long *counts_dev, const uint32_t *perm,
                                      uint32_t V, uint64_t *out, bool accumulate,
                                      uint32_t extra_vertex, uint64_t extra, hipStream_t s)
{
    uint64_t blocks = ((uint64_t)V + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(counts_to_graph_ids_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, counts_dev,
                       perm, V, out, accumulate ? 1 : 0, extra_vertex, extra);
    return hipGetLastError();
}

hipError_t launch_stationary_starts(const uint32_t *row_ptr, uint32_t V, uint64_t E, uint64_t n, uint32_t seed_lo,
                                   uint32_t seed_hi, uint32_t *out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(stationary_starts_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, row_ptr, V, E, n, seed_lo,
                       seed_hi, out);
    return hipGetLastError();
}

hipError_t launch_step_bench(const uint32_t *row_ptr, const uint32_t *col, const uint32_t *starts, uint64_t n,
                             uint32_t seed_lo, uint32_t seed_hi, unsigned long long *sink, int num_cus, hipStream_t s)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, step_bench_kernel, 256, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    hipLaunchKernelGGL(step_bench_kernel, dim3(per_cu * num_cus), dim3(256), 0, s, row_ptr, col, starts, n, seed_lo,
                       seed_hi, sink);
    return hipGetLastError();
}

hipError_t launch_copy_bench(const uint4 *src, uint4 *dst, uint64_t n16, int num_cus, hipStream_t s)
{
    hipLaunchKernelGGL(copy_bench_kernel, dim3(num_cus * 16), dim3(256), 0, s, src, dst, n16);
    return hipGetLastError();
}

hipError_t launch_gather_bench(const uint32_t *table, uint64_t table_words, uint64_t gathers,
                               uint32_t seed_lo, uint32_t seed_hi, unsigned long long *sink,
                               int num_cus, hipStream_t s)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gather_bench_kernel, 256, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    hipLaunchKernelGGL(gather_bench_kernel, dim3(per_cu * num_cus), dim3(256), 0, s, table, table_words,
                       gathers, seed_lo, seed_hi, sink);
    return hipGetLastError();
}

}  // namespace ptw
