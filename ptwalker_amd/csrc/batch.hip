// Batched single-source walks for gfx950: PPR queries count every visit
// (type 0, src/core/solver.hpp:242-262), RWR queries count each walker's
// last vertex (type 1, src/core/solver.hpp:284-305, src/core/RWR.cpp:99-105).
//
// All queries of a batch run in two persistent launches.  The host cuts
// every query into chunks of walker ids, those of dense queries first.
// batch_dense_kernel: workgroups take the dense chunks from one queue
// counter; a workgroup's LDS holds direct counters for the highest-degree
// vertices and a hash cache for the rest, both flushed to the query's slice
// at a barrier when the workgroup moves to another query.
// batch_walk_kernel: wavefronts take the remaining (hash) chunks and refill
// idle lanes from their current chunk, so lanes of one wavefront may hold
// walkers of different queries; each wavefront keeps kWaveHubs direct LDS
// counters and a small LDS cache of (query, vertex) counts.
// Query q's walker w draws from Philox(key = seed, ctr = (w, step, stream_q)):
// a query's counts are the same alone (ptw_walk with stream = stream_q) and
// in any mix.  A query's visits to its own source are counted in a register.
//
// Counts go to per-query slices:
//   dense slice: u32[V] in device ids
//   hash slice:  open addressing over {u32 key, u32 count}, capacity a power
//                of two sized from the query's expected visits
// After the walk a compaction pass lists each query's visited vertices as
// (vertex, count) pairs, so top-k and the cleanup read only those.
//
// The compaction also sums each query's count statistics: entries per count
// bin (8 bins per power of two), the largest count and the entropy terms.
// Top-k: the bin holding the k-th count is read from them; when that bin and
// the ones above fit the candidate slots, every workgroup gathers its share
// of the candidates (counts >= the bin's lowest) and one workgroup per query
// bitonic-sorts them in LDS by (count desc, device id asc).  Otherwise that
// workgroup narrows the count range with LDS histograms over the query's
// entries until the candidates fit; ties that do not fit the tile are cut by
// an exact radix select on the device id.
#include <type_traits>

#include "../../include/ptw.h"
#include "engine.hpp"
#include "philox.hpp"
#include "wave.hpp"

namespace ptw {

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kBins = 2048;

constexpr unsigned long long kEmpty64 = ~0ull;
constexpr uint32_t kWaveSlots = 256;  // per-wavefront LDS cache slots
constexpr uint32_t kWaveHubs = 1024;  // per-wavefront direct hub counters
constexpr int kCacheProbes = 8;
constexpr uint32_t kSeg = 65536;  // dense-slice compaction segment

// 16 B vector for streaming stores that bypass the caches' reuse
typedef uint32_t vec4u __attribute__((ext_vector_type(4)));

// Where a query's counts go: its dense slice, or its hub slice and table.
struct SliceRef {
    uint32_t kind = 0, mask = 0, hub_len = 0;
    uint64_t off = 0, hub_off = 0;
};
__device__ __forceinline__ SliceRef slice_ref(const BatchQuery &bq)
{
    SliceRef r;
    r.kind = bq.kind;
    r.mask = bq.mask;
    r.hub_len = bq.hub_len;
    r.off = bq.slice_off;
    r.hub_off = bq.hub_off;
    return r;
}

struct LaneQuery {
    uint32_t q = kEmpty;
    uint32_t src = 0, thr = 0, stream = 0;
    uint32_t ends = 0;  // count only the walker's last vertex
    SliceRef r;
};

__device__ __forceinline__ void slice_add(const BatchLaunch &p, uint32_t q, const SliceRef &r, uint32_t v,
                                          uint32_t n)
{
    if (r.kind == SLICE_DENSE) {
        atomicAdd(&p.dense[r.off + v], n);
        return;
    }
    if (v < r.hub_len) {
        atomicAdd(&p.dense[r.hub_off + v], n);
        return;
    }
    const uint64_t off = r.off;
    const uint32_t mask = r.mask;
    // a slot {key, count} is one little-endian u64 (count << 32 | key): an
    // existing key takes one 64-bit add on the count half, an empty slot one
    // compare-and-swap that installs the key with its first count
    constexpr unsigned long long kEmptySlot = kEmpty;  // {kEmpty, 0}
    uint32_t h = (v * 0x9E3779B1u) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        unsigned long long *slot = reinterpret_cast<unsigned long long *>(&p.hash[off + h]);
        unsigned long long cur = *slot;
        if (cur == kEmptySlot) {
            cur = atomicCAS(slot, kEmptySlot, ((unsigned long long)n << 32) | v);
            if (cur == kEmptySlot) return;
        }
        if ((uint32_t)cur == v) {
            atomicAdd(slot, (unsigned long long)n << 32);
            return;
        }
        h = (h + 1) & mask;
    }
    atomicAdd(&p.overflow[q], n);
}

// Per-wavefront LDS cache of (query, vertex) counts in front of the slices:
// the lanes of a wavefront walk one chunk (one query) at a time, whose hot
// vertices -- the source's neighbourhood -- would otherwise collide in global
// atomics.  Each wavefront owns its region, so no barrier is needed; the
// region is flushed whenever the wavefront moves to a chunk of another query.
struct WaveCache {
    unsigned long long *keys;
    uint32_t *cnt;
    __device__ __forceinline__ bool add(uint32_t q, uint32_t v) const
    {
        const unsigned long long key = ((unsigned long long)q << 32) | v;
        uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 56) & (kWaveSlots - 1);
#pragma unroll 1
        for (int probe = 0; probe < kCacheProbes; ++probe) {
            unsigned long long k = keys[h];
            if (k == kEmpty64) k = atomicCAS(&keys[h], kEmpty64, key);
            if (k == kEmpty64 || k == key) {
                atomicAdd(&cnt[h], 1u);
                return true;
            }
            h = (h + 1) & (kWaveSlots - 1);
        }
        return false;
    }
    // Wavefront-uniform: every lane of the wavefront calls it.
    __device__ __forceinline__ void flush(const BatchLaunch &p, uint32_t lane) const
    {
        for (uint32_t s = lane; s < kWaveSlots; s += 64) {
            const unsigned long long k = keys[s];
            if (k != kEmpty64) {
                const uint32_t q = (uint32_t)(k >> 32);
                slice_add(p, q, slice_ref(p.queries[q]), (uint32_t)k, cnt[s]);
                keys[s] = kEmpty64;
                cnt[s] = 0;
            }
        }
    }
};

__global__ void batch_begin_kernel(BatchLaunch p, uint32_t max_queries)
{
    const uint32_t nq = (uint32_t)p.hdr[1];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        p.queue[0] = 0;
        p.queue[1] = 0;
        p.queue[2] = 0;
        p.queue[3] = 0;
        p.queue[4] = 0;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < max_queries && i < nq; i += gridDim.x * blockDim.x) {
        p.touched_count[i] = 0;
        p.overflow[i] = 0;
        p.steps[i] = 0;
        if (p.vlog) {
            for (uint32_t w = 0; w < kBatchLogQWords; ++w) p.qlog[(uint64_t)i * kBatchLogQWords + w] = 0;
        }
    }
    if (p.vlog) {
        // the per-(query, bucket) entry counts the walk adds to
        const uint64_t n = (uint64_t)(nq < max_queries ? nq : max_queries) * p.log_nb;
        for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n; w += (uint64_t)gridDim.x * blockDim.x)
            p.part_cnt[w] = 0;
    }
    // the top-k statistics, as one flat array of 16 B words
    const uint32_t q_used = nq < max_queries ? nq : max_queries;
    const uint64_t words = (uint64_t)q_used * (sizeof(TopkStats) / 16);
    uint4 *st4 = reinterpret_cast<uint4 *>(p.tk_stats);
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x)
        st4[w] = uint4{0u, 0u, 0u, 0u};
}

__global__ __launch_bounds__(1024) void batch_walk_kernel(BatchLaunch p)
{
    extern __shared__ unsigned long long wave_keys[];
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x / 64, waves = blockDim.x / 64;
    uint32_t *const wave_cnt = reinterpret_cast<uint32_t *>(wave_keys + (size_t)waves * kWaveSlots);
    const WaveCache C{wave_keys + (size_t)wave * kWaveSlots, wave_cnt + wave * kWaveSlots};
    uint32_t *const hub = wave_cnt + (size_t)waves * kWaveSlots + (size_t)wave * kWaveHubs;
    for (uint32_t s = lane; s < kWaveSlots; s += 64) {
        C.keys[s] = kEmpty64;
        C.cnt[s] = 0;
    }
    for (uint32_t s = lane; s < kWaveHubs; s += 64) hub[s] = 0;
    const uint64_t num_chunks = p.hdr[0];
    const uint64_t first_chunk = p.hdr[6];  // the workgroup chunks of dense queries come first
    const uint32_t seed_lo = (uint32_t)p.hdr[3], seed_hi = (uint32_t)(p.hdr[3] >> 32);
    const bool textbook = p.hdr[4] == PTW_SEM_TEXTBOOK;
    const uint4 *wedge = p.hdr[5] ? p.wedge : nullptr;
    const uint4 *nrec = wedge ? nullptr : p.nrec;  // neighbour records (unweighted batches)
    uint32_t cur_beg = 0, cur_deg = 0;             // the row of v, from its record
    uint64_t c_next = 0, c_end = 0;  // wave-uniform current chunk
    uint32_t cq = kEmpty;
    bool exhausted = false;
    bool active = false;
    LaneQuery Q;
    uint64_t w = 0;
    uint32_t v = 0, step = 0;
    uint32_t my_src = 0;
    uint64_t my_steps = 0;

    // One count for the lane's query: its own source in a register, the
    // wavefront's hub counters (only while they hold the lane's query), the
    // query's hub slice, the LDS cache, then the query's slice.
    auto count = [&](uint32_t x) {
        if (x == Q.src)
            ++my_src;
        else if (x < kWaveHubs && Q.q == cq)
            atomicAdd(&hub[x], 1u);
        else if (x < Q.r.hub_len)
            atomicAdd(&p.dense[Q.r.hub_off + x], 1u);  // no reply to wait for
        else if (!C.add(Q.q, x))
            slice_add(p, Q.q, Q.r, x, 1u);
    };

    // hub counters belong to cq, the query of the wavefront's current chunk
    auto flush_hubs = [&]() {
        if (cq == kEmpty) return;
        const SliceRef r = slice_ref(p.queries[cq]);
        for (uint32_t s = lane; s < kWaveHubs; s += 64) {
            if (hub[s]) {
                slice_add(p, cq, r, s, hub[s]);
                hub[s] = 0;
            }
        }
    };
    auto flush = [&]() {
        if (Q.q == kEmpty) return;
        if (my_src) slice_add(p, Q.q, Q.r, Q.src, my_src);
        if (my_steps) atomicAdd(&p.steps[Q.q], (unsigned long long)my_steps);
        my_src = 0;
        my_steps = 0;
    };

    for (;;) {
        uint64_t idle = __ballot(!active);
        while (idle != 0 && !exhausted) {
            if (c_next == c_end) {
                unsigned long long ci = 0;
                if (lane == 0) ci = atomicAdd(&p.queue[0], 1ull);
                ci = first_chunk + wave_broadcast_first(ci);
                if (ci >= num_chunks) {
                    exhausted = true;
                    break;
                }
                const BatchChunk ch = p.chunks[ci];
                if (ch.q != cq && cq != kEmpty) {  // wave-uniform
                    C.flush(p, lane);
                    flush_hubs();
                }
                cq = ch.q;
                c_next = ch.start;
                c_end = ch.start + ch.len;
            }
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            const uint64_t avail = c_end - c_next;
            const uint32_t take = avail < n_idle ? (uint32_t)avail : n_idle;
            const uint32_t rank = lane_rank(idle);
            if (!active && rank < take) {
                if (Q.q != cq) {
                    flush();
                    const BatchQuery bq = p.queries[cq];
                    Q.q = cq;
                    Q.src = bq.source_dev;
                    Q.thr = bq.thr;
                    Q.stream = bq.stream;
                    Q.ends = bq.ends_only;
                    Q.r = slice_ref(bq);
                }
                w = c_next + rank;
                v = Q.src;
                if (nrec) {
                    cur_beg = p.row_ptr[v];
                    cur_deg = p.row_ptr[v + 1] - cur_beg;
                }
                step = 0;
                if (!Q.ends) ++my_src;  // the start visit
                active = true;
            }
            c_next += take;
            idle = __ballot(!active);
        }
        if (__ballot(active) == 0) break;

        if (active) {
            const u32x4 r = philox4x32_10((uint32_t)w, (uint32_t)(w >> 32), step, Q.stream, seed_lo, seed_hi);
            const bool coin = textbook || step >= 1;
            bool stop = coin && r.z >= Q.thr;
            if (!stop) {
                const uint32_t beg = nrec ? cur_beg : p.row_ptr[v];
                const uint32_t deg = nrec ? cur_deg : p.row_ptr[v + 1] - beg;
                if (deg == 0) {
                    stop = true;  // batches end walkers at vertices without out-edges
                } else {
                    const uint32_t j = mulhi64(r.x, r.y, deg);
                    if (wedge) {
                        const uint4 e = wedge[beg + j];
                        v = r.w >= e.y ? e.z : e.x;
                    } else if (nrec) {
                        const uint4 e = nrec[(uint64_t)beg + j];
                        v = e.x;
                        cur_beg = e.y;
                        cur_deg = e.z;
                    } else {
                        v = p.col[beg + j];
                    }
                    ++step;
                    if (!Q.ends) count(v);
                }
            }
            if (stop) {
                if (Q.ends) count(v);
                my_steps += step;
                active = false;
            }
        }
    }
    flush();
    C.flush(p, lane);
    flush_hubs();
}

// Dense queries, workgroup-coherent.  A workgroup takes one chunk of one
// query at a time from the workgroup queue, and its wavefronts refill idle
// lanes from the chunk through an LDS offset.  Visits of the workgroup's
// query go to a register (the query's own source), to LDS direct counters for
// device ids below `hubs` (the top-degree vertices after the relabel), to an
// LDS first-come hash cache of `slots` other vertices, and otherwise to the
// query's dense slice.  The LDS counts are flushed to the slice at a barrier
// when the workgroup's next chunk belongs to another query, and at the end.
// Every wavefront reaches each barrier: a chunk runs dry and every walk ends.
// The direct counters are u16, two per LDS word (vertex x in half x & 1 of
// word x >> 1); the increment that takes one to 2^15 moves 2^15 to the slice.
// The cache counts are u32 and cannot wrap: none exceeds its query's visit
// total, which the host checks is below 2^32.
__global__ __launch_bounds__(1024) void batch_dense_kernel(BatchLaunch p, uint32_t hubs, uint32_t slots,
                                                          uint32_t nb)
{
    extern __shared__ uint32_t dense_lds[];
    const uint32_t hub_words = (hubs + 1) >> 1;
    uint32_t *const direct = dense_lds;
    uint32_t *const keys = dense_lds + hub_words;
    uint32_t *const vals = keys + slots;
    const uint32_t mask = slots ? slots - 1 : 0u;
    // launches with logged queries (nb > 0): entries per id bucket of the
    // workgroup's query when it is logged (the cache serves sliced queries)
    uint32_t *const bkt = dense_lds + hub_words + 2 * slots;
    const bool logging = nb != 0;
    __shared__ uint32_t s_q, s_len, s_off, s_done;
    __shared__ unsigned long long s_start;

    for (uint32_t i = threadIdx.x; i < hub_words; i += blockDim.x) direct[i] = 0;
    for (uint32_t i = threadIdx.x; i < slots; i += blockDim.x) {
        keys[i] = kEmpty;
        vals[i] = 0;
    }
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) bkt[i] = 0;
    const uint32_t lane = __lane_id();
    const uint64_t n_dense = p.hdr[6];
    const uint32_t seed_lo = (uint32_t)p.hdr[3], seed_hi = (uint32_t)(p.hdr[3] >> 32);
    const bool textbook = p.hdr[4] == PTW_SEM_TEXTBOOK;
    const uint4 *wedge = p.hdr[5] ? p.wedge : nullptr;
    const uint4 *nrec = wedge ? nullptr : p.nrec;  // neighbour records (unweighted batches)
    uint32_t cq = kEmpty;  // workgroup-uniform: the query whose counts the LDS holds
    uint32_t src = 0, thr = 0, stream = 0;
    uint32_t src_beg = 0, src_deg = 0;  // workgroup-uniform: the source's row
    uint32_t cur_beg = 0, cur_deg = 0;  // the row of v, from its record
    bool ends = false;  // workgroup-uniform: count only each walker's last vertex
    bool qlogged = false;  // workgroup-uniform: the query is SLICE_LOG
    uint64_t off = 0;  // workgroup-uniform: the query's dense slice, or its hub slice when logged
    uint32_t my_src = 0;
    uint64_t my_steps = 0;
    // A visit the LDS does not take waits in `pend` for the wavefront's next
    // log append (at most one per lane between appends).
    uint32_t pend = kNoVisit;
    // The wavefront's log segment: entries [seg_pos, seg_end) of segment
    // seg_id (wave-uniform; seg_id = ~0 when it holds none).
    uint64_t seg_pos = 0, seg_end = 0, seg_id = ~0ull;
    // segments this batch may claim (hdr[8] <= the log's allocation)
    const uint64_t seg_limit = logging ? (p.hdr[8] < p.log_segs ? p.hdr[8] : p.log_segs) : 0;

    // One count of the workgroup's query: its source in a register, the LDS
    // hub counters, then the LDS cache and the query's dense slice, or the log.
    auto count = [&](uint32_t x) {
        if (x == src) {
            ++my_src;
        } else if (x < hubs) {
            // the u16 spill rule of walk.hip's lds_add16, into the u32 slice
            const uint32_t sh = (x & 1u) << 4;
            const uint32_t old = atomicAdd(&direct[x >> 1], 1u << sh);
            if (((old >> sh) & 0xFFFFu) == 0x7FFFu) {
                atomicSub(&direct[x >> 1], 0x8000u << sh);
                atomicAdd(&p.dense[off + x], 0x8000u);
            }
        } else if (qlogged) {
            pend = x;
        } else {
            if (slots) {
                uint32_t h = (x * 0x9E3779B1u) & mask;
#pragma unroll 1
                for (int probe = 0; probe < kCacheProbes; ++probe) {
                    uint32_t k = keys[h];
                    if (k == kEmpty) k = atomicCAS(&keys[h], kEmpty, x);
                    if (k == kEmpty || k == x) {
                        atomicAdd(&vals[h], 1u);
                        return;
                    }
                    h = (h + 1) & mask;
                }
            }
            atomicAdd(&p.dense[off + x], 1u);
        }
    };

    // Appends the wavefront's pending visits to its log segment as
    // consecutive u32 stores; a full segment is closed and a new one claimed
    // for the workgroup's query.  A claim beyond the log's capacity drops the
    // visits and marks the query in p.overflow (the batch then fails).
    // Called by all lanes of the wavefront together.
    auto log_append = [&]() {
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
            unsigned long long s = 0;
            if (lane == 0) {
                if (seg_id != ~0ull) p.seg_n[seg_id] = kBatchLogSeg;  // filled to the end
                s = atomicAdd(&p.queue[2], 1ull);
                if (s < seg_limit) {
                    p.seg_q[s] = cq;
                    atomicAdd(&p.qlog[(uint64_t)cq * kBatchLogQWords], 1u);
                } else {
                    atomicAdd(&p.overflow[cq], n - (uint32_t)room);
                }
            }
            s = wave_broadcast_first(s);
            if (s < seg_limit) {
                const uint64_t base = s * kBatchLogSeg;
                if (r >= room) slot = base + (r - room);
                seg_id = s;
                seg_pos = base + (n - room);
                seg_end = base + kBatchLogSeg;
            } else {
                seg_id = ~0ull;
                seg_pos = seg_end = 0;
            }
        }
        if (pend != kNoVisit) {
            if (slot != ~0ull) {
                p.vlog[slot] = pend;
                atomicAdd(&bkt[pend >> p.log_shift], 1u);
            }
            pend = kNoVisit;
        }
    };

    // Workgroup-uniform: every thread calls it, between barriers.
    auto flush = [&]() {
        if (cq == kEmpty) return;
        uint64_t hot = my_src, st = my_steps;
        for (int o = 32; o > 0; o >>= 1) {
            hot += __shfl_xor(hot, o);
            st += __shfl_xor(st, o);
        }
        if (lane == 0) {
            if (hot) {
                if (qlogged)
                    atomicAdd(&p.qlog[(uint64_t)cq * kBatchLogQWords + 4], (uint32_t)hot);
                else
                    atomicAdd(&p.dense[off + src], (uint32_t)hot);
            }
            if (st) atomicAdd(&p.steps[cq], (unsigned long long)st);
            // the wavefront's segment holds this query's entries only
            if (logging && seg_id != ~0ull) p.seg_n[seg_id] = (uint32_t)(seg_pos - seg_id * kBatchLogSeg);
        }
        seg_id = ~0ull;
        seg_pos = seg_end = 0;
        my_src = 0;
        my_steps = 0;
        for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
            const uint32_t c = bkt[b];
            if (c) {
                atomicAdd(&p.part_cnt[(uint64_t)cq * nb + b], c);
                bkt[b] = 0;
            }
        }
        for (uint32_t i = threadIdx.x; i < hub_words; i += blockDim.x) {
            const uint32_t c = direct[i];
            if (c) {
                if (c & 0xFFFFu) atomicAdd(&p.dense[off + 2 * i], c & 0xFFFFu);
                if (c >> 16) atomicAdd(&p.dense[off + 2 * i + 1], c >> 16);
                direct[i] = 0;
            }
        }
        for (uint32_t i = threadIdx.x; i < slots; i += blockDim.x) {
            const uint32_t k = keys[i];
            if (k != kEmpty) {
                atomicAdd(&p.dense[off + k], vals[i]);
                keys[i] = kEmpty;
                vals[i] = 0;
            }
        }
    };

    for (;;) {
        __syncthreads();  // the previous chunk is walked; the chunk state is free
        if (threadIdx.x == 0) {
            const unsigned long long ci = atomicAdd(&p.queue[1], 1ull);
            if (ci >= n_dense) {
                s_done = 1;
            } else {
                const BatchChunk ch = p.chunks[ci];
                s_q = ch.q;
                s_start = ch.start;
                s_len = ch.len;
                s_off = 0;
                s_done = 0;
            }
        }
        __syncthreads();
        if (s_done) break;
        const uint32_t q = s_q;
        if (q != cq) {
            flush();
            __syncthreads();  // LDS is clear before counting for q
            cq = q;
            const BatchQuery bq = p.queries[q];
            src = bq.source_dev;
            thr = bq.thr;
            stream = bq.stream;
            ends = bq.ends_only != 0;
            qlogged = bq.kind == SLICE_LOG;
            off = qlogged ? bq.hub_off : bq.slice_off;
            src_beg = p.row_ptr[src];
            src_deg = p.row_ptr[src + 1] - src_beg;
        }
        const uint64_t start = s_start;
        const uint32_t len = s_len;
        bool exhausted = false, active = false;
        uint64_t w = 0;
        uint32_t v = 0, step = 0;
        for (;;) {
            const uint64_t idle = __ballot(!active);
            if (idle != 0 && !exhausted) {  // wave-uniform
                const uint32_t n_idle = (uint32_t)__popcll(idle);
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&s_off, n_idle);
                base = (uint32_t)wave_broadcast_first(base);
                if (base >= len) {
                    exhausted = true;
                } else {
                    const uint32_t take = len - base < n_idle ? len - base : n_idle;
                    const uint32_t rank = lane_rank(idle);
                    if (!active && rank < take) {
                        w = start + base + rank;
                        v = src;
                        cur_beg = src_beg;
                        cur_deg = src_deg;
                        step = 0;
                        if (!ends) ++my_src;  // the start visit
                        active = true;
                    }
                }
            }
            if (__ballot(active) == 0) break;
            if (active) {
                const u32x4 r = philox4x32_10((uint32_t)w, (uint32_t)(w >> 32), step, stream, seed_lo, seed_hi);
                const bool coin = textbook || step >= 1;
                bool stop = coin && r.z >= thr;
                if (!stop) {
                    const uint32_t beg = nrec ? cur_beg : p.row_ptr[v];
                    const uint32_t deg = nrec ? cur_deg : p.row_ptr[v + 1] - beg;
                    if (deg == 0) {
                        stop = true;  // batches end walkers at vertices without out-edges
                    } else {
                        const uint32_t j = mulhi64(r.x, r.y, deg);
                        if (wedge) {
                            const uint4 e = wedge[beg + j];
                            v = r.w >= e.y ? e.z : e.x;
                        } else if (nrec) {
                            const uint4 e = nrec[(uint64_t)beg + j];
                            v = e.x;
                            cur_beg = e.y;
                            cur_deg = e.z;
                        } else {
                            v = p.col[beg + j];
                        }
                        ++step;
                        if (!ends) count(v);
                    }
                }
                if (stop) {
                    if (ends) count(v);
                    my_steps += step;
                    active = false;
                }
            }
            if (logging) log_append();
        }
    }
    flush();
}

// A lane's share of a query's count statistics (TopkStats): counts of 1, 2
// and 3 (most entries) in registers, larger counts in the workgroup's LDS
// bins; c ln c in 2^-20 fixed point, so the sums do not depend on the order
// of the additions.
__device__ __forceinline__ unsigned long long clogc_fixed(uint32_t c)
{
    return c > 1 ? (unsigned long long)((double)c * log((double)c) * 1048576.0 + 0.5) : 0ull;
}
// c ln c in 2^-20 fixed point for counts below kClogcTable, from the same
// expression, for kernels whose lanes meet scattered counts (a table read
// instead of a divergent double-precision log).
constexpr uint32_t kClogcTable = 512;
__device__ __forceinline__ void fill_clogc_table(unsigned long long *tab)
{
    for (uint32_t c = threadIdx.x; c < kClogcTable; c += blockDim.x) tab[c] = clogc_fixed(c);
}

struct LaneStats {
    uint32_t maxc = 0, n1 = 0, n2 = 0, n3 = 0;
    unsigned long long sum = 0, clogc = 0;
    __device__ __forceinline__ void add(uint32_t c, uint32_t *s_hist, const unsigned long long *tab = nullptr)
    {
        maxc = c > maxc ? c : maxc;
        sum += c;
        clogc += tab && c < kClogcTable ? tab[c] : clogc_fixed(c);
        if (c >= 4) {
            atomicAdd(&s_hist[count_bin(c)], 1u);
        } else {
            n1 += c == 1;
            n2 += c == 2;
            n3 += c == 3;
        }
    }
};

// Adds the workgroup's statistics of one task to the query's TopkStats and
// clears the LDS bins.  Workgroup-uniform: every thread calls it.
__device__ void flush_stats(LaneStats &ls, uint32_t *s_hist, uint32_t *s_max, unsigned long long *s_sums,
                            TopkStats *st)
{
    const uint32_t lane = threadIdx.x & 63;
    for (int o = 32; o > 0; o >>= 1) {
        ls.maxc = max(ls.maxc, (uint32_t)__shfl_xor(ls.maxc, o));
        ls.sum += __shfl_xor(ls.sum, o);
        ls.clogc += __shfl_xor(ls.clogc, o);
        ls.n1 += __shfl_xor(ls.n1, o);
        ls.n2 += __shfl_xor(ls.n2, o);
        ls.n3 += __shfl_xor(ls.n3, o);
    }
    if (lane == 0) {
        if (ls.maxc) atomicMax(s_max, ls.maxc);
        if (ls.sum) atomicAdd(&s_sums[0], ls.sum);
        if (ls.clogc) atomicAdd(&s_sums[1], ls.clogc);
        if (ls.n1) atomicAdd(&s_hist[1], ls.n1);
        if (ls.n2) atomicAdd(&s_hist[2], ls.n2);
        if (ls.n3) atomicAdd(&s_hist[3], ls.n3);
    }
    ls = LaneStats{};
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kCountBins; b += blockDim.x) {
        const uint32_t h = s_hist[b];
        if (h) {
            atomicAdd(&st->hist[b], h);
            s_hist[b] = 0;
        }
    }
    if (threadIdx.x == 0) {
        if (*s_max) atomicMax(&st->maxc, *s_max);
        if (s_sums[0]) atomicAdd(&st->sum, s_sums[0]);
        if (s_sums[1]) atomicAdd(&st->clogc, s_sums[1]);
        *s_max = 0;
        s_sums[0] = 0;
        s_sums[1] = 0;
    }
    __syncthreads();
}

constexpr uint32_t kCompactBlock = 256;
constexpr uint32_t kCompactRound = 1024;  // pair places a wavefront round can fill

// One round of a wavefront's compaction: each lane brings up to 16 pairs
// (valid[j]); they are ranked in lane order, staged in the wavefront's LDS
// row and written to the query's list from one reservation, as consecutive
// 8 B stores, and their counts are added to the lanes' statistics on the
// way out.  Every lane of the wavefront calls it.
__device__ __forceinline__ void compact_round(const uint2 (&e)[16], const bool (&valid)[16], uint2 *stage,
                                              uint32_t *count, uint2 *list, uint32_t cap, LaneStats &ls,
                                              uint32_t *s_hist)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t n = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) n += valid[j];
    uint32_t incl = n;  // inclusive prefix over the wavefront
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    const uint32_t total = __shfl(incl, 63);
    if (total == 0) return;  // wave-uniform
    uint32_t base = 0;
    if (lane == 63) base = atomicAdd(count, total);
    base = __shfl(base, 63);
    uint32_t at = incl - n;
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (valid[j]) stage[at++] = e[j];
    wave_lds_sync();
    for (uint32_t i = lane; i < total; i += 64) {
        const uint2 x = stage[i];
        ls.add(x.y, s_hist);
        if (base + i < cap) list[base + i] = x;
    }
    wave_lds_sync();  // the row is reused by the next round
}

// After the walk: the visited vertices of each query, listed as (device id,
// count) pairs in the query's touched list, so that top-k streams the pairs
// instead of whole slices, and the query's count statistics for top-k and
// the entropy.  One read of each slice or table: the tasks are (query,
// segment) with the segments of a dense slice kSeg counters long and those
// of a hash table 1/segs of it (a multiple of 16 slots).  Wavefronts work
// independently within a task: each takes 1024 counters (16 per lane) or
// 1024 slots (16 per lane) per round with 16 B loads, and writes the round's
// pairs coalesced through LDS with one reservation on the query's count; the
// statistics are added to the query's once per task.  A list that would
// overflow marks the query for full-slice scans (touched_count > capacity);
// the statistics cover every entry either way.
__global__ __launch_bounds__(kCompactBlock) void batch_compact_kernel(BatchLaunch p)
{
    __shared__ uint2 stage[kCompactBlock / 64][kCompactRound];
    __shared__ uint32_t s_hist[kCountBins];
    __shared__ uint32_t s_max;
    __shared__ unsigned long long s_sums[2];
    for (uint32_t b = threadIdx.x; b < kCountBins; b += blockDim.x) s_hist[b] = 0;
    if (threadIdx.x == 0) {
        s_max = 0;
        s_sums[0] = 0;
        s_sums[1] = 0;
    }
    __syncthreads();
    LaneStats ls;
    const uint32_t nq = (uint32_t)p.hdr[1];
    const uint32_t pitch = dense_pitch(p.V);
    const uint32_t segs = (pitch + kSeg - 1) / kSeg;
    const uint64_t tasks = (uint64_t)nq * segs;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    uint2 *row = stage[wave];
    for (uint64_t t = blockIdx.x; t < tasks; t += gridDim.x) {
        const uint32_t q = (uint32_t)(t / segs);
        const uint32_t seg = (uint32_t)(t % segs);
        const BatchQuery bq = p.queries[q];
        if (bq.kind == SLICE_LOG) continue;  // listed by batch_log_count_kernel
        uint2 *list = p.touched + bq.touched_off;
        uint2 e[16];
        bool valid[16];
        // counters [lo, hi) of a dense range (a dense slice, or a hash
        // query's hub slice) starting at `base` in the dense pool; lo and hi
        // are multiples of 64
        auto compact_dense = [&](uint64_t base_off, uint32_t lo, uint32_t hi) {
            const uint4 *slice4 = reinterpret_cast<const uint4 *>(p.dense + base_off);
            const uint32_t step = waves * 1024;
            // the next round's 64 B per lane is loaded before this round is
            // placed, so two rounds of loads are in flight per wavefront
            uint4 nx[4];
            uint32_t base = lo + wave * 1024;
            {
                const uint32_t my0 = base + lane * 16;  // my0 < hi implies my0 + 16 <= hi
#pragma unroll
                for (int j = 0; j < 4; ++j) nx[j] = my0 < hi ? slice4[(my0 >> 2) + j] : uint4{0u, 0u, 0u, 0u};
            }
            for (; base < hi; base += step) {  // wave-uniform
                const uint32_t my0 = base + lane * 16;
                uint4 cur[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) cur[j] = nx[j];
                const uint32_t nmy0 = my0 + step;
#pragma unroll
                for (int j = 0; j < 4; ++j) nx[j] = nmy0 < hi ? slice4[(nmy0 >> 2) + j] : uint4{0u, 0u, 0u, 0u};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    e[4 * j] = uint2{my0 + 4 * j, cur[j].x};
                    e[4 * j + 1] = uint2{my0 + 4 * j + 1, cur[j].y};
                    e[4 * j + 2] = uint2{my0 + 4 * j + 2, cur[j].z};
                    e[4 * j + 3] = uint2{my0 + 4 * j + 3, cur[j].w};
                }
#pragma unroll
                for (int j = 0; j < 16; ++j) valid[j] = e[j].y != 0;
                compact_round(e, valid, row, &p.touched_count[q], list, bq.touched_cap, ls, s_hist);
            }
        };
        if (bq.kind == SLICE_DENSE) {
            const uint32_t lo = seg * kSeg;
            compact_dense(bq.slice_off, lo, lo + kSeg < pitch ? lo + kSeg : pitch);
        } else {
            if ((uint64_t)seg * kSeg < bq.hub_len) {
                const uint32_t lo = seg * kSeg;
                compact_dense(bq.hub_off, lo, lo + kSeg < bq.hub_len ? lo + kSeg : bq.hub_len);
            }
            // tables hold a power of two >= 2 slots and start at an even
            // slot, so 16 B loads of slot pairs are aligned
            const uint64_t cap = (uint64_t)bq.mask + 1;
            const uint64_t span = ((cap + segs - 1) / segs + 15) & ~15ull;
            const uint64_t lo = (uint64_t)seg * span;
            const uint64_t hi = lo + span < cap ? lo + span : cap;
            const uint2 *tab = p.hash + bq.slice_off;
            for (uint64_t base = lo + wave * 1024; base < hi; base += waves * 1024) {  // wave-uniform
                const uint64_t my0 = base + lane * 16;
                if (my0 + 16 <= hi) {
                    const uint4 *t4 = reinterpret_cast<const uint4 *>(tab + my0);
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const uint4 x = t4[j];
                        e[2 * j] = uint2{x.x, x.y};
                        e[2 * j + 1] = uint2{x.z, x.w};
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 16; ++j) e[j] = my0 + j < hi ? tab[my0 + j] : uint2{kEmpty, 0u};
                }
#pragma unroll
                for (int j = 0; j < 16; ++j) valid[j] = e[j].x != kEmpty && e[j].y != 0;
                compact_round(e, valid, row, &p.touched_count[q], list, bq.touched_cap, ls, s_hist);
            }
        }
        flush_stats(ls, s_hist, &s_max, s_sums, p.tk_stats + q);
    }
}

// ---- visit log of SLICE_LOG queries ------------------------------------------

// Exclusive prefix of x over the workgroup (at most 1024 threads) and the
// total.  Every thread calls it; s_wave holds 16 words.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t x, unsigned long long *s_wave, uint64_t &total)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = (blockDim.x + 63) >> 6;
    uint64_t incl = x;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint64_t before = 0;
    total = 0;
    for (uint32_t w = 0; w < waves; ++w) {
        const uint64_t t = s_wave[w];
        before += w < wave ? t : 0;
        total += t;
    }
    __syncthreads();  // s_wave is reused by the next scan
    return before + incl - x;
}

// Per logged query: the exclusive prefix of its bucket entries (the bucket's
// first entry in the query's partition, and the scatter cursor) and its
// entry total.
__global__ __launch_bounds__(1024) void batch_log_rows_kernel(BatchLaunch p)
{
    __shared__ unsigned long long s_wave[16];
    const uint32_t nq = (uint32_t)p.hdr[1], nb = p.log_nb;
    const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
    for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
        if (p.queries[q].kind != SLICE_LOG) continue;  // workgroup-uniform
        const uint64_t row = (uint64_t)q * nb;
        const uint32_t b0 = threadIdx.x * per;
        uint64_t mine = 0;
        for (uint32_t b = b0; b < b0 + per && b < nb; ++b) mine += p.part_cnt[row + b];
        uint64_t total = 0;
        uint64_t at = block_excl_scan(mine, s_wave, total);
        if (total >> 32) {
            // u32 places cannot hold it (no walk comes near: the plan keeps
            // every query's expected counts below 2^32 / 1.5): the query is
            // marked as failed and its log is not partitioned
            for (uint32_t b = b0; b < b0 + per && b < nb; ++b) p.part_cnt[row + b] = 0;
            if (threadIdx.x == 0) {
                atomicAdd(&p.overflow[q], 1u);
                p.qlog[(uint64_t)q * kBatchLogQWords] = 0;
                p.qlog[(uint64_t)q * kBatchLogQWords + 5] = 0;
            }
            continue;
        }
        for (uint32_t b = b0; b < b0 + per && b < nb; ++b) {
            p.part_loc[row + b] = (uint32_t)at;
            p.part_cur[row + b] = (uint32_t)at;
            at += p.part_cnt[row + b];
        }
        if (threadIdx.x == 0) p.qlog[(uint64_t)q * kBatchLogQWords + 5] = (uint32_t)total;
    }
}

// One workgroup: places every logged query's partition, segment list and
// pair list one after the other (exclusive prefixes over the queries), and
// gives the pair list its place in the touched pool (from hdr[7]) and its
// capacity: a bucket holds at most its entries plus its hub counters plus
// the source as distinct vertices.
__global__ __launch_bounds__(1024) void batch_log_plan_kernel(BatchLaunch p)
{
    __shared__ unsigned long long s_wave[16];
    const uint32_t nq = (uint32_t)p.hdr[1];
    BatchQuery *qs = const_cast<BatchQuery *>(p.queries);
    uint64_t c_ent = 0, c_seg = 0, c_pair = 0;
    for (uint32_t base = 0; base < nq; base += blockDim.x) {
        const uint32_t q = base + threadIdx.x;
        uint64_t ent = 0, segs = 0, pairs = 0;
        uint32_t *ql = p.qlog + (uint64_t)q * kBatchLogQWords;
        bool logged = false;
        if (q < nq && qs[q].kind == SLICE_LOG) {
            logged = true;
            ent = ql[5];
            segs = ql[0];
            pairs = ent + qs[q].hub_len + 1;
            if (pairs > p.V) pairs = p.V;
        }
        uint64_t t_ent, t_seg, t_pair;
        const uint64_t x_ent = block_excl_scan(ent, s_wave, t_ent);
        const uint64_t x_seg = block_excl_scan(segs, s_wave, t_seg);
        const uint64_t x_pair = block_excl_scan(pairs, s_wave, t_pair);
        if (logged) {
            p.q_base[q] = c_ent + x_ent;
            ql[2] = (uint32_t)(c_seg + x_seg);
            ql[1] = (uint32_t)(c_seg + x_seg);
            qs[q].touched_off = p.hdr[7] + c_pair + x_pair;
            qs[q].touched_cap = (uint32_t)pairs;
        }
        c_ent += t_ent;
        c_seg += t_seg;
        c_pair += t_pair;
    }
}

// Lists every claimed segment under its query.
__global__ void batch_log_seglist_kernel(BatchLaunch p)
{
    uint64_t used = p.queue[2];
    if (used > p.hdr[8]) used = p.hdr[8];
    if (used > p.log_segs) used = p.log_segs;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < used; s += stride) {
        const uint32_t q = p.seg_q[s];
        const uint32_t at = atomicAdd(&p.qlog[(uint64_t)q * kBatchLogQWords + 1], 1u);
        p.seg_list[at] = (uint32_t)s;
    }
}

// The logged queries after the walk, one query per workgroup at a time,
// taken from queue[qword] in the order of log_order (largest first).  The
// workgroup owns the query, so cursors and pair places are LDS values, not
// device atomics.
//
// batch_log_partition_kernel: tiles of kBatchLogTile of the query's
// segments (thread t holds entry t of each) are counted per id bucket; each
// bucket's run continues at the bucket's LDS cursor, the tile is ordered by
// bucket in LDS and goes out as u16 offsets (id - bucket start).
//
// batch_log_count_kernel: bucket by bucket, 2^log_shift u32 LDS counters
// start from the hub slice where it covers the bucket, take the source's
// register-counted visits and the bucket's entries, and the nonzero ones are
// written in id order as the bucket's run of (vertex, count) pairs, after the
// runs of the buckets before it.  The run's place and length are kept for
// lookups; the counts go into the query's statistics.

// Workgroup barrier that orders LDS accesses only: global loads and stores
// in flight stay in flight across it (the kernels using it exchange nothing
// through global memory within a launch).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// block_excl_scan with LDS-only barriers.
__device__ __forceinline__ uint64_t block_excl_scan_lds(uint64_t x, unsigned long long *s_wave, uint64_t &total)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = (blockDim.x + 63) >> 6;
    uint64_t incl = x;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) s_wave[wave] = incl;
    lds_barrier();
    uint64_t before = 0;
    total = 0;
    for (uint32_t w = 0; w < waves; ++w) {
        const uint64_t t = s_wave[w];
        before += w < wave ? t : 0;
        total += t;
    }
    lds_barrier();  // s_wave is reused by the next scan
    return before + incl - x;
}

// Next logged query for this workgroup (kEmpty when none is left), as a
// wave-uniform value.  Every thread calls it.
__device__ __forceinline__ uint32_t take_logged_query(const BatchLaunch &p, uint32_t qword, uint32_t *s_take)
{
    lds_barrier();  // the previous query is done with the LDS
    if (threadIdx.x == 0) {
        const unsigned long long i = atomicAdd(&p.queue[qword], 1ull);
        *s_take = i < p.hdr[9] ? p.log_order[i] : kEmpty;
    }
    lds_barrier();
    return __builtin_amdgcn_readfirstlane(*s_take);
}

// Segments whose ids and fills a partition workgroup stages in LDS at once.
constexpr uint32_t kPartSegBatch = 1024;
// Dynamic LDS of a partition workgroup with tiles of `tile` segments and nb
// id buckets, and the most it may take (160 KiB less its static variables).
constexpr size_t kPartitionLdsBytes = 163840 - 256;
inline size_t partition_lds(uint32_t tile, uint32_t nb)
{
    return ((size_t)tile * kBatchLogSeg + 5 * (size_t)nb + 2 * kPartSegBatch) * 4;
}
static_assert(((size_t)kBatchLogTile / 2 * kBatchLogSeg + 5 * (size_t)kBatchLogMaxBuckets + 2 * kPartSegBatch) * 4 <=
                  kPartitionLdsBytes,
              "half tiles fit beside the most buckets");

// Dynamic LDS: the tile's staging (TILE * kBatchLogSeg words), five arrays of
// log_nb words, then the ids and fills of up to kPartSegBatch of the query's
// segments.  The next tile's entries are loaded while a tile is placed.  A
// tile costs about 7 us beside its entries (barriers, the bucket scan, the
// wait for the next tile), so tiles are as large as the LDS allows.
template <uint32_t TILE>
__global__ __launch_bounds__(1024) void batch_log_partition_kernel(BatchLaunch p)
{
    extern __shared__ uint32_t lp_lds[];
    __shared__ unsigned long long s_wave[16];
    __shared__ uint32_t s_take;
    const uint32_t nb = p.log_nb, shift = p.log_shift;
    const uint32_t off_mask = (1u << shift) - 1;
    uint32_t *const stage = lp_lds;
    uint32_t *const hist = stage + TILE * kBatchLogSeg;
    uint32_t *const start = hist + nb;
    uint32_t *const fill = start + nb;
    uint32_t *const base = fill + nb;
    uint32_t *const cur = base + nb;
    uint32_t *const seg_id = cur + nb;
    uint32_t *const seg_fill = seg_id + kPartSegBatch;
    const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
    for (;;) {
        const uint32_t q = take_logged_query(p, 3, &s_take);
        if (q == kEmpty) break;
        const uint32_t *ql = p.qlog + (uint64_t)q * kBatchLogQWords;
        const uint32_t nseg = ql[0], seg0 = ql[2];
        const uint64_t qb = p.q_base[q];
        const uint64_t row = (uint64_t)q * nb;
        for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) cur[b] = p.part_loc[row + b];
        for (uint32_t g0 = 0; g0 < nseg; g0 += kPartSegBatch) {
            const uint32_t ng = nseg - g0 < kPartSegBatch ? nseg - g0 : kPartSegBatch;
            lds_barrier();  // the previous group's ids are no longer read
            for (uint32_t i = threadIdx.x; i < ng; i += blockDim.x) {
                const uint32_t sg = p.seg_list[seg0 + g0 + i];
                seg_id[i] = sg;
                seg_fill[i] = p.seg_n[sg];
            }
            lds_barrier();
            // entry threadIdx.x of each segment of tile t0 (kNoVisit past its fill)
            auto load_tile = [&](uint32_t t0, uint32_t (&e)[TILE]) {
#pragma unroll
                for (uint32_t k = 0; k < TILE; ++k) {
                    e[k] = kNoVisit;
                    if (t0 + k < ng && threadIdx.x < seg_fill[t0 + k])
                        e[k] = p.vlog[(uint64_t)seg_id[t0 + k] * kBatchLogSeg + threadIdx.x];
                }
            };
            uint32_t nxt[TILE];
            load_tile(0, nxt);
            for (uint32_t t0 = 0; t0 < ng; t0 += TILE) {
                uint32_t ent[TILE];
#pragma unroll
                for (uint32_t k = 0; k < TILE; ++k) ent[k] = nxt[k];
                for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) hist[b] = 0;
                lds_barrier();
#pragma unroll
                for (uint32_t k = 0; k < TILE; ++k)
                    if (ent[k] != kNoVisit) atomicAdd(&hist[ent[k] >> shift], 1u);
                if (t0 + TILE < ng) load_tile(t0 + TILE, nxt);  // in flight meanwhile
                lds_barrier();
                const uint32_t b0 = threadIdx.x * per;
                uint64_t mine = 0;
                for (uint32_t b = b0; b < b0 + per && b < nb; ++b) mine += hist[b];
                uint64_t n_tile = 0;
                uint32_t at = (uint32_t)block_excl_scan_lds(mine, s_wave, n_tile);
                for (uint32_t b = b0; b < b0 + per && b < nb; ++b) {
                    const uint32_t h = hist[b];
                    start[b] = at;
                    fill[b] = 0;
                    base[b] = cur[b];
                    cur[b] += h;
                    at += h;
                }
                lds_barrier();
#pragma unroll
                for (uint32_t k = 0; k < TILE; ++k) {
                    if (ent[k] != kNoVisit) {
                        const uint32_t b = ent[k] >> shift;
                        const uint32_t pos = start[b] + atomicAdd(&fill[b], 1u);
                        stage[pos] = (b << 16) | (ent[k] & off_mask);
                    }
                }
                lds_barrier();
                for (uint32_t i = threadIdx.x; i < (uint32_t)n_tile; i += blockDim.x) {
                    const uint32_t e = stage[i];
                    const uint32_t b = e >> 16;
                    p.part16[qb + base[b] + (i - start[b])] = (uint16_t)(e & 0xFFFFu);
                }
                lds_barrier();  // stage, hist and the cursors are reused by the next tile
            }
        }
    }
}

// Dynamic LDS: 2^log_shift u32 counters, a bitmap of the nonzero ones, then
// the query's entry count and first entry per bucket (2 * log_nb words).
// The counters are zero between buckets: a bucket sets the counters and bits
// of its hub values, source visits and entries, so it costs its entries, not
// its span.  Output: wavefront w owns the bucket's w-th run of
// kBatchRunIds ids (one bitmap word per lane), reads and clears each set
// counter and writes the run's pairs in id order at a place it reserves on
// the query's LDS pair cursor; runs of one query land in any order, which
// the pair list allows (top-k reads it whole, lookups search a run).  Two
// barriers per bucket; the next nonempty bucket's entries are loaded while
// this one is counted.  512-thread workgroups: two fit a CU, and one's
// barriers overlap the other's work.  Small buckets come last, one per
// wavefront (below).
constexpr uint32_t kLogCountBlock = 512;
// Buckets of at most kSmallBucket entries, with no hub counter and not the
// source's, are counted by one wavefront each, without LDS counters or
// workgroup barriers: a sort of their entries in registers (kSmallRegs per
// lane; one where the bucket has at most 64 entries), then the runs of equal
// ids.  On 4096 PPR queries on R-MAT scale 24 the counting took 9.4 ms
// without this path and 7.8, 7.5, 8.0 and 10.5 ms with 1, 2, 4 and 8 keys per
// lane (with every query logged: 8.0 ms at 2 keys, 7.8 ms with 1 key for the
// buckets of at most 64 entries).
constexpr uint32_t kSmallRegs = 2;
constexpr uint32_t kSmallBucket = 64 * kSmallRegs;
static_assert((1u << kBatchLogShift) / 32 == kLogCountBlock, "one bitmap word per thread");

// Ascending bitonic sort of the wavefront's 64 * R keys, element k * 64 +
// lane in key[k]: partners 64 or more apart are in the same lane's
// registers, nearer ones in another lane of the same register.
template <uint32_t R>
__device__ __forceinline__ void wave_sort(uint32_t (&key)[R])
{
    static_assert((R & (R - 1)) == 0, "a power of two of registers");
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (uint32_t size = 2; size <= 64 * R; size <<= 1) {
#pragma unroll
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= 64) {
                const uint32_t d = stride >> 6;
#pragma unroll
                for (uint32_t k = 0; k < R; ++k) {
                    if (k & d) continue;  // k is the lower element of the pair (k, k | d)
                    const bool asc = ((k * 64 + lane) & size) == 0;
                    const uint32_t a = key[k], b = key[k | d];
                    key[k] = asc ? min(a, b) : max(a, b);
                    key[k | d] = asc ? max(a, b) : min(a, b);
                }
            } else {
#pragma unroll
                for (uint32_t k = 0; k < R; ++k) {
                    const uint32_t other = (uint32_t)__shfl_xor(key[k], stride);
                    const bool asc = ((k * 64 + lane) & size) == 0;
                    const bool lower = (lane & stride) == 0;
                    key[k] = (lower == asc) ? min(key[k], other) : max(key[k], other);
                }
            }
        }
    }
}
static_assert(kBatchRunIds == 64 * 32, "a wavefront's run is one bitmap word per lane");
__global__ __launch_bounds__(kLogCountBlock) void batch_log_count_kernel(BatchLaunch p)
{
    extern __shared__ uint32_t lc_lds[];
    __shared__ uint32_t s_hist[kCountBins];
    __shared__ uint32_t s_max;
    __shared__ unsigned long long s_sums[2];
    __shared__ uint32_t s_take, s_pairs;
    __shared__ unsigned long long s_clogc[kClogcTable];
    fill_clogc_table(s_clogc);
    const uint32_t nb = p.log_nb, shift = p.log_shift;
    const uint32_t span = 1u << shift;
    const uint32_t off_mask = span - 1;
    const uint32_t runs_per_bucket = span / kBatchRunIds;
    uint32_t *const cnt = lc_lds;
    uint32_t *const bits = lc_lds + span;
    uint32_t *const b_cnt = bits + span / 32;
    uint32_t *const b_loc = b_cnt + nb;
    for (uint32_t i = threadIdx.x; i < span + span / 32; i += blockDim.x) lc_lds[i] = 0;
    for (uint32_t b = threadIdx.x; b < kCountBins; b += blockDim.x) s_hist[b] = 0;
    if (threadIdx.x == 0) {
        s_max = 0;
        s_sums[0] = 0;
        s_sums[1] = 0;
    }
    LaneStats ls;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint4 *part4 = reinterpret_cast<const uint4 *>(p.part16);
    auto mark = [&](uint32_t i, uint32_t c) {
        atomicAdd(&cnt[i], c);
        atomicOr(&bits[i >> 5], 1u << (i & 31));
    };
    for (;;) {
        const uint32_t q = take_logged_query(p, 4, &s_take);
        if (q == kEmpty) break;
        // per-query values in scalar registers: the loop below then waits on
        // no load of them (its prefetched entries stay in flight)
        const uint32_t src = __builtin_amdgcn_readfirstlane(p.queries[q].source_dev);
        const uint32_t hub_len = __builtin_amdgcn_readfirstlane(p.queries[q].hub_len);
        const uint64_t hub_off = wave_broadcast_first(p.queries[q].hub_off);
        const uint64_t touched_off = wave_broadcast_first(p.queries[q].touched_off);
        const uint32_t src_visits = __builtin_amdgcn_readfirstlane(p.qlog[(uint64_t)q * kBatchLogQWords + 4]);
        const uint64_t qb = wave_broadcast_first(p.q_base[q]);
        const uint64_t row = (uint64_t)q * nb;
        for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
            b_cnt[b] = p.part_cnt[row + b];
            b_loc[b] = p.part_loc[row + b];
        }
        if (threadIdx.x == 0) s_pairs = 0;
        lds_barrier();
        // buckets with something to count: entries, hub counters or the
        // source; the small ones (entries only) go to one wavefront each after
        // the others
        auto special = [&](uint32_t b) {
            const uint32_t lo = b << shift;
            return lo < hub_len || (src >= lo && src - lo < span);
        };
        auto small = [&](uint32_t b) { return b_cnt[b] != 0 && b_cnt[b] <= kSmallBucket && !special(b); };
        auto busy = [&](uint32_t b) { return (b_cnt[b] != 0 || special(b)) && !small(b); };
        auto next_busy = [&](uint32_t b) {
            while (b < nb && !busy(b)) ++b;
            return b;
        };
        // the first aligned 16 B chunk of bucket b's entries for this thread
        auto first_chunk = [&](uint32_t b) {
            uint4 v = uint4{0u, 0u, 0u, 0u};
            if (b < nb) {
                const uint64_t e0 = qb + b_loc[b];
                const uint64_t a0 = e0 & ~7ull;
                if (threadIdx.x < (uint32_t)((e0 + b_cnt[b] - a0 + 7) >> 3)) v = part4[(a0 >> 3) + threadIdx.x];
            }
            return v;
        };
        uint32_t b = next_busy(0);
        for (uint32_t e = 0; e < b; ++e)  // buckets before the first busy one: no runs (small ones later)
            if (threadIdx.x < runs_per_bucket && !small(e)) p.run_len[(row + e) * runs_per_bucket + threadIdx.x] = 0;
        uint4 first = first_chunk(b);
        while (b < nb) {  // workgroup-uniform
            const uint32_t lo = b << shift;
            const uint32_t hi = p.V - lo < span ? p.V : lo + span;
            const uint32_t n = b_cnt[b];
            const uint64_t e0 = qb + b_loc[b];
            const uint64_t a0 = e0 & ~7ull;
            const uint32_t chunks = (uint32_t)((e0 + n - a0 + 7) >> 3);
            // The entries first: the chunk loaded ahead is used before the next
            // bucket's load is issued, so waiting for it does not wait for that.
            auto mark_chunk = [&](uint4 v, uint32_t j) {
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
                const uint64_t at = a0 + 8ull * j;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint64_t idx = at + e;
                    if (idx >= e0 && idx < e0 + n) mark(((w[e >> 1] >> ((e & 1) * 16)) & 0xFFFFu) & off_mask, 1u);
                }
            };
            if (threadIdx.x < chunks) mark_chunk(first, threadIdx.x);
            for (uint32_t j = threadIdx.x + kLogCountBlock; j < chunks; j += kLogCountBlock)
                mark_chunk(part4[(a0 >> 3) + j], j);
            const uint32_t nxt = next_busy(b + 1);
            first = first_chunk(nxt);  // in flight while this bucket is finished
            // hub counters [lo, hub_end) of the bucket (the log holds none of
            // those ids, the source's visits are in src_visits), eight loads
            // in flight per thread
            const uint32_t hub_end = hub_len < hi ? hub_len : hi;
            const uint32_t hub_n = hub_end > lo ? hub_end - lo : 0u;
            for (uint32_t i0 = 0; i0 < hub_n; i0 += 8 * kLogCountBlock) {
                uint32_t c[8];
#pragma unroll
                for (uint32_t k = 0; k < 8; ++k) {
                    const uint32_t i = i0 + k * kLogCountBlock + threadIdx.x;
                    c[k] = i < hub_n ? p.dense[hub_off + lo + i] : 0u;
                }
#pragma unroll
                for (uint32_t k = 0; k < 8; ++k)
                    if (c[k]) mark(i0 + k * kLogCountBlock + threadIdx.x, c[k]);
            }
            if (src >= lo && src < hi && threadIdx.x == 0 && src_visits) mark(src - lo, src_visits);
            lds_barrier();
            // output: wavefront w's run, lane l's bitmap word
            uint32_t word = bits[threadIdx.x];
            const uint32_t pc = (uint32_t)__popc(word);
            uint32_t incl = pc;
#pragma unroll
            for (uint32_t o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            const uint32_t run_n = __shfl(incl, 63);
            uint32_t run_at = 0;
            if (lane == 0 && run_n) run_at = atomicAdd(&s_pairs, run_n);
            run_at = __shfl(run_at, 0);
            uint2 *out = p.touched + touched_off + run_at + (incl - pc);
            while (word) {
                const uint32_t i = 32 * threadIdx.x + (uint32_t)(__ffs(word) - 1);
                word &= word - 1;
                const uint32_t c = cnt[i];
                cnt[i] = 0;
                *out++ = uint2{lo + i, c};
                ls.add(c, s_hist, s_clogc);
            }
            bits[threadIdx.x] = 0;
            if (lane == 0) {
                const uint64_t ri = (row + b) * runs_per_bucket + wave;
                p.run_loc[ri] = run_at;
                p.run_len[ri] = run_n;
            }
            lds_barrier();  // the counters and bits are clear for the next bucket
            // buckets skipped on the way: no runs (small ones later)
            for (uint32_t e = b + 1; e < nxt; ++e)
                if (threadIdx.x < runs_per_bucket && !small(e)) p.run_len[(row + e) * runs_per_bucket + threadIdx.x] = 0;
            b = nxt;
        }
        // small buckets, one per wavefront: sort the entries, count the runs
        // of equal ids, write the distinct ids in order with their counts
        // one small bucket sb of n entries from e0, R keys per lane
        auto count_small = [&](auto regs, uint32_t sb, uint32_t n, uint64_t e0) {
            constexpr uint32_t R = decltype(regs)::value;
            uint32_t key[R];
#pragma unroll
            for (uint32_t k = 0; k < R; ++k) {
                const uint32_t e = k * 64 + lane;
                key[k] = e < n ? (uint32_t)p.part16[e0 + e] : 0xFFFFFFFFu;
            }
            wave_sort(key);
            // first element of each run of equal keys (padding sorts last)
            uint64_t fm[R];
#pragma unroll
            for (uint32_t k = 0; k < R; ++k) {
                // the element before: lane - 1 of this register, or lane 63 of the previous one
                const uint32_t up = (uint32_t)__shfl_up(key[k], 1);
                const uint32_t tail = k ? (uint32_t)__shfl(key[k ? k - 1 : 0], 63) : 0u;
                const uint32_t before = lane ? up : tail;
                const bool first = key[k] != 0xFFFFFFFFu && (k * 64 + lane == 0 || before != key[k]);
                fm[k] = __ballot(first);
            }
            uint32_t m = 0;
#pragma unroll
            for (uint32_t k = 0; k < R; ++k) m += (uint32_t)__popcll(fm[k]);
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&s_pairs, m);
            base = __shfl(base, 0);
            uint2 *out = p.touched + touched_off + base;
            const uint32_t lo = sb << shift;
            const uint64_t below = (lane ? (~0ull >> (64 - lane)) : 0ull);  // lanes below this one
            uint32_t rank0 = 0;  // firsts in the registers before k
#pragma unroll
            for (uint32_t k = 0; k < R; ++k) {
                if ((fm[k] >> lane) & 1ull) {
                    // the next first element (or n) ends this run
                    uint32_t next = n;
                    const uint64_t above = lane < 63 ? fm[k] >> (lane + 1) : 0ull;
                    if (above) {
                        next = k * 64 + lane + 1 + (uint32_t)__ffsll((long long)above) - 1;
                    } else {
#pragma unroll
                        for (uint32_t k2 = R - 1; k2 > k; --k2)
                            if (fm[k2]) next = k2 * 64 + (uint32_t)__ffsll((long long)fm[k2]) - 1;
                    }
                    const uint32_t c = next - (k * 64 + lane);
                    out[rank0 + (uint32_t)__popcll(fm[k] & below)] = uint2{lo + key[k], c};
                    ls.add(c, s_hist, s_clogc);
                }
                rank0 += (uint32_t)__popcll(fm[k]);
            }
            // the bucket's runs of kBatchRunIds ids: place and length
            uint32_t at = 0;
#pragma unroll
            for (uint32_t r = 0; r < (1u << kBatchLogShift) / kBatchRunIds; ++r) {
                uint32_t cnt_r = 0;
#pragma unroll
                for (uint32_t k = 0; k < R; ++k)
                    cnt_r += (uint32_t)__popcll(fm[k] & __ballot(key[k] != 0xFFFFFFFFu && key[k] / kBatchRunIds == r));
                if (lane == r) {
                    const uint64_t ri = (row + sb) * runs_per_bucket + r;
                    p.run_loc[ri] = base + at;
                    p.run_len[ri] = cnt_r;
                }
                at += cnt_r;
            }
        };
        for (uint32_t sb = wave; sb < nb; sb += kLogCountBlock / 64) {
            if (!small(sb)) continue;  // wave-uniform
            const uint32_t n = b_cnt[sb];
            const uint64_t e0 = qb + b_loc[sb];
            if (n <= 64)  // wave-uniform: a one-register sort is cheaper
                count_small(std::integral_constant<uint32_t, 1>{}, sb, n, e0);
            else
                count_small(std::integral_constant<uint32_t, kSmallRegs>{}, sb, n, e0);
        }
        lds_barrier();  // every wavefront's pairs are reserved
        if (threadIdx.x == 0) p.touched_count[q] = s_pairs;
        flush_stats(ls, s_hist, &s_max, s_sums, p.tk_stats + q);
    }
}

// ---- top-k -------------------------------------------------------------------

struct Entries {
    const BatchLaunch *p;
    BatchQuery Q;
    uint32_t mode;  // 0 pair list, 1 dense full slice, 2 hash hub slice then full table
    uint32_t n;
    uint32_t V;
    __device__ __forceinline__ void get(uint32_t i, uint32_t &v, uint32_t &c) const
    {
        if (mode == 0) {
            const uint2 e = p->touched[Q.touched_off + i];
            v = e.x;
            c = e.y;
        } else if (mode == 1) {
            v = i;
            c = p->dense[Q.slice_off + i];
        } else if (i < Q.hub_len) {
            v = i;
            c = p->dense[Q.hub_off + i];
        } else {
            const uint2 s = p->hash[Q.slice_off + (i - Q.hub_len)];
            v = s.x;
            c = s.x == kEmpty ? 0u : s.y;
        }
    }
};

// A query's entries after the compaction: its pair list, or the whole slice
// or table when the list overflowed.
__device__ __forceinline__ Entries entries_of(const BatchLaunch &p, uint32_t q)
{
    Entries E;
    E.p = &p;
    E.Q = p.queries[q];
    E.V = p.V;
    const uint32_t t = p.touched_count[q];
    if (t <= E.Q.touched_cap || E.Q.kind == SLICE_LOG) {  // a logged list fits by construction
        E.mode = 0;
        E.n = t;
    } else if (E.Q.kind == SLICE_DENSE) {
        E.mode = 1;
        E.n = p.V;
    } else {
        E.mode = 2;
        E.n = E.Q.hub_len + E.Q.mask + 1;
    }
    return E;
}

// The count bin holding the k-th largest count (bin 1, the count 1, when
// fewer than k entries exist), the entries in higher bins and in it.  Every
// lane of the wavefront calls it: lane l reads bins 4l..4l+3.
struct Bucket {
    uint32_t b, above, inbin;
};
__device__ Bucket kth_bucket(const TopkStats *st, uint32_t k)
{
    static_assert(kCountBins == 256, "four bins per lane");
    const uint32_t lane = threadIdx.x & 63;
    const uint4 h4 = reinterpret_cast<const uint4 *>(st->hist)[lane];
    const uint32_t h[4] = {h4.x, h4.y, h4.z, h4.w};
    const uint32_t tot = h[0] + h[1] + h[2] + h[3];
    uint32_t suf = tot;  // entries in this lane's bins and all higher ones
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_down(suf, o);
        if (lane + o < 64) suf += y;
    }
    const uint32_t higher = suf - tot;
    const uint64_t m = __ballot(higher < k && suf >= k);
    Bucket r;
    if (m) {  // wave-uniform
        uint32_t cum = higher, b = 0, inbin = 0;
#pragma unroll
        for (int j = 3; j >= 0; --j) {
            if (cum + h[j] >= k) {
                b = 4 * lane + (uint32_t)j;
                inbin = h[j];
                break;
            }
            cum += h[j];
        }
        const int L = __ffsll((unsigned long long)m) - 1;
        r.b = __shfl(b, L);
        r.above = __shfl(cum, L);
        r.inbin = __shfl(inbin, L);
    } else {
        const uint32_t total = __shfl(suf, 0);
        const uint32_t ones = __shfl(h[1], 0);
        r.b = 1;
        r.inbin = ones;
        r.above = total - ones;
    }
    return r;
}

constexpr uint32_t kTopkSeg = 32768;   // entries per (query, segment) task
constexpr uint32_t kTopkBlock = 512;  // workgroup size of the task kernels

__device__ __forceinline__ uint32_t topk_segments(const Entries &E)
{
    return (uint32_t)(((uint64_t)E.n + kTopkSeg - 1) / kTopkSeg);
}

// Calls fn(q, E, s) for this workgroup's share of the (query, segment)
// tasks: segments of kTopkSeg entries numbered across the queries in order,
// task t going to workgroup t mod gridDim.x, so a large query is spread over
// many workgroups.  Every thread calls it; fn is called workgroup-uniformly.
template <typename F>
__device__ void for_topk_segments(const BatchLaunch &p, F &&fn)
{
    __shared__ uint32_t s_wave[kTopkBlock / 64];
    __shared__ uint32_t s_n;
    __shared__ uint32_t s_q[kTopkBlock], s_first[kTopkBlock];
    const uint32_t nq = (uint32_t)p.hdr[1];
    const uint32_t G = gridDim.x, me = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    uint64_t carry = 0;  // tasks of the queries before this round's
    for (uint32_t base = 0; base < nq; base += blockDim.x) {
        const uint32_t q = base + threadIdx.x;
        const uint32_t segs = q < nq ? topk_segments(entries_of(p, q)) : 0u;
        uint32_t incl = segs;
#pragma unroll
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == 63) s_wave[wave] = incl;
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < waves; ++w) {
            before += w < wave ? s_wave[w] : 0u;
            total += s_wave[w];
        }
        const uint64_t t0 = carry + before + incl - segs;  // the query's first task
        const uint32_t r = (uint32_t)(t0 % G);
        const uint32_t first = me >= r ? me - r : me + G - r;  // first s with (t0 + s) mod G == me
        if (first < segs) {
            const uint32_t i = atomicAdd(&s_n, 1u);
            s_q[i] = q;
            s_first[i] = first;
        }
        __syncthreads();
        const uint32_t n = s_n;
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t qq = s_q[i];
            const Entries E = entries_of(p, qq);
            const uint32_t segs_q = topk_segments(E);
            for (uint32_t sg = s_first[i]; sg < segs_q; sg += G) fn(qq, E, sg);
        }
        carry += total;
        __syncthreads();  // the round's shared state is reused by the next
    }
}

// The candidates of every query whose entries in and above the k-th count's
// bucket fit its candidate slots: every entry with count >= 2^b, written in
// any order to the query's part of the candidate pool.
__global__ __launch_bounds__(kTopkBlock) void batch_topk_gather_kernel(BatchLaunch p)
{
    __shared__ uint32_t s_lo, s_ok;
    const uint32_t k = (uint32_t)p.hdr[2];
    const uint32_t lane = threadIdx.x & 63;
    for_topk_segments(p, [&](uint32_t q, const Entries &E, uint32_t sg) {
        TopkStats *st = p.tk_stats + q;
        if (threadIdx.x < 64) {
            const Bucket bk = kth_bucket(st, k);
            if (threadIdx.x == 0) {
                s_ok = k > 0 && st->maxc > 0 && (uint64_t)bk.above + bk.inbin <= E.Q.cand_cap;
                s_lo = (uint32_t)count_bin_lo(bk.b);
            }
        }
        __syncthreads();
        if (s_ok) {  // workgroup-uniform
            const uint32_t lo_c = s_lo;
            const uint64_t lo = (uint64_t)sg * kTopkSeg;
            const uint64_t hi = lo + kTopkSeg < E.n ? lo + kTopkSeg : (uint64_t)E.n;
            for (uint64_t b0 = lo; b0 < hi; b0 += blockDim.x) {
                const uint64_t i = b0 + threadIdx.x;
                uint32_t v = 0, c = 0;
                if (i < hi) E.get((uint32_t)i, v, c);
                const bool take = c >= lo_c && c != 0;
                const uint64_t m = __ballot(take);
                if (m) {  // wave-uniform
                    uint32_t slot = 0;
                    if (lane == 0) slot = atomicAdd(&st->ncand, (uint32_t)__popcll(m));
                    slot = (uint32_t)wave_broadcast_first(slot) + lane_rank(m);
                    if (take && slot < E.Q.cand_cap)
                        p.tk_cand[E.Q.cand_off + slot] = ((unsigned long long)c << 32) | (uint32_t)~v;
                }
            }
        }
        __syncthreads();  // s_lo and s_ok are reused by the next task
    });
}

__device__ void block_bitonic_desc(unsigned long long *a, uint32_t n)
{
    for (uint32_t size = 2; size <= n; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                const uint32_t j = i ^ stride;
                if (j > i) {
                    const bool desc = (i & size) == 0;
                    const unsigned long long x = a[i], y = a[j];
                    if (desc ? (x < y) : (x > y)) {
                        a[i] = y;
                        a[j] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// One workgroup per query: the entropy from the statistics, then the top-k.
// Gathered candidates are loaded into the LDS tile; otherwise the count
// range of the k-th count is narrowed with LDS histograms over the query's
// entries, starting from its bucket, before the candidates are collected.
__global__ __launch_bounds__(1024) void batch_topk_kernel(BatchLaunch p, const uint32_t *inv, uint32_t *out_v,
                                                         unsigned long long *out_c, uint32_t tile)
{
    __shared__ uint32_t hist[kBins];
    __shared__ uint32_t s_u32[8];
    extern __shared__ unsigned long long cand[];
    const uint32_t nq = (uint32_t)p.hdr[1];
    const uint32_t k = (uint32_t)p.hdr[2];

    for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
        const Entries E = entries_of(p, q);
        const TopkStats *st = p.tk_stats + q;
        if (threadIdx.x < 64) {
            const Bucket bk = kth_bucket(st, k);
            if (threadIdx.x == 0) {
                const double N = (double)st->sum;
                p.entropy[q] = N > 0 ? log(N) - ((double)st->clogc / 1048576.0) / N : 0.0;
                s_u32[0] = st->maxc;
                s_u32[1] = bk.b;
                s_u32[2] = bk.above;
                s_u32[3] = bk.inbin;
                s_u32[4] = st->ncand;
            }
        }
        __syncthreads();
        const uint32_t maxc = s_u32[0];
        const uint32_t b0 = s_u32[1];
        uint32_t above = s_u32[2];  // entries with count > hi
        uint32_t inbin = s_u32[3];
        const uint32_t gathered = s_u32[4];
        __syncthreads();
        if (k == 0) continue;  // entropy only
        const bool pooled = maxc == 0 || (uint64_t)above + inbin <= E.Q.cand_cap;
        if (pooled) {
            uint32_t nc = maxc ? gathered : 0u;
            if (nc > tile) nc = tile;  // cannot happen: the gather fits the slots
            for (uint32_t i = threadIdx.x; i < nc; i += blockDim.x) cand[i] = p.tk_cand[E.Q.cand_off + i];
            if (threadIdx.x == 0) s_u32[6] = nc;
        } else {
            uint64_t lo = count_bin_lo(b0), hi = count_bin_hi(b0);
            if (hi > maxc) hi = maxc;
            bool overflow_ties = false;
            // narrow [lo, hi] until the candidates (every count >= lo) fit the
            // tile or one value remains
            const bool fits = (uint64_t)above + inbin <= tile;
            while (!fits && lo <= hi) {
                const uint64_t range = hi - lo + 1;
                const uint64_t width = (range + kBins - 1) / kBins;
                for (uint32_t b = threadIdx.x; b < kBins; b += blockDim.x) hist[b] = 0;
                __syncthreads();
                // The lowest bin is not counted: it holds most entries (counts of
                // one or two visits) and would serialize the LDS atomics, yet it
                // matters only when fewer than k entries lie above it -- then the
                // next pass narrows into it.
                for (uint32_t i = threadIdx.x; i < E.n; i += blockDim.x) {
                    uint32_t v, c;
                    E.get(i, v, c);
                    if (c >= lo && c <= hi) {
                        const uint32_t bin = (uint32_t)((c - lo) / width);
                        if (bin) atomicAdd(&hist[bin], 1u);
                    }
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    uint32_t cum = above, b = 0;
                    const uint32_t nb = (uint32_t)((range + width - 1) / width);
                    for (int bb = (int)nb - 1; bb >= 0; --bb) {
                        if (cum + hist[bb] >= k || bb == 0) {
                            b = (uint32_t)bb;
                            break;
                        }
                        cum += hist[bb];
                    }
                    s_u32[1] = b;
                    s_u32[2] = cum;        // entries above bin b
                    s_u32[3] = hist[b];    // entries in bin b (bin 0: not counted)
                }
                __syncthreads();
                const uint32_t b = s_u32[1];
                const uint64_t nlo = lo + (uint64_t)b * width;
                uint64_t nhi = nlo + width - 1;
                if (nhi > hi) nhi = hi;
                above = s_u32[2];
                inbin = s_u32[3];
                lo = nlo;
                hi = nhi;
                __syncthreads();
                if (b != 0 && (uint64_t)above + inbin <= tile) break;
                if (lo == hi) {
                    // one value left whose count is not known to fit: the exact
                    // id select below takes the k - above smallest ids among it
                    overflow_ties = true;
                    break;
                }
            }
            // collect candidates: everything above hi, and the bin [lo, hi]
            uint32_t tie_limit = 0xFFFFFFFFu;  // ties at count lo == hi keep ids <= tie_limit
            if (overflow_ties) {
                // the narrowing can end on an uncounted value: count its ties, and
                // take them all when they fit beside the entries above
                if (threadIdx.x == 0) s_u32[7] = 0;
                __syncthreads();
                uint32_t ties = 0;
                for (uint32_t i = threadIdx.x; i < E.n; i += blockDim.x) {
                    uint32_t v, c;
                    E.get(i, v, c);
                    ties += c == (uint32_t)lo;
                }
                if (ties) atomicAdd(&s_u32[7], ties);
                __syncthreads();
                if ((uint64_t)above + s_u32[7] <= tile) overflow_ties = false;
                __syncthreads();
            }
            if (overflow_ties) {
                // radix select of the (k - above) smallest device ids among count == lo
                const uint32_t need = k - above;
                uint32_t prefix = 0, pmask = 0;
                const int shifts[3] = {21, 10, 0};
                const uint32_t widths[3] = {11, 11, 10};
                uint32_t rank_left = need;
                for (int pass = 0; pass < 3; ++pass) {
                    for (uint32_t bb = threadIdx.x; bb < kBins; bb += blockDim.x) hist[bb] = 0;
                    __syncthreads();
                    const uint32_t dmask = (1u << widths[pass]) - 1;
                    for (uint32_t i = threadIdx.x; i < E.n; i += blockDim.x) {
                        uint32_t v, c;
                        E.get(i, v, c);
                        if (c == (uint32_t)lo && (v & pmask) == prefix) atomicAdd(&hist[(v >> shifts[pass]) & dmask], 1u);
                    }
                    __syncthreads();
                    if (threadIdx.x == 0) {
                        uint32_t cum = 0, d = 0;
                        for (uint32_t dd = 0; dd <= dmask; ++dd) {
                            if (cum + hist[dd] >= rank_left) {
                                d = dd;
                                break;
                            }
                            cum += hist[dd];
                        }
                        s_u32[4] = d;
                        s_u32[5] = cum;
                    }
                    __syncthreads();
                    prefix |= s_u32[4] << shifts[pass];
                    pmask |= dmask << shifts[pass];
                    rank_left -= s_u32[5];
                    __syncthreads();
                }
                tie_limit = prefix;
            }
            if (threadIdx.x == 0) s_u32[6] = 0;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < E.n; i += blockDim.x) {
                uint32_t v, c;
                E.get(i, v, c);
                if (c == 0) continue;
                bool take = c > hi;
                if (!take && c >= lo && c <= hi) take = !overflow_ties || v <= tie_limit;
                if (take) {
                    const uint32_t slot = atomicAdd(&s_u32[6], 1u);
                    if (slot < tile) cand[slot] = ((unsigned long long)c << 32) | (uint32_t)~v;
                }
            }
        }
        __syncthreads();
        uint32_t nc = s_u32[6];
        if (nc > tile) nc = tile;  // cannot happen: the narrowing bounds the count
        uint32_t n2 = 1;
        while (n2 < nc) n2 <<= 1;
        for (uint32_t i = nc + threadIdx.x; i < n2; i += blockDim.x) cand[i] = 0;
        __syncthreads();
        block_bitonic_desc(cand, n2);
        for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
            uint32_t gv = PTW_INVALID_VERTEX;
            unsigned long long c = 0;
            if (i < nc) {
                const unsigned long long key = cand[i];
                const uint32_t dv = ~(uint32_t)key;
                c = key >> 32;
                gv = inv ? inv[dv] : dv;
            }
            out_v[(uint64_t)q * k + i] = gv;
            out_c[(uint64_t)q * k + i] = c;
        }
        __syncthreads();
    }
}

// ---- cleanup and export --------------------------------------------------------

// Zeroes what a batch left in its slices, over (query, segment) tasks that
// spread every slice across the grid (the compaction's segments): a dense
// slice with few touched entries has them cleared one by one from its pair
// list (a share of the list per segment); otherwise, and for hash tables,
// each task streams zeros over its segment with 16 B stores.
__global__ __launch_bounds__(256) void batch_reset_kernel(BatchLaunch p)
{
    const uint32_t nq = (uint32_t)p.hdr[1];
    const uint32_t pitch = dense_pitch(p.V);
    const uint32_t segs = (pitch + kSeg - 1) / kSeg;
    const uint64_t tasks = (uint64_t)nq * segs;
    for (uint64_t t = blockIdx.x; t < tasks; t += gridDim.x) {
        const uint32_t q = (uint32_t)(t / segs);
        const uint32_t seg = (uint32_t)(t % segs);
        const BatchQuery Q = p.queries[q];
        if (Q.kind == SLICE_DENSE) {
            const uint32_t n = p.touched_count[q];
            if (n <= Q.touched_cap && n < pitch / 16) {
                // a scattered 4 B store moves a whole line: only for sparse slices
                const uint32_t lo = (uint32_t)((uint64_t)n * seg / segs);
                const uint32_t hi = (uint32_t)((uint64_t)n * (seg + 1) / segs);
                for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x)
                    p.dense[Q.slice_off + p.touched[Q.touched_off + i].x] = 0;
            } else {
                const uint32_t lo = seg * kSeg;
                const uint32_t hi = lo + kSeg < pitch ? lo + kSeg : pitch;  // multiples of 64
                vec4u *d4 = reinterpret_cast<vec4u *>(p.dense + Q.slice_off);
                const vec4u z = {0u, 0u, 0u, 0u};
                for (uint32_t i = (lo >> 2) + threadIdx.x; i < (hi >> 2); i += blockDim.x)
                    __builtin_nontemporal_store(z, &d4[i]);
            }
        } else {  // hash or logged: the hub slice, then a hash query's table
            if ((uint64_t)seg * kSeg < Q.hub_len) {  // the hub slice, streamed
                const uint32_t lo = seg * kSeg;
                const uint32_t hi = lo + kSeg < Q.hub_len ? lo + kSeg : Q.hub_len;  // multiples of 64
                vec4u *d4 = reinterpret_cast<vec4u *>(p.dense + Q.hub_off);
                const vec4u z = {0u, 0u, 0u, 0u};
                for (uint32_t i = (lo >> 2) + threadIdx.x; i < (hi >> 2); i += blockDim.x)
                    __builtin_nontemporal_store(z, &d4[i]);
            }
            if (Q.kind == SLICE_LOG) continue;
            // tables hold a power of two >= 2 slots from an even slot: 16 B
            // stores of slot pairs are aligned
            const uint64_t cap = (uint64_t)Q.mask + 1;
            const uint64_t span = ((cap + segs - 1) / segs + 15) & ~15ull;
            const uint64_t lo = (uint64_t)seg * span;
            if (lo >= cap) continue;
            const uint64_t hi = lo + span < cap ? lo + span : cap;
            vec4u *h4 = reinterpret_cast<vec4u *>(p.hash + Q.slice_off);
            const vec4u e = {kEmpty, 0u, kEmpty, 0u};
            for (uint64_t i = (lo >> 1) + threadIdx.x; i < (hi >> 1); i += blockDim.x)
                __builtin_nontemporal_store(e, &h4[i]);
        }
    }
}

__global__ void batch_gather_kernel(BatchLaunch p, uint64_t n, const uint32_t *qidx, const uint32_t *vertices,
                                    const uint32_t *perm, uint64_t *out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const BatchQuery Q = p.queries[qidx[i]];
        const uint32_t v = perm ? perm[vertices[i]] : vertices[i];
        uint64_t c = 0;
        if (Q.kind == SLICE_DENSE) {
            c = p.dense[Q.slice_off + v];
        } else if (Q.kind == SLICE_LOG) {
            // binary search in the run of v's kBatchRunIds ids (ascending)
            const uint64_t pi = (uint64_t)qidx[i] * p.log_nb * ((1u << p.log_shift) / kBatchRunIds) +
                                v / kBatchRunIds;
            const uint2 *run = p.touched + Q.touched_off + p.run_loc[pi];
            uint32_t lo = 0, hi = p.run_len[pi];
            while (lo < hi) {
                const uint32_t mid = lo + ((hi - lo) >> 1);
                if (run[mid].x < v)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (lo < p.run_len[pi] && run[lo].x == v) c = run[lo].y;
        } else if (v < Q.hub_len) {
            c = p.dense[Q.hub_off + v];
        } else {
            uint32_t h = (v * 0x9E3779B1u) & Q.mask;
            for (uint32_t probe = 0; probe <= Q.mask; ++probe) {
                const uint2 s = p.hash[Q.slice_off + h];
                if (s.x == v) {
                    c = s.y;
                    break;
                }
                if (s.x == kEmpty) break;
                h = (h + 1) & Q.mask;
            }
        }
        out[i] = c;
    }
}

__global__ void batch_export_kernel(BatchLaunch p, uint32_t q, const uint32_t *perm, const uint32_t *inv,
                                    uint64_t *out, int phase)
{
    const BatchQuery Q = p.queries[q];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (Q.kind == SLICE_DENSE) {
        if (phase != 0) return;
        for (uint64_t g = t0; g < p.V; g += stride) out[g] = p.dense[Q.slice_off + (perm ? perm[g] : g)];
    } else if (phase == 0) {
        for (uint64_t g = t0; g < p.V; g += stride) out[g] = 0;
    } else if (Q.kind == SLICE_LOG) {
        const uint32_t n = p.touched_count[q];
        for (uint64_t i = t0; i < n; i += stride) {
            const uint2 e = p.touched[Q.touched_off + i];
            out[inv ? inv[e.x] : e.x] = e.y;
        }
    } else {
        for (uint64_t i = t0; i <= Q.mask; i += stride) {
            const uint2 s = p.hash[Q.slice_off + i];
            if (s.x != kEmpty) out[inv ? inv[s.x] : s.x] = s.y;
        }
        for (uint64_t i = t0; i < Q.hub_len && i < p.V; i += stride) {
            const uint32_t c = p.dense[Q.hub_off + i];
            if (c) out[inv ? inv[i] : i] = c;
        }
    }
}

}  // namespace

hipError_t launch_batch_begin(const BatchLaunch &p, uint32_t max_queries, hipStream_t s)
{
    uint32_t blocks = (max_queries + 255) / 256;
    if (blocks == 0) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(batch_begin_kernel, dim3(blocks), dim3(256), 0, s, p, max_queries);
    return hipGetLastError();
}

hipError_t launch_batch_walk(const BatchLaunch &p, int block, int num_cus, int walk_cus, int lds_hubs,
                             int cache_slots, hipStream_t s)
{
    const uint32_t nb = p.vlog ? p.log_nb : 0u;
    const uint32_t slots = (uint32_t)cache_slots;
    const uint32_t hubs = p.vlog ? batch_log_hubs(lds_hubs, nb, slots) : (uint32_t)lds_hubs;
    const size_t dlds = (((size_t)hubs + 1) / 2 + 2 * (size_t)slots + nb) * 4;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_dense_kernel, block, dlds) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    hipLaunchKernelGGL(batch_dense_kernel, dim3(per_cu * walk_cus), dim3(block), dlds, s, p, hubs, slots, nb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t lds = (size_t)(block / 64) * (kWaveSlots * 12 + kWaveHubs * 4);
    per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_walk_kernel, block, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    hipLaunchKernelGGL(batch_walk_kernel, dim3(per_cu * walk_cus), dim3(block), lds, s, p);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (p.vlog) {
        hipLaunchKernelGGL(batch_log_rows_kernel, dim3(num_cus * 2), dim3(1024), 0, s, p);
        hipLaunchKernelGGL(batch_log_plan_kernel, dim3(1), dim3(1024), 0, s, p);
        hipLaunchKernelGGL(batch_log_seglist_kernel, dim3(num_cus * 4), dim3(256), 0, s, p);
        // one query per workgroup at a time: grids of what fits resident
        const uint32_t tile = partition_lds(kBatchLogTile, nb) <= kPartitionLdsBytes ? kBatchLogTile : kBatchLogTile / 2;
        const size_t lp_lds = partition_lds(tile, nb);
        const size_t lc_lds = ((size_t)1 << p.log_shift) * 4 + ((size_t)1 << p.log_shift) / 8 + 8 * (size_t)nb;
        auto partition = tile == kBatchLogTile ? batch_log_partition_kernel<kBatchLogTile>
                                               : batch_log_partition_kernel<kBatchLogTile / 2>;
        int lp = 0, lc = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&lp, partition, kBatchLogSeg, lp_lds) != hipSuccess || lp < 1)
            lp = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&lc, batch_log_count_kernel, kLogCountBlock, lc_lds) !=
                hipSuccess || lc < 1)
            lc = 1;
        hipLaunchKernelGGL(partition, dim3(num_cus * lp), dim3(kBatchLogSeg), lp_lds, s, p);
        hipLaunchKernelGGL(batch_log_count_kernel, dim3(num_cus * lc), dim3(kLogCountBlock), lc_lds, s, p);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(batch_compact_kernel, dim3(num_cus * 8), dim3(kCompactBlock), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_batch_topk(const BatchLaunch &p, const uint32_t *inv, uint32_t *out_v, unsigned long long *out_c,
                             int tile, int num_cus, hipStream_t s)
{
    hipLaunchKernelGGL(batch_topk_gather_kernel, dim3(num_cus * 4), dim3(kTopkBlock), 0, s, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_topk_kernel, dim3(num_cus * 2), dim3(1024), (size_t)tile * 8, s, p, inv, out_v, out_c,
                       (uint32_t)tile);
    return hipGetLastError();
}

hipError_t launch_batch_reset(const BatchLaunch &p, int num_cus, hipStream_t s)
{
    hipLaunchKernelGGL(batch_reset_kernel, dim3(num_cus * 8), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_batch_gather(const BatchLaunch &p, uint64_t n, const uint32_t *qidx, const uint32_t *vertices,
                               const uint32_t *perm, uint64_t *out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(batch_gather_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, qidx, vertices, perm, out);
    return hipGetLastError();
}

hipError_t launch_batch_export(const BatchLaunch &p, uint32_t q, const uint32_t *perm, const uint32_t *inv,
                               uint64_t *out, hipStream_t s)
{
    hipLaunchKernelGGL(batch_export_kernel, dim3(1024), dim3(256), 0, s, p, q, perm, inv, out, 0);
    hipLaunchKernelGGL(batch_export_kernel, dim3(1024), dim3(256), 0, s, p, q, perm, inv, out, 1);
    return hipGetLastError();
}

}  // namespace ptw
