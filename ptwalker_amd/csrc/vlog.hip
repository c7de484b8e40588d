// Visit-log counting for gfx950: turns the log of visits that the walk
// kernel could not count in LDS into per-vertex counts.
//
// The walk appends the device id of each such visit to the log with
// coalesced stores and counts it in its id bucket (id >> shift, 2^13 to
// 2^15 ids per bucket).  Counting the log directly with device atomics
// would cost what the walk saved (one memory-side atomic per visit), so the
// log is partitioned by bucket first:
//
//   plan     one workgroup scans the bucket counts into bucket starts and
//            cuts every bucket into count tasks of at most `seg` entries;
//   scatter  each tile of 16384 entries is counting-sorted by bucket in
//            LDS, reserves its run in every bucket it holds with one atomic
//            per bucket, and writes the runs out with consecutive stores as
//            u16 offsets inside the bucket (the bucket holds the rest);
//   count    each task counts its bucket's ids in LDS u32 counters (one per
//            id of the bucket) and adds the nonzero ones to the device
//            counters: one atomic per vertex a task saw, not per visit.
//
// Order inside a bucket does not matter (integer sums), so the partition is
// not stable and the counts equal those of per-visit atomics exactly.
#include "engine.hpp"

namespace ptw {

namespace {

constexpr uint32_t kThreads = 1024;
constexpr uint32_t kScatterTile = 16384;  // entries per scatter tile (16 per thread)
constexpr uint32_t kMinTaskEntries = 16384;

// Exclusive prefix of x over the workgroup's threads in thread order;
// `total` receives the sum.  wtot holds one word per wavefront.
__device__ uint32_t block_exclusive(uint32_t x, uint32_t *wtot, uint32_t &total)
{
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t y = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(y, off);
        if (lane >= (uint32_t)off) y += t;
    }
    if (lane == 63u) wtot[w] = y;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t i = 0; i < nw; ++i) {
        const uint32_t t = wtot[i];
        if (i < w) before += t;
        all += t;
    }
    __syncthreads();  // wtot may be reused
    total = all;
    return before + y - x;
}

// plan = hist[nb] | start[nb + 1] | cursor[nb] | task_first[nb + 1] | seg.
// One workgroup of kThreads; thread t owns buckets [t * per, t * per + per).
__global__ __launch_bounds__(kThreads) void vlog_plan_kernel(uint32_t *plan, uint32_t nb, uint32_t parallel)
{
    __shared__ uint32_t wtot[kThreads / 64];
    const uint32_t *hist = plan;
    uint32_t *start = plan + nb;
    uint32_t *cursor = start + nb + 1;
    uint32_t *task_first = cursor + nb;
    uint32_t *seg = task_first + nb + 1;
    const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = threadIdx.x * per;
    const uint32_t b1 = b0 + per < nb ? b0 + per : nb;

    uint32_t local = 0;
    for (uint32_t b = b0; b < b1; ++b) local += hist[b];
    uint32_t total = 0;
    uint32_t ex = block_exclusive(local, wtot, total);
    for (uint32_t b = b0; b < b1; ++b) {
        start[b] = ex;
        cursor[b] = ex;
        ex += hist[b];
    }
    // tasks of about total / parallel entries, so every count workgroup
    // gets work; a bucket smaller than that is one task
    uint32_t s = (uint32_t)(((uint64_t)total + parallel - 1) / (parallel ? parallel : 1));
    if (s < kMinTaskEntries) s = kMinTaskEntries;
    local = 0;
    for (uint32_t b = b0; b < b1; ++b) local += (hist[b] + s - 1) / s;
    uint32_t ntasks = 0;
    ex = block_exclusive(local, wtot, ntasks);
    for (uint32_t b = b0; b < b1; ++b) {
        task_first[b] = ex;
        ex += (hist[b] + s - 1) / s;
    }
    if (threadIdx.x == 0) {
        start[nb] = total;
        task_first[nb] = ntasks;
        seg[0] = s;
    }
}

// Partition of the log's first min(head, cap) entries (a multiple of 4;
// kNoVisit entries are skipped) into `out` by bucket, runs placed from the
// plan's cursors, each entry as its offset inside its bucket (shift <= 16).
// LDS: the tile's keys, then cnt, off and gb per bucket.
__global__ __launch_bounds__(kThreads) void vlog_scatter_kernel(const uint32_t *log, const unsigned long long *head,
                                                               uint64_t cap, uint32_t shift, uint32_t nb,
                                                               uint32_t *cursor, uint16_t *out)
{
    extern __shared__ uint32_t sm[];
    __shared__ uint32_t wtot[kThreads / 64];
    uint32_t *keys = sm;
    uint32_t *cnt = keys + kScatterTile;
    uint32_t *off = cnt + nb;
    uint32_t *gb = off + nb;
    const uint64_t h = *head;
    const uint64_t n = h < cap ? h : cap;
    const uint64_t tiles = (n + kScatterTile - 1) / kScatterTile;
    const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = threadIdx.x * per;
    const uint32_t b1 = b0 + per < nb ? b0 + per : nb;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) cnt[b] = 0;
        __syncthreads();
        const uint64_t base = t * kScatterTile;
        uint32_t x[16], r[16];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t i = base + ((uint64_t)k * blockDim.x + threadIdx.x) * 4;
            uint4 q = make_uint4(kNoVisit, kNoVisit, kNoVisit, kNoVisit);
            if (i < n) q = *reinterpret_cast<const uint4 *>(log + i);  // n % 4 == 0: i + 3 < n
            x[4 * k] = q.x;
            x[4 * k + 1] = q.y;
            x[4 * k + 2] = q.z;
            x[4 * k + 3] = q.w;
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) r[j] = x[j] != kNoVisit ? atomicAdd(&cnt[x[j] >> shift], 1u) : 0u;
        __syncthreads();
        uint32_t local = 0;
        for (uint32_t b = b0; b < b1; ++b) local += cnt[b];
        uint32_t total = 0;
        uint32_t ex = block_exclusive(local, wtot, total);
        for (uint32_t b = b0; b < b1; ++b) {
            const uint32_t c = cnt[b];
            off[b] = ex;
            ex += c;
            if (c) gb[b] = atomicAdd(&cursor[b], c);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; ++j)
            if (x[j] != kNoVisit) keys[off[x[j] >> shift] + r[j]] = x[j];
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
            const uint32_t y = keys[i];
            const uint32_t b = y >> shift;
            out[gb[b] + (i - off[b])] = (uint16_t)(y & ((1u << shift) - 1u));
        }
        __syncthreads();  // the next tile reuses keys, cnt, off and gb
    }
}

// Count tasks: task t of bucket b covers entries [start[b] + k seg,
// min(start[b] + (k + 1) seg, start[b + 1])) of the partition, k = t -
// task_first[b].  LDS: 2^shift u32 counters.
__global__ __launch_bounds__(kThreads) void vlog_count_kernel(const uint16_t *parted, const uint32_t *plan, uint32_t shift,
                                                             uint32_t nb, unsigned long long *counts)
{
    extern __shared__ uint32_t c[];
    const uint32_t *start = plan + nb;
    const uint32_t *task_first = start + nb + 1 + nb;
    const uint32_t seg = task_first[nb + 1];
    const uint32_t ntasks = task_first[nb];
    const uint32_t width = 1u << shift;
    for (uint32_t t = blockIdx.x; t < ntasks; t += gridDim.x) {
        uint32_t lo = 0, hi = nb;  // task_first[lo] <= t < task_first[hi]
        while (hi - lo > 1) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (task_first[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
        const uint64_t e0 = (uint64_t)start[lo] + (uint64_t)(t - task_first[lo]) * seg;
        const uint64_t end = start[lo + 1];
        const uint64_t e1 = e0 + seg < end ? e0 + seg : end;
        for (uint32_t i = threadIdx.x; i < width; i += blockDim.x) c[i] = 0;
        __syncthreads();
        uint64_t e = e0 + threadIdx.x;
        for (; e + 3ull * blockDim.x < e1; e += 4ull * blockDim.x) {
            const uint32_t y0 = parted[e], y1 = parted[e + blockDim.x];
            const uint32_t y2 = parted[e + 2ull * blockDim.x], y3 = parted[e + 3ull * blockDim.x];
            atomicAdd(&c[y0], 1u);
            atomicAdd(&c[y1], 1u);
            atomicAdd(&c[y2], 1u);
            atomicAdd(&c[y3], 1u);
        }
        for (; e < e1; e += blockDim.x) atomicAdd(&c[parted[e]], 1u);
        __syncthreads();
        const uint64_t v0 = (uint64_t)lo << shift;
        for (uint32_t i = threadIdx.x; i < width; i += blockDim.x) {
            const uint32_t ci = c[i];
            if (ci) atomicAdd(&counts[v0 + i], (unsigned long long)ci);
        }
        __syncthreads();  // the next task clears c
    }
}

}  // namespace

hipError_t launch_vlog_begin(const VlogBuffers &b, uint32_t nb, hipStream_t s)
{
    return launch_fill_u32(b.hist(), nb, 0u, s);
}

hipError_t launch_vlog_count(const VlogBuffers &b, const unsigned long long *head, uint64_t cap, uint32_t shift,
                             uint32_t nb, unsigned long long *counts_dev, int num_cus, hipStream_t s)
{
    if (nb == 0 || nb > kVlogMaxBuckets || shift < kVlogMinShift || shift > kVlogMaxShift || nb > b.plan_buckets ||
        cap > b.cap || cap % kVlogSeg)
        return hipErrorInvalidValue;
    const uint32_t grid = (uint32_t)(2 * (num_cus > 0 ? num_cus : 1));
    hipLaunchKernelGGL(vlog_plan_kernel, dim3(1), dim3(kThreads), 0, s, b.plan, nb, grid);
    const size_t scatter_lds = ((size_t)kScatterTile + 3ull * nb) * sizeof(uint32_t);
    uint32_t *cursor = b.plan + nb + nb + 1;
    hipLaunchKernelGGL(vlog_scatter_kernel, dim3(grid), dim3(kThreads), scatter_lds, s, b.log, head, cap, shift,
                       nb, cursor, b.parted);
    const size_t count_lds = ((size_t)1 << shift) * sizeof(uint32_t);
    hipLaunchKernelGGL(vlog_count_kernel, dim3(grid), dim3(kThreads), count_lds, s, b.parted, b.plan, shift, nb,
                       counts_dev);
    return hipGetLastError();
}

}  // namespace ptw
