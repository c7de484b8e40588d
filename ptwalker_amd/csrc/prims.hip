// Device-wide scan, stable radix sort, select and max for graph ingest.
//
// Scan: a tile of 4096 items (256 threads x 16 contiguous items) is summed,
// the tile sums are scanned by the same routine (one level per factor of
// 4096), and each tile then scans its own items from its offset.
//
// Sort: least-significant digit first, 8 bits per pass.  A pass counts each
// tile's digits (LDS atomics), scans the digit-major table of counts
// (digit d of tile t at d * tiles + t, so the scan yields every (digit,
// tile) start in the output), and scatters.  The scatter is stable: a tile
// is taken in 16 rounds of 256 consecutive items, one per thread in
// position order; within a wavefront the lanes holding the same digit are
// found with one ballot per digit bit, so an item's rank among them is a
// popcount of the lower lanes, and the wavefronts of a round add their
// per-digit counts in order before the round's running offsets advance.
#include "prims.hpp"

#include "wave.hpp"

namespace ptw {
namespace prims {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = (int)(kTileItems / kThreads);  // 16
constexpr int kWaves = kThreads / 64;

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

inline uint32_t grid_for_tiles(uint64_t tiles) { return (uint32_t)(tiles < 16384 ? (tiles ? tiles : 1) : 16384); }

template <typename U>
__device__ U block_sum(U x, U *s)
{
    s[threadIdx.x] = x;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    const U r = s[0];
    __syncthreads();
    return r;
}

// Exclusive prefix of x over the block's threads, in thread order.
template <typename U>
__device__ U block_exclusive(U x, U *s)
{
    s[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const U y = (int)threadIdx.x >= off ? s[threadIdx.x - off] : U(0);
        __syncthreads();
        s[threadIdx.x] += y;
        __syncthreads();
    }
    const U r = s[threadIdx.x] - x;
    __syncthreads();
    return r;
}

template <typename T, typename U>
__global__ __launch_bounds__(kThreads) void tile_sums_kernel(const T *in, uint64_t n, uint64_t tiles, U *sums)
{
    __shared__ U s[kThreads];
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint64_t base = tile * kTileItems;
        U acc = 0;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const uint64_t i = base + (uint64_t)j * kThreads + threadIdx.x;
            if (i < n) acc += (U)in[i];
        }
        const U t = block_sum<U>(acc, s);
        if (threadIdx.x == 0) sums[tile] = t;
    }
}

// offs == nullptr: a single tile starting at 0.  out may equal in: a thread
// reads all of its items before it writes them.
template <typename T, typename U>
__global__ __launch_bounds__(kThreads) void tile_scan_kernel(const T *in, U *out, uint64_t n, uint64_t tiles,
                                                             const U *offs)
{
    __shared__ U s[kThreads];
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint64_t first = tile * kTileItems + (uint64_t)threadIdx.x * kItems;
        U v[kItems];
        U sum = 0;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            v[k] = first + k < n ? (U)in[first + k] : U(0);
            sum += v[k];
        }
        U run = block_exclusive<U>(sum, s) + (offs ? offs[tile] : U(0));
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            if (first + k < n) out[first + k] = run;
            run += v[k];
        }
    }
}

template <typename T, typename U>
hipError_t scan_impl(const T *in, U *out, uint64_t n, void *temp, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint64_t tiles = (n + kTileItems - 1) / kTileItems;
    const uint32_t grid = grid_for_tiles(tiles);
    if (tiles == 1) {
        hipLaunchKernelGGL((tile_scan_kernel<T, U>), dim3(1), dim3(kThreads), 0, s, in, out, n, tiles,
                           (const U *)nullptr);
        return hipGetLastError();
    }
    U *sums = (U *)temp;
    void *rest = (char *)temp + align256(tiles * sizeof(U));
    hipLaunchKernelGGL((tile_sums_kernel<T, U>), dim3(grid), dim3(kThreads), 0, s, in, n, tiles, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = scan_impl<U, U>(sums, sums, tiles, rest, s)) != hipSuccess) return e;
    hipLaunchKernelGGL((tile_scan_kernel<T, U>), dim3(grid), dim3(kThreads), 0, s, in, out, n, tiles,
                       (const U *)sums);
    return hipGetLastError();
}

__global__ __launch_bounds__(kThreads) void digit_counts_kernel(const uint32_t *keys, uint64_t n, int shift,
                                                                uint32_t mask, uint64_t tiles, uint32_t *counts)
{
    __shared__ uint32_t s[256];
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        s[threadIdx.x] = 0;
        __syncthreads();
        const uint64_t base = tile * kTileItems;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const uint64_t i = base + (uint64_t)j * kThreads + threadIdx.x;
            if (i < n) atomicAdd(&s[(keys[i] >> shift) & mask], 1u);
        }
        __syncthreads();
        counts[(uint64_t)threadIdx.x * tiles + tile] = s[threadIdx.x];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kThreads) void digit_scatter_kernel(const uint32_t *kin, const uint32_t *vin,
                                                                 uint32_t *kout, uint32_t *vout, uint64_t n,
                                                                 int shift, uint32_t mask, int digit_bits,
                                                                 uint64_t tiles, const uint32_t *starts)
{
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_cnt[kWaves][256];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (uint64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        s_base[t] = starts[(uint64_t)t * tiles + tile];
        for (int r = 0; r < kItems; ++r) {
#pragma unroll
            for (int q = 0; q < kWaves; ++q) s_cnt[q][t] = 0;
            const uint64_t i = tile * kTileItems + (uint64_t)r * kThreads + t;
            const bool valid = i < n;
            const uint32_t k = valid ? kin[i] : 0u;
            const uint32_t v = valid ? vin[i] : 0u;
            const uint32_t d = (k >> shift) & mask;
            // lanes of this wavefront holding the same digit (all lanes take
            // part in every ballot)
            uint64_t peer = __ballot(valid);
            for (int b = 0; b < digit_bits; ++b) {
                const bool bit = (d >> b) & 1u;
                const uint64_t m = __ballot(bit);
                peer &= bit ? m : ~m;
            }
            __syncthreads();  // s_cnt cleared, s_base set
            const uint32_t rank = lane_rank(peer);
            if (valid && 63 - __clzll((long long)peer) == lane) s_cnt[w][d] = (uint32_t)__popcll(peer);
            __syncthreads();
            {  // thread t owns digit t: the wavefronts' starts, in order
                uint32_t run = s_base[t];
#pragma unroll
                for (int q = 0; q < kWaves; ++q) {
                    const uint32_t c = s_cnt[q][t];
                    s_cnt[q][t] = run;
                    run += c;
                }
                s_base[t] = run;
            }
            __syncthreads();
            if (valid) {
                const uint32_t pos = s_cnt[w][d] + rank;
                kout[pos] = k;
                vout[pos] = v;
            }
            __syncthreads();  // s_cnt is cleared again by the next round
        }
    }
}

__global__ void scatter_flagged_kernel(const uint2 *in, const uint8_t *flags, const uint64_t *offs, uint64_t n,
                                       uint2 *out, uint64_t *count)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (flags[i]) out[offs[i]] = in[i];
        if (i == n - 1) *count = offs[i] + (flags[i] ? 1u : 0u);
    }
}

__global__ __launch_bounds__(kThreads) void max_u32_kernel(const uint2 *in, uint64_t n, uint32_t *out)
{
    __shared__ uint32_t s[kThreads];
    uint32_t m = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint2 e = in[i];
        m = max(m, max(e.x, e.y));
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] = max(s[threadIdx.x], s[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, s[0]);
}

uint32_t grid_for_items(uint64_t n)
{
    const uint64_t b = (n + kThreads - 1) / kThreads;
    return (uint32_t)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

}  // namespace

size_t scan_temp_bytes(uint64_t n, size_t acc_bytes)
{
    const uint64_t tiles = (n + kTileItems - 1) / kTileItems;
    if (tiles <= 1) return 0;
    return align256(tiles * acc_bytes) + scan_temp_bytes(tiles, acc_bytes);
}

hipError_t exclusive_scan(const uint32_t *in, uint32_t *out, uint64_t n, void *temp, hipStream_t s)
{
    return scan_impl<uint32_t, uint32_t>(in, out, n, temp, s);
}

hipError_t exclusive_scan(const uint64_t *in, uint64_t *out, uint64_t n, void *temp, hipStream_t s)
{
    return scan_impl<uint64_t, uint64_t>(in, out, n, temp, s);
}

hipError_t exclusive_scan(const uint8_t *in, uint64_t *out, uint64_t n, void *temp, hipStream_t s)
{
    return scan_impl<uint8_t, uint64_t>(in, out, n, temp, s);
}

size_t sort_temp_bytes(uint64_t n)
{
    const uint64_t tiles = (n + kTileItems - 1) / kTileItems;
    return 2 * align256(n * 4) + align256(256 * tiles * 4) + scan_temp_bytes(256 * tiles, 4);
}

hipError_t sort_pairs(const uint32_t *kin, uint32_t *kout, const uint32_t *vin, uint32_t *vout, uint64_t n, int bits,
                      void *temp, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (n > 0xFFFFFFFFull || bits < 1 || bits > 32) return hipErrorInvalidValue;
    const uint64_t tiles = (n + kTileItems - 1) / kTileItems;
    char *p = (char *)temp;
    uint32_t *tk = (uint32_t *)p;
    p += align256(n * 4);
    uint32_t *tv = (uint32_t *)p;
    p += align256(n * 4);
    uint32_t *counts = (uint32_t *)p;
    p += align256(256 * tiles * 4);
    void *scan_tmp = p;
    const int passes = (bits + 7) / 8;
    const uint32_t grid = grid_for_tiles(tiles);
    const uint32_t *sk = kin, *sv = vin;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = 8 * pass;
        const int db = bits - shift < 8 ? bits - shift : 8;
        const uint32_t mask = (1u << db) - 1u;
        // the last pass lands in kout/vout
        uint32_t *dk = ((passes - 1 - pass) % 2 == 0) ? kout : tk;
        uint32_t *dv = ((passes - 1 - pass) % 2 == 0) ? vout : tv;
        hipLaunchKernelGGL(digit_counts_kernel, dim3(grid), dim3(kThreads), 0, s, sk, n, shift, mask, tiles, counts);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        // digits beyond the mask have zero counts: scan only the live ones
        if ((e = exclusive_scan(counts, counts, (uint64_t)(mask + 1) * tiles, scan_tmp, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(digit_scatter_kernel, dim3(grid), dim3(kThreads), 0, s, sk, sv, dk, dv, n, shift, mask, db,
                           tiles, (const uint32_t *)counts);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        sk = dk;
        sv = dv;
    }
    return hipSuccess;
}

size_t select_temp_bytes(uint64_t n) { return align256(n * 8) + scan_temp_bytes(n, 8); }

hipError_t select_flagged(const uint2 *in, const uint8_t *flags, uint2 *out, uint64_t n, uint64_t *count_dev,
                          void *temp, hipStream_t s)
{
    if (n == 0) return hipMemsetAsync(count_dev, 0, 8, s);
    uint64_t *offs = (uint64_t *)temp;
    hipError_t e = exclusive_scan(flags, offs, n, (char *)temp + align256(n * 8), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(scatter_flagged_kernel, dim3(grid_for_items(n)), dim3(kThreads), 0, s, in, flags,
                       (const uint64_t *)offs, n, out, count_dev);
    return hipGetLastError();
}

hipError_t max_u32(const uint2 *in, uint64_t n, uint32_t *out_dev, hipStream_t s)
{
    hipError_t e = hipMemsetAsync(out_dev, 0, 4, s);
    if (e != hipSuccess || n == 0) return e;
    const uint64_t b = (n + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(max_u32_kernel, dim3((uint32_t)(b < 2048 ? b : 2048)), dim3(kThreads), 0, s, in, n, out_dev);
    return hipGetLastError();
}

}  // namespace prims
}  // namespace ptw

// ---- self-check ------------------------------------------------------------

#include <algorithm>
#include <numeric>
#include <vector>

namespace ptw {
namespace prims {

namespace {

// splitmix64: the check's input generator
inline uint64_t mix(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct DevMem {
    std::vector<void *> p;
    ~DevMem()
    {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename X>
    X *get(uint64_t count)
    {
        void *q = nullptr;
        if (hipMalloc(&q, (count ? count : 1) * sizeof(X)) != hipSuccess) return nullptr;
        p.push_back(q);
        return (X *)q;
    }
};

template <typename X>
bool differ(const std::vector<X> &a, const std::vector<X> &b, const char *what, std::string &msg)
{
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i] != b[i]) {
            msg = std::string(what) + ": first difference at " + std::to_string(i) + " (device " +
                  std::to_string((unsigned long long)a[i]) + ", host " + std::to_string((unsigned long long)b[i]) +
                  ")";
            return true;
        }
    return false;
}

}  // namespace

int self_check(uint64_t n, int bits, uint64_t seed, hipStream_t s, std::string &msg)
{
#define PC_TRY(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            msg = std::string(#expr) + ": " + hipGetErrorString(e_);          \
            return -2;                                                        \
        }                                                                     \
    } while (0)
    if (n >= 0xFFFFFFFFull || bits < 1 || bits > 32) {
        msg = "self-check needs n < 2^32 - 1 and bits in 1..32";
        return -1;
    }
    uint64_t x = seed;
    const uint32_t kmask = bits == 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    std::vector<uint32_t> k(n), v(n), w32(n);
    std::vector<uint64_t> w64(n);
    std::vector<uint8_t> f(n);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t r = mix(x);
        // runs of a repeated key test stability across lanes, waves and tiles
        k[i] = (i && (r & 7) < 3) ? k[i - 1] : (uint32_t)(r >> 32) & kmask;
        v[i] = (uint32_t)i;
        w32[i] = (uint32_t)(r & 0xFFFF);
        w64[i] = r >> 20;
        f[i] = (r >> 8) & 1;
    }
    DevMem M;
    uint32_t *dk = M.get<uint32_t>(n), *dv = M.get<uint32_t>(n), *ok = M.get<uint32_t>(n), *ov = M.get<uint32_t>(n);
    uint32_t *d32 = M.get<uint32_t>(n), *dmax = M.get<uint32_t>(1);
    uint64_t *d64 = M.get<uint64_t>(n), *o64 = M.get<uint64_t>(n), *dcount = M.get<uint64_t>(1);
    uint8_t *df = M.get<uint8_t>(n);
    uint2 *dp = M.get<uint2>(n), *dsel = M.get<uint2>(n);
    const size_t tbytes = std::max({sort_temp_bytes(n), scan_temp_bytes(n, 8), select_temp_bytes(n)});
    uint8_t *tmp = M.get<uint8_t>(tbytes);
    if (!dk || !dv || !ok || !ov || !d32 || !dmax || !d64 || !o64 || !dcount || !df || !dp || !dsel || !tmp) {
        msg = "self-check: device allocation failed";
        return -2;
    }
    std::vector<uint2> pairs(n);
    for (uint64_t i = 0; i < n; ++i) pairs[i] = uint2{k[i], v[i] ^ 0x5A5A5A5Au};
    PC_TRY(hipMemcpyAsync(dk, k.data(), n * 4, hipMemcpyHostToDevice, s));
    PC_TRY(hipMemcpyAsync(dv, v.data(), n * 4, hipMemcpyHostToDevice, s));
    PC_TRY(hipMemcpyAsync(d32, w32.data(), n * 4, hipMemcpyHostToDevice, s));
    PC_TRY(hipMemcpyAsync(d64, w64.data(), n * 8, hipMemcpyHostToDevice, s));
    PC_TRY(hipMemcpyAsync(df, f.data(), n, hipMemcpyHostToDevice, s));
    PC_TRY(hipMemcpyAsync(dp, pairs.data(), n * 8, hipMemcpyHostToDevice, s));
    // device
    PC_TRY(sort_pairs(dk, ok, dv, ov, n, bits, tmp, s));
    std::vector<uint32_t> gk(n), gv(n), g32(n), gmax(1);
    std::vector<uint64_t> g64(n), gf(n), gcount(1);
    std::vector<uint2> gsel(n);
    PC_TRY(hipMemcpyAsync(gk.data(), ok, n * 4, hipMemcpyDeviceToHost, s));
    PC_TRY(hipMemcpyAsync(gv.data(), ov, n * 4, hipMemcpyDeviceToHost, s));
    PC_TRY(exclusive_scan(d32, d32, n, tmp, s));  // in place
    PC_TRY(hipMemcpyAsync(g32.data(), d32, n * 4, hipMemcpyDeviceToHost, s));
    PC_TRY(exclusive_scan(d64, o64, n, tmp, s));
    PC_TRY(hipMemcpyAsync(g64.data(), o64, n * 8, hipMemcpyDeviceToHost, s));
    PC_TRY(exclusive_scan(df, o64, n, tmp, s));
    PC_TRY(hipMemcpyAsync(gf.data(), o64, n * 8, hipMemcpyDeviceToHost, s));
    PC_TRY(select_flagged(dp, df, dsel, n, dcount, tmp, s));
    PC_TRY(hipMemcpyAsync(gsel.data(), dsel, n * 8, hipMemcpyDeviceToHost, s));
    PC_TRY(hipMemcpyAsync(gcount.data(), dcount, 8, hipMemcpyDeviceToHost, s));
    PC_TRY(max_u32(dp, n, dmax, s));
    PC_TRY(hipMemcpyAsync(gmax.data(), dmax, 4, hipMemcpyDeviceToHost, s));
    PC_TRY(hipStreamSynchronize(s));
    // host
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
    std::vector<uint32_t> hk(n), hv(n), h32(n);
    std::vector<uint64_t> h64(n), hf(n);
    for (uint64_t i = 0; i < n; ++i) {
        hk[i] = k[idx[i]];
        hv[i] = v[idx[i]];
    }
    uint32_t a32 = 0;
    uint64_t a64 = 0, af = 0;
    std::vector<uint2> hsel;
    uint32_t hmax = 0;
    for (uint64_t i = 0; i < n; ++i) {
        h32[i] = a32;
        a32 += w32[i];
        h64[i] = a64;
        a64 += w64[i];
        hf[i] = af;
        af += f[i];
        if (f[i]) hsel.push_back(pairs[i]);
        hmax = std::max(hmax, std::max(pairs[i].x, pairs[i].y));
    }
    if (differ(gk, hk, "sort keys", msg) || differ(gv, hv, "sort values (stability)", msg) ||
        differ(g32, h32, "u32 scan (in place)", msg) || differ(g64, h64, "u64 scan", msg) ||
        differ(gf, hf, "u8 -> u64 scan", msg))
        return -1;
    if (gcount[0] != hsel.size()) {
        msg = "select: count " + std::to_string(gcount[0]) + ", host " + std::to_string(hsel.size());
        return -1;
    }
    for (size_t i = 0; i < hsel.size(); ++i)
        if (gsel[i].x != hsel[i].x || gsel[i].y != hsel[i].y) {
            msg = "select: first difference at " + std::to_string(i);
            return -1;
        }
    if (gmax[0] != (n ? hmax : 0u)) {
        msg = "max: device " + std::to_string(gmax[0]) + ", host " + std::to_string(hmax);
        return -1;
    }
    return 0;
#undef PC_TRY
}

}  // namespace prims
}  // namespace ptw
