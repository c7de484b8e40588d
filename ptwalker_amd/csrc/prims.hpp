// Device-wide primitives of graph ingest (prims.hip): exclusive scans, a
// stable LSD radix sort of (u32 key, u32 value) pairs, an order-preserving
// select and a u32 max.  Each takes a caller-provided temporary buffer of
// the size its *_temp_bytes function returns and queues its kernels on the
// given stream without waiting.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace ptw {
namespace prims {

// Items per tile of the scan and sort kernels (256 threads x 16 items).
constexpr uint64_t kTileItems = 4096;

// out[i] = sum of in[0..i) (out may equal in).  n < 2^40.
size_t scan_temp_bytes(uint64_t n, size_t acc_bytes);
hipError_t exclusive_scan(const uint32_t *in, uint32_t *out, uint64_t n, void *temp, hipStream_t s);
hipError_t exclusive_scan(const uint64_t *in, uint64_t *out, uint64_t n, void *temp, hipStream_t s);
hipError_t exclusive_scan(const uint8_t *in, uint64_t *out, uint64_t n, void *temp, hipStream_t s);

// Stable sort of n < 2^32 pairs by the low `bits` (1..32) bits of the key,
// 8 bits per pass.  kin/vin are not written; kout/vout receive the result.
size_t sort_temp_bytes(uint64_t n);
hipError_t sort_pairs(const uint32_t *kin, uint32_t *kout, const uint32_t *vin, uint32_t *vout, uint64_t n, int bits,
                      void *temp, hipStream_t s);

// out[0..count) = the in[i] with flags[i] != 0, in order; *count_dev = count.
size_t select_temp_bytes(uint64_t n);
hipError_t select_flagged(const uint2 *in, const uint8_t *flags, uint2 *out, uint64_t n, uint64_t *count_dev,
                          void *temp, hipStream_t s);

// *out_dev = max over the 2n u32 words of `in` (0 for n = 0).
hipError_t max_u32(const uint2 *in, uint64_t n, uint32_t *out_dev, hipStream_t s);

// Runs every primitive above on n pseudo-random items and compares it with a
// host computation; returns 0, or -1 with the first difference in `msg`
// (-2: a HIP error, in `msg`).
int self_check(uint64_t n, int bits, uint64_t seed, hipStream_t s, std::string &msg);

}  // namespace prims
}  // namespace ptw
