// Philox4x32-10 counter-based generator and the integer-only draws the walk
// engine takes from it.  Usable from host and device code.
//
// Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11.  Each round multiplies words 0 and 2 by the
// constants M0, M1, mixes the high halves with words 1 and 3 and the round
// key, and the key advances by the Weyl constants W0, W1.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptw {

struct u32x4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                        uint32_t c3, uint32_t k0, uint32_t k1)
{
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c0;
        const uint64_t p1 = (uint64_t)M1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += W0;
        k1 += W1;
    }
    return u32x4{c0, c1, c2, c3};
}

// floor((hi * 2^32 + lo) * n / 2^64): an unbiased-to-2^-32 pick in [0, n)
// from 64 random bits, integer only, identical on host and device.
__host__ __device__ __forceinline__ uint32_t mulhi64(uint32_t lo, uint32_t hi, uint32_t n)
{
    const uint64_t a = (uint64_t)hi * n;
    const uint64_t b = ((uint64_t)lo * n) >> 32;
    return (uint32_t)((a + b) >> 32);
}

// Step counter value reserved for the start-vertex draw.
constexpr uint32_t kInitStep = 0xFFFFFFFFu;

}  // namespace ptw
