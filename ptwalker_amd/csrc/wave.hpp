// Wavefront (wave64) helpers shared by the walk kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptw {

// Number of set bits of `mask` below this lane (wave-level exclusive prefix).
__device__ __forceinline__ uint32_t lane_rank(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Value of `x` in the first active lane, for every lane.
__device__ __forceinline__ uint64_t wave_broadcast_first(uint64_t x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Lane order within a wavefront: LDS writes of the wavefront are complete
// and visible to its other lanes after this point.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace ptw
