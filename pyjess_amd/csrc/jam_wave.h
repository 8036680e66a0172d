// jam_wave.h -- wave-level helpers (64-lane wavefronts of gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jam_common.h"

namespace jam {

__device__ static inline int lane_id() { return (int)(threadIdx.x & (JAM_WAVE - 1)); }

__device__ static inline unsigned long long ballot(bool p) { return __ballot(p); }

// number of set bits of m below this lane
__device__ static inline int rank_below(unsigned long long m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ static inline int first_bit(unsigned long long m) { return __ffsll((long long)m) - 1; }

// Orders the LDS accesses of one wave around this point (other lanes' writes
// before it are visible to reads after it); waves of 64 lanes run in lockstep,
// so no s_barrier is needed.
__device__ static inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace jam
