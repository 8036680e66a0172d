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

}  // namespace jam
