// Wavefront helpers shared by the kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

namespace cuvhs {

// The double lane k holds, for k the same in every lane: two lane reads into
// scalar registers, no LDS.
__device__ __forceinline__ double lane_value(double v, int k) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, k);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), k);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Wavefront-scope ordering of LDS stores before later loads by other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace cuvhs
