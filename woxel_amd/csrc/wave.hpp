// Wave-level helpers shared by the build kernels (64-lane waves).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wxw {

// Exclusive prefix sum over the wave's lanes of v, and the wave's total (in every
// lane). Every lane of the wave must be active.
__device__ __forceinline__ uint32_t wave_scan(uint32_t v, uint32_t &total) {
    const uint32_t lane = __lane_id();
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, (unsigned)d);
        if (lane >= (uint32_t)d) x += y;
    }
    total = __shfl(x, 63);
    return x - v;
}

}  // namespace wxw
