// Block-wide exclusive scan shared by the Classic and payload kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nxg {

constexpr uint32_t kScanThreads = 256u;

// Exclusive scan of one value per thread of a 256-thread block, and the
// block's total and OR of `flags`.
static __device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t flags, uint32_t &total, uint32_t &all_flags) {
    __shared__ uint32_t sums[kScanThreads / 64u], ors[kScanThreads / 64u];
    const uint32_t lane = threadIdx.x % 64u, wave = threadIdx.x / 64u;
    uint32_t inc = v, fl = flags;
#pragma unroll
    for (uint32_t d = 1u; d < 64u; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
        fl |= __shfl_xor(fl, d, 64);
    }
    if (lane == 63u) {
        sums[wave] = inc;
        ors[wave] = fl;
    }
    __syncthreads();
    uint32_t before = 0u;
    total = 0u;
    all_flags = 0u;
#pragma unroll
    for (uint32_t w = 0u; w < kScanThreads / 64u; ++w) {
        if (w < wave) before += sums[w];
        total += sums[w];
        all_flags |= ors[w];
    }
    __syncthreads();
    return before + inc - v;
}

}  // namespace nxg
