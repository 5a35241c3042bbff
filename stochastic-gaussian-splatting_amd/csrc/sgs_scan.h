// Single-workgroup exclusive scan shared by the hybrid and batched binning passes.
#pragma once

#include "sgs_internal.h"

namespace sgs {

constexpr uint32_t kScanThreads = 1024;

// Exclusive scan of n values (n from the device when n_device is set, capped at capacity):
// out[0..n] with out[n] the total; the total also goes to total_out, and overflow_out is set
// to total > overflow_limit, judged on *total64 instead when given (a 64-bit total counted
// elsewhere, so a 32-bit wrap of the scan cannot hide an overflow). One workgroup of 1024
// threads, 4096 values per iteration.
struct CountInput {
    const uint32_t* values;
    __device__ uint32_t operator()(uint32_t i) const { return values[i]; }
};

template <typename Input>
__global__ __launch_bounds__(kScanThreads) void exclusive_scan_kernel(Input in, uint32_t* __restrict__ out,
                                                                      const uint32_t* __restrict__ n_device,
                                                                      uint32_t n_host, uint32_t capacity,
                                                                      uint32_t* __restrict__ total_out,
                                                                      uint32_t* __restrict__ overflow_out,
                                                                      uint32_t overflow_limit,
                                                                      const unsigned long long* __restrict__ total64)
{
    __shared__ uint32_t wave_sums[kScanThreads / 64];
    uint32_t n = n_device ? *n_device : n_host;
    if (n > capacity) n = capacity;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += kScanThreads * 4u) {
        uint32_t v[4];
        uint32_t local = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = base + tid * 4u + (uint32_t)k;
            v[k] = i < n ? in(i) : 0u;
            local += v[k];
        }
        // inclusive wave scan of the per-thread totals
        uint32_t incl = local;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, off, 64);
            if ((int)lane >= off) incl += y;
        }
        if (lane == 63u) wave_sums[wave] = incl;
        __syncthreads();
        uint32_t wave_off = 0, block_total = 0;
        for (uint32_t w = 0; w < kScanThreads / 64; w++) {
            const uint32_t s = wave_sums[w];
            if (w < wave) wave_off += s;
            block_total += s;
        }
        uint32_t run = carry + wave_off + incl - local;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = base + tid * 4u + (uint32_t)k;
            if (i < n) out[i] = run;
            run += v[k];
        }
        carry += block_total;
        __syncthreads();
    }
    if (tid == 0) {
        out[n] = carry;
        if (total_out) *total_out = carry;
        if (overflow_out) {
            const unsigned long long t = total64 ? *total64 : (unsigned long long)carry;
            *overflow_out = t > (unsigned long long)overflow_limit ? 1u : 0u;
        }
    }
}

}  // namespace sgs
