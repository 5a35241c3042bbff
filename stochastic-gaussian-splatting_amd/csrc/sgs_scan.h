// Single-workgroup exclusive scan shared by the hybrid and batched binning passes.
#pragma once

#include "sgs_internal.h"

namespace sgs {

constexpr uint32_t kScanThreads = 1024;

// Exclusive scan of n values (n from the device when n_device is set, capped at capacity):
// out[0..n] with out[n] the total; the total also goes to total_out, and overflow_out is set
// to total > overflow_limit, judged on the 64-bit sum of the values (so a 32-bit wrap of the
// scan cannot hide an overflow), which also goes to total64_out when given. One workgroup of
// 1024 threads, 4096 values per iteration.
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
                                                                      unsigned long long* __restrict__ total64_out)
{
    __shared__ uint32_t wave_sums[kScanThreads / 64];
    __shared__ unsigned long long wave_sums64[kScanThreads / 64];
    uint32_t n = n_device ? *n_device : n_host;
    if (n > capacity) n = capacity;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t carry = 0;
    unsigned long long sum64 = 0;  // this thread's values, summed without wrap
    for (uint32_t base = 0; base < n; base += kScanThreads * 4u) {
        uint32_t v[4];
        uint32_t local = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t i = base + tid * 4u + (uint32_t)k;
            v[k] = i < n ? in(i) : 0u;
            local += v[k];
            sum64 += v[k];
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
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)sum64, off, 64);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(sum64 >> 32), off, 64);
        sum64 += ((unsigned long long)hi << 32) | lo;
    }
    if (lane == 0u) wave_sums64[wave] = sum64;
    __syncthreads();
    if (tid == 0) {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < kScanThreads / 64; w++) t += wave_sums64[w];
        out[n] = carry;
        if (total_out) *total_out = carry;
        if (total64_out) *total64_out = t;
        if (overflow_out) *overflow_out = t > (unsigned long long)overflow_limit ? 1u : 0u;
    }
}

}  // namespace sgs
