// Measurement kernels (include/sgs_microbench.h): the HBM copy ceiling and the random 64-bit
// atomicMin rate that the renderer's roofline figures are read against.
#include <hip/hip_runtime.h>

#include <string>

#include "sgs_microbench.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

constexpr int kBlock = 256;

// grid of 2048 workgroups x 256 lanes: 8 per CU on 256 CUs, enough to saturate HBM and the
// memory-side atomic units
constexpr int kGrid = 2048;

__global__ __launch_bounds__(kBlock) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * kBlock;
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    // four independent 16-byte loads in flight per lane
    for (; i + 3 * stride < n; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014): a bijective 64-bit mixer
__device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void atomic_min_kernel(unsigned long long* __restrict__ words, uint64_t n_words,
                                                             uint64_t n_ops, uint64_t seed)
{
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t op = (uint64_t)blockIdx.x * kBlock + threadIdx.x; op < n_ops; op += stride) {
        const uint64_t h = mix64(seed + 0x9e3779b97f4a7c15ull * (op + 1));
        // the address from the high bits by multiply-shift: always in [0, n_words)
        const uint64_t a = __umul64hi(h, n_words);
        atomicMin(&words[a], (unsigned long long)mix64(h ^ 0x5851f42d4c957f2dull));
    }
}

}  // namespace

extern "C" {

const char* sgs_mb_last_error(void) { return g_error.c_str(); }

int sgs_mb_copy(void* dst, const void* src, size_t bytes, void* stream)
{
    if (!dst || !src) return fail(-1, "NULL pointer");
    if (bytes % 16u != 0 || reinterpret_cast<uintptr_t>(dst) % 16u || reinterpret_cast<uintptr_t>(src) % 16u)
        return fail(-1, "bytes and both pointers must be multiples of 16");
    const char* s = static_cast<const char*>(src);
    const char* d = static_cast<const char*>(dst);
    if (s < d + bytes && d < s + bytes) return fail(-1, "source and destination overlap");
    if (bytes == 0) return 0;
    hipLaunchKernelGGL(copy_kernel, dim3(kGrid), dim3(kBlock), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint4*>(src), static_cast<uint4*>(dst), bytes / 16u);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail(-3, std::string("copy launch: ") + hipGetErrorString(e));
}

int sgs_mb_atomic_min_u64(uint64_t* words, uint64_t n_words, uint64_t n_ops, uint64_t seed, void* stream)
{
    if (!words || n_words == 0) return fail(-1, "words must be a non-empty device buffer");
    if (n_ops == 0) return 0;
    hipLaunchKernelGGL(atomic_min_kernel, dim3(kGrid), dim3(kBlock), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<unsigned long long*>(words), n_words, n_ops, seed);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail(-3, std::string("atomicMin launch: ") + hipGetErrorString(e));
}

}  // extern "C"
