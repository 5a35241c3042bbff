// Scene loading on the device: decoding of 3DGS PLY vertex records into the renderer's
// arrays, Morton ordering of the means, and the permutation of every array into that order.
// Replaces the array assembly of Model.py:24-75 (plyfile columns -> torch tensors, SH
// re-layout) and its morton_encode + argsort + indexing (Model.py:32-34,43,53,58,69,74).
//
// All three kernels are HBM-bound byte movement; lanes are mapped to (record, field) pairs so
// that consecutive lanes read consecutive bytes of one record and write consecutive floats
// of one output row.
#include <hipcub/hipcub.hpp>

#include "sgs_internal.h"

namespace sgs {

namespace {

constexpr uint32_t kBlock = 256;

// Output column of decoded/gathered field f (0 <= f < 11 + 3B) of one Gaussian:
// 0-2 means, 3-5 log scales, 6-9 rotation, 10 opacity, 11.. SH. PLY files store the
// degree-0 colour as f_dc_0..2 and the rest channel-major, f_rest_k with k = c*(B-1) + (b-1)
// (Model.py:61-75 reshapes to (3, B-1) and swaps the axes); fields 11.. follow the file
// order and are routed here to sh[b][c].
struct FieldDst {
    float* base;
    uint32_t row_floats;
    uint32_t col;
};

__device__ __forceinline__ FieldDst field_dst(const SceneOut& o, uint32_t f)
{
    const uint32_t B = o.sh_bases;
    if (f < 3) return {o.means, 3u, f};
    if (f < 6) return {o.scales, 3u, f - 3u};
    if (f < 10) return {o.rotations, 4u, f - 6u};
    if (f < 11) return {o.opacities, 1u, 0u};
    if (f < 14) return {o.sh, 3u * B, f - 11u};  // f_dc_c -> sh[0][c]
    const uint32_t k = f - 14u;
    const uint32_t c = k / (B - 1u), b = 1u + k % (B - 1u);
    return {o.sh, 3u * B, 3u * b + c};
}

__device__ __forceinline__ uint32_t load_bytes(const uint8_t* p, uint32_t n, bool big_endian)
{
    uint32_t v = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t byte = p[i];
        v |= big_endian ? byte << (8u * (n - 1u - i)) : byte << (8u * i);
    }
    return v;
}

// One scalar property converted to float32 the way numpy's astype(float32) does (round to
// nearest even for 32-bit integers and doubles, exact for the rest).
__device__ float load_scalar(const uint8_t* p, uint32_t type, bool big_endian)
{
    switch (type) {
    case SGS_SCALAR_INT8: return (float)(int8_t)p[0];
    case SGS_SCALAR_UINT8: return (float)p[0];
    case SGS_SCALAR_INT16: return (float)(int16_t)(uint16_t)load_bytes(p, 2, big_endian);
    case SGS_SCALAR_UINT16: return (float)(uint16_t)load_bytes(p, 2, big_endian);
    case SGS_SCALAR_INT32: return (float)(int32_t)load_bytes(p, 4, big_endian);
    case SGS_SCALAR_UINT32: return (float)load_bytes(p, 4, big_endian);
    case SGS_SCALAR_FLOAT32: return __uint_as_float(load_bytes(p, 4, big_endian));
    default: {  // SGS_SCALAR_FLOAT64
        const uint64_t lo = load_bytes(p + (big_endian ? 4 : 0), 4, big_endian);
        const uint64_t hi = load_bytes(p + (big_endian ? 0 : 4), 4, big_endian);
        return (float)__longlong_as_double((long long)(hi << 32 | lo));
    }
    }
}

// Lanes of a wave take (record, field) pairs: 64 / F records per wave, F = 11 + 3B fields.
__global__ void __launch_bounds__(kBlock) ply_decode_kernel(const uint8_t* __restrict__ records, uint64_t n_records,
                                                            PlyDecodeTable table, uint64_t first, SceneOut out)
{
    const uint32_t F = 11u + 3u * out.sh_bases;
    const uint32_t per_wave = 64u / F;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = lane / F, f = lane - r * F;
    if (r >= per_wave) return;
    const uint64_t wave = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const uint64_t rec = wave * per_wave + r;
    if (rec >= n_records) return;
    const float v = load_scalar(records + rec * table.stride + table.offset[f], table.type[f], table.big_endian != 0);
    const FieldDst d = field_dst(out, f);
    d.base[(first + rec) * d.row_floats + d.col] = v;
}

// Permutation of every array: row i of dst = row order[i] of src, same lane mapping.
__global__ void __launch_bounds__(kBlock) gather_rows_kernel(SceneArgs src, const uint32_t* __restrict__ order,
                                                             SceneOut dst)
{
    const uint32_t F = 11u + 3u * dst.sh_bases;
    const uint32_t per_wave = 64u / F;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = lane / F, f = lane - r * F;
    if (r >= per_wave) return;
    const uint64_t wave = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const uint64_t row = wave * per_wave + r;
    if (row >= dst.n) return;
    const uint64_t s = order[row];
    if (s >= src.n) return;  // an index outside the source leaves the row untouched
    // the SH row is copied as it is (no re-layout): fields >= 11 map to sh columns 0..3B-1
    const float* sbase;
    float* dbase;
    uint32_t w, col;
    if (f < 3) { sbase = src.means; dbase = dst.means; w = 3; col = f; }
    else if (f < 6) { sbase = src.scales; dbase = dst.scales; w = 3; col = f - 3; }
    else if (f < 10) { sbase = src.rotations; dbase = dst.rotations; w = 4; col = f - 6; }
    else if (f < 11) { sbase = src.opacities; dbase = dst.opacities; w = 1; col = 0; }
    else { sbase = src.sh; dbase = dst.sh; w = 3u * dst.sh_bases; col = f - 11u; }
    dbase[row * w + col] = sbase[s * w + col];
}

// Order-preserving map of float32 to uint32 (negative values reversed below positives), so
// integer min/max atomics give the float min/max.
__device__ __forceinline__ uint32_t ordered_key(float x)
{
    const uint32_t b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__host__ __device__ __forceinline__ float key_to_float(uint32_t k)
{
    const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float(b);
#else
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
#endif
}

// bounds[0..2] = ordered keys of the per-axis minima, bounds[3..5] of the maxima; NaN and
// infinite coordinates are left out.
__global__ void bounds_init_kernel(uint32_t* bounds)
{
    if (threadIdx.x < 6) bounds[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
}

__global__ void __launch_bounds__(kBlock) bounds_kernel(const float* __restrict__ means, uint64_t n, uint32_t* bounds)
{
    uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        for (int a = 0; a < 3; ++a) {
            const float x = means[3 * i + a];
            if (__builtin_isfinite(x)) {
                const uint32_t k = ordered_key(x);
                lo[a] = min(lo[a], k);
                hi[a] = max(hi[a], k);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], (uint32_t)__shfl_xor((int)lo[a], off, 64));
            hi[a] = max(hi[a], (uint32_t)__shfl_xor((int)hi[a], off, 64));
        }
    }
    if ((threadIdx.x & 63u) == 0) {
        for (int a = 0; a < 3; ++a) {
            atomicMin(bounds + a, lo[a]);
            atomicMax(bounds + 3 + a, hi[a]);
        }
    }
}

// Morton code with 21 bits per axis of the bounding cube: bit 3k of the code is bit k of the
// x cell, 3k+1 of y, 3k+2 of z.
__device__ __forceinline__ uint64_t spread3(uint32_t v)
{
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ void __launch_bounds__(kBlock) morton_codes_kernel(const float* __restrict__ means, uint64_t n,
                                                              const uint32_t* __restrict__ bounds,
                                                              uint64_t* __restrict__ codes, uint32_t* __restrict__ index)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float lo[3];
    bool has[3];  // the axis has at least one finite coordinate
    float extent = 0.0f;
    for (int a = 0; a < 3; ++a) {
        has[a] = bounds[a] != 0xffffffffu;
        lo[a] = has[a] ? key_to_float(bounds[a]) : 0.0f;
        const float span = has[a] ? key_to_float(bounds[3 + a]) - lo[a] : 0.0f;
        extent = span > extent ? span : extent;
    }
    uint64_t code = 0;
    if (extent > 0.0f) {
        for (int a = 0; a < 3; ++a) {
            const float t = (means[3 * i + a] - lo[a]) / extent;
            // NaN and t <= 0 give cell 0; the top face and +inf give the last cell
            const uint32_t q = has[a] && t > 0.0f ? (uint32_t)fminf(t * 2097152.0f, 2097151.0f) : 0u;
            code |= spread3(q) << a;
        }
    }
    codes[i] = code;
    index[i] = (uint32_t)i;
}

uint32_t rows_grid(uint64_t rows, uint32_t fields)
{
    const uint64_t per_wave = 64u / fields;
    const uint64_t waves = (rows + per_wave - 1) / per_wave;
    return (uint32_t)((waves + kBlock / 64 - 1) / (kBlock / 64));
}

}  // namespace

hipError_t launch_ply_decode(const void* records, uint64_t n_records, const PlyDecodeTable& table, uint64_t first,
                             const SceneOut& out, hipStream_t stream)
{
    if (n_records == 0) return hipSuccess;
    hipLaunchKernelGGL(ply_decode_kernel, dim3(rows_grid(n_records, 11u + 3u * out.sh_bases)), dim3(kBlock), 0, stream,
                       (const uint8_t*)records, n_records, table, first, out);
    return hipGetLastError();
}

hipError_t launch_gather_rows(const SceneArgs& src, const uint32_t* order, const SceneOut& dst, hipStream_t stream)
{
    if (dst.n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(rows_grid(dst.n, 11u + 3u * dst.sh_bases)), dim3(kBlock), 0, stream,
                       src, order, dst);
    return hipGetLastError();
}

size_t morton_sort_temp_bytes(uint64_t n)
{
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                           (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 63) != hipSuccess)
        return 0;
    return bytes;
}

hipError_t launch_morton_order(const float* means, uint64_t n, uint32_t* order, uint64_t* sorted_codes,
                               const MortonBuffers& ws, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(bounds_init_kernel, dim3(1), dim3(64), 0, stream, ws.bounds);
    const uint64_t blocks = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(bounds_kernel, dim3((uint32_t)std::min<uint64_t>(blocks, 4096)), dim3(kBlock), 0, stream, means, n,
                       ws.bounds);
    hipLaunchKernelGGL(morton_codes_kernel, dim3((uint32_t)blocks), dim3(kBlock), 0, stream, means, n, ws.bounds,
                       ws.codes_in, ws.index_in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t temp = ws.temp_bytes;
    // LSD radix sort: stable, so equal codes keep their file order
    return hipcub::DeviceRadixSort::SortPairs(ws.temp, temp, ws.codes_in, sorted_codes, ws.index_in, order, (int)n, 0,
                                              63, stream);
}

}  // namespace sgs
