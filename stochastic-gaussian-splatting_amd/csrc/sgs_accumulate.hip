// Accumulation kernels: closest accepted fragment per framebuffer (sub)pixel with a
// 64-bit atomicMin on (float_bits(depth) << 32 | gaussian_index).
//
// Scatter schedule for wave64: one lane per Gaussian projects it; Gaussians with at most
// `seq` fragments (points) are finished by their own lane; every larger one is then taken
// by the whole wave in ballot order, its parameters broadcast through SGPRs with
// v_readlane, the lanes striding 64 fragments at a time in row-major AABB order with an
// incremental (x, y) walk (no per-fragment division). atomicMin is order-independent, so
// the framebuffer does not depend on this schedule.
//
// Restates accumulate_scatter.cu:19-146 and scatter_utils.cuh:14-183; the binning variants
// restate accumulate_hybrid.cu:34-117 and accumulate_hybrid_gps.cu:27-123.
#include "sgs_internal.h"

namespace sgs {

constexpr uint32_t kBlock = 256;

__global__ void fill_u64_kernel(uint64_t* __restrict__ p, size_t n, uint64_t v)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x * 2u;
    for (size_t k = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 2u; k < n; k += stride) {
        if (k + 1u < n) {
            ulonglong2 pair;
            pair.x = v;
            pair.y = v;
            *reinterpret_cast<ulonglong2*>(p + k) = pair;
        } else {
            p[k] = v;
        }
    }
}

__global__ void zero_u32_kernel(uint32_t* __restrict__ p, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) p[k] = 0u;
}

static unsigned grid_for(size_t items, unsigned per_block, unsigned cap)
{
    size_t g = (items + per_block - 1u) / per_block;
    if (g < 1u) g = 1u;
    return (unsigned)(g > cap ? cap : g);
}

hipError_t launch_fill_u64(uint64_t* p, size_t n, uint64_t value, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(n, 2u * kBlock, 8192u)), dim3(kBlock), 0, stream, p, n, value);
    return hipGetLastError();
}

__global__ void fill_u64_if_kernel(uint64_t* __restrict__ p, size_t n, uint64_t v, const uint32_t* __restrict__ flag)
{
    if (!*flag) return;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) p[k] = v;
}

hipError_t launch_fill_u64_if(uint64_t* p, size_t n, uint64_t value, const uint32_t* run_if_set, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u64_if_kernel, dim3(grid_for(n, kBlock, 8192u)), dim3(kBlock), 0, stream, p, n, value,
                       run_if_set);
    return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, size_t n, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_for(n, kBlock, 8192u)), dim3(kBlock), 0, stream, p, n);
    return hipGetLastError();
}

// sum over the wave, then one 64-bit atomic per wave
__device__ __forceinline__ void wave_count(uint64_t* counter, uint32_t v)
{
    uint32_t s = v;
    for (int off = 32; off > 0; off >>= 1) s += (uint32_t)__shfl_xor((int)s, off, 64);
    if (lane_id() == 0u && s) atomicAdd(reinterpret_cast<unsigned long long*>(counter), (unsigned long long)s);
}

// ------------------------------------------------------------------------------------------
// stochastic transparency
// ------------------------------------------------------------------------------------------
template <bool kBlue>
__device__ __forceinline__ bool st_accept(uint32_t px, uint32_t py, float msx, float msy, const ConicExp2 ce, float o,
                                          uint32_t g, uint32_t frame, const uint8_t* __restrict__ bn, uint32_t bn_rot)
{
    const float dx = (float)px - msx;
    const float dy = (float)py - msy;
    const float alpha = fragment_alpha(ce, o, dx, dy);
    if (alpha < kMinAlpha) return false;
    const float u = kBlue ? blue_noise(bn, bn_rot, g, px, py) : rand_uniform(frame, g, px, py);
    return !(u >= alpha);
}

template <bool kBlue, bool kBin, bool kCount>
__global__ __launch_bounds__(kBlock) void st_scatter_kernel(SceneArgs sc, const float4* __restrict__ w2c, FbCamera cam,
                                                            uint64_t* __restrict__ fb, uint32_t frame,
                                                            const uint8_t* __restrict__ bn, uint32_t bn_rot,
                                                            uint32_t seq, BinArgs bin, uint64_t* __restrict__ counters)
{
    const uint32_t lane = lane_id();
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    const float4 r1 = w2c[0], r2 = w2c[1], r3 = w2c[2];
    Projected p;
    bool active = g < sc.n && project_scene(sc, g, r1, r2, r3, cam, p);

    uint32_t x0 = 0, y0 = 0, w = 0, count = 0;
    float msx = 0.0f, msy = 0.0f, o = 0.0f;
    ConicExp2 ce = { 0.0f, 0.0f, 0.0f };
    uint64_t frag = 0;
    if (active) {
        const Aabb b = pixel_aabb(p, cam);
        x0 = b.x0;
        y0 = b.y0;
        w = b.x1 - b.x0;
        count = w * (b.y1 - b.y0);
        msx = p.mean_x - 0.5f;
        msy = p.mean_y - 0.5f;
        ce = conic_exp2(p.conic_x, p.conic_y, p.conic_z);
        o = p.opacity;
        frag = fragment_key(p.depth, g);
    }
    uint32_t n_tested = 0, n_accepted = 0, n_binned = 0;
    const uint32_t n_projected = active ? 1u : 0u;

    if constexpr (kBin) {
        // wave-aggregated slot reservation; beyond the capacity Gaussians are scattered
        const bool want = active && count > bin.threshold;
        const uint64_t m = __ballot(want);
        if (m) {
            const int leader = __builtin_ctzll(m);
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&bin.counters->n_binned, (uint32_t)__popcll(m));
            base = bcast(base, leader);
            const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (want && slot < bin.capacity) {
                float4* dst = reinterpret_cast<float4*>(static_cast<BinnedST*>(bin.binned) + slot);
                const Aabb b = pixel_aabb(p, cam);
                dst[0] = make_float4(__uint_as_float((uint32_t)frag), __uint_as_float((uint32_t)(frag >> 32)), msx, msy);
                dst[1] = make_float4(p.conic_x, p.conic_y, p.conic_z, o);
                dst[2] = make_float4(p.power_threshold, __uint_as_float(b.x0 | (b.x1 << 16)),
                                     __uint_as_float(b.y0 | (b.y1 << 16)), 0.0f);
                active = false;
                n_binned = 1;
            }
        }
    }

    // small Gaussians: the owning lane walks its AABB
    if (active && count <= seq) {
        uint32_t r = 0, q = 0;
        for (uint32_t f = 0; f < count; f++) {
            const uint32_t px = x0 + r, py = y0 + q;
            if (st_accept<kBlue>(px, py, msx, msy, ce, o, g, frame, bn, bn_rot)) {
                atomicMin(reinterpret_cast<unsigned long long*>(&fb[py * cam.width_i + px]), (unsigned long long)frag);
                if constexpr (kCount) n_accepted++;
            }
            if constexpr (kCount) n_tested++;
            if (++r == w) {
                r = 0;
                q++;
            }
        }
    }

    // larger Gaussians: the whole wave, one Gaussian at a time
    uint64_t mask = __ballot(active && count > seq);
    while (mask) {
        const int l = __builtin_ctzll(mask);
        mask &= mask - 1ull;
        const uint32_t cx0 = bcast(x0, l), cy0 = bcast(y0, l), cw = bcast(w, l), cn = bcast(count, l);
        const uint32_t cg = bcast(g, l);
        const float cmx = bcast(msx, l), cmy = bcast(msy, l), co = bcast(o, l);
        const ConicExp2 cce = { bcast(ce.a, l), bcast(ce.b, l), bcast(ce.c, l) };
        const uint64_t cfrag = bcast(frag, l);
        uint32_t q, r, qs, rs;
        if (cw >= 64u) {
            q = 0;
            r = lane;
            qs = 0;
            rs = 64u;
        } else {
            q = lane / cw;
            r = lane - q * cw;
            qs = 64u / cw;
            rs = 64u - qs * cw;
        }
        for (uint32_t f = lane; f < cn; f += 64u) {
            const uint32_t px = cx0 + r, py = cy0 + q;
            if (st_accept<kBlue>(px, py, cmx, cmy, cce, co, cg, frame, bn, bn_rot)) {
                atomicMin(reinterpret_cast<unsigned long long*>(&fb[py * cam.width_i + px]), (unsigned long long)cfrag);
                if constexpr (kCount) n_accepted++;
            }
            if constexpr (kCount) n_tested++;
            r += rs;
            q += qs;
            if (r >= cw) {
                r -= cw;
                q++;
            }
        }
    }

    if constexpr (kCount) {
        wave_count(&counters[0], n_tested);
        wave_count(&counters[1], n_accepted);
        wave_count(&counters[2], n_projected);
        wave_count(&counters[3], n_binned);
    }
}

template <bool kBlue, bool kBin, bool kCount>
static void st_dispatch(unsigned grid, hipStream_t stream, const SceneArgs& sc, const float4* w2c, const FbCamera& cam,
                        uint64_t* fb, uint32_t frame, const uint8_t* bn, uint32_t rot, uint32_t seq, const BinArgs& bin,
                        uint64_t* counters)
{
    hipLaunchKernelGGL((st_scatter_kernel<kBlue, kBin, kCount>), dim3(grid), dim3(kBlock), 0, stream, sc, w2c, cam, fb,
                       frame, bn, rot, seq, bin, counters);
}

hipError_t launch_st_scatter(const SceneArgs& sc, const float* w2c, const FbCamera& cam, uint64_t* fb, uint32_t frame,
                             const uint8_t* bn, uint32_t rot, uint32_t seq, const BinArgs& bin, uint64_t* counters,
                             hipStream_t stream)
{
    if (sc.n == 0) return hipSuccess;
    const unsigned grid = (sc.n + kBlock - 1u) / kBlock;
    const float4* w = reinterpret_cast<const float4*>(w2c);
    const bool blue = bn != nullptr, binning = bin.binned != nullptr, count = counters != nullptr;
    // 8 instantiations: blue noise x binning x counting
    if (blue) {
        if (binning) {
            if (count) st_dispatch<true, true, true>(grid, stream, sc, w, cam, fb, frame, bn, rot, seq, bin, counters);
            else st_dispatch<true, true, false>(grid, stream, sc, w, cam, fb, frame, bn, rot, seq, bin, counters);
        } else {
            if (count) st_dispatch<true, false, true>(grid, stream, sc, w, cam, fb, frame, bn, rot, seq, bin, counters);
            else st_dispatch<true, false, false>(grid, stream, sc, w, cam, fb, frame, bn, rot, seq, bin, counters);
        }
    } else {
        if (binning) {
            if (count) st_dispatch<false, true, true>(grid, stream, sc, w, cam, fb, frame, bn, rot, seq, bin, counters);
            else st_dispatch<false, true, false>(grid, stream, sc, w, cam, fb, frame, bn, rot, seq, bin, counters);
        } else {
            if (count) st_dispatch<false, false, true>(grid, stream, sc, w, cam, fb, frame, bn, rot, seq, bin, counters);
            else st_dispatch<false, false, false>(grid, stream, sc, w, cam, fb, frame, bn, rot, seq, bin, counters);
        }
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Gaussian point splatting
// ------------------------------------------------------------------------------------------
template <bool kBin, bool kCount>
__global__ __launch_bounds__(kBlock) void gps_scatter_kernel(SceneArgs sc, const float4* __restrict__ w2c, FbCamera cam,
                                                             uint64_t* __restrict__ fb, uint32_t frame,
                                                             uint32_t max_points, uint32_t seq, BinArgs bin,
                                                             uint64_t* __restrict__ counters)
{
    const uint32_t lane = lane_id();
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    const float4 r1 = w2c[0], r2 = w2c[1], r3 = w2c[2];
    Projected p;
    bool active = g < sc.n && project_scene(sc, g, r1, r2, r3, cam, p);
    const uint32_t n_projected = active ? 1u : 0u;

    uint32_t count = 0, seed_x = 0, seed_y = 0;
    float mx = 0.0f, my = 0.0f, L0 = 0.0f, L1 = 0.0f, L2 = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f, pt = 0.0f;
    float o_rcp = 0.0f, o_dilog = 0.0f;
    uint64_t frag = 0;
    if (active) {
        // importance and Poisson point count (accumulate_scatter.cu:120-129)
        o_dilog = dilog(p.opacity);
        const float importance = kTwoPi * sqrtf(p.determinant) * o_dilog;
        uint2 seed = make_seed(g, frame);
        seed_x = seed.x;
        seed_y = seed.y;
        const uint32_t ct = poisson(seed, importance);
        count = ct < max_points ? ct : max_points;
        if (count == 0u) active = false;
        L0 = sqrtf(p.cov_x);
        L1 = p.cov_y / L0;
        L2 = sqrtf(p.cov_z - L1 * L1);
        o_rcp = 1.0f / fmaxf(p.opacity, 1.0e-37f);
        mx = p.mean_x;
        my = p.mean_y;
        cx = p.conic_x;
        cy = p.conic_y;
        cz = p.conic_z;
        pt = p.power_threshold;
        frag = fragment_key(p.depth, g);
    }
    uint32_t n_tested = 0, n_accepted = 0, n_binned = 0;

    if constexpr (kBin) {
        const bool want = active && count > bin.threshold;
        const uint64_t m = __ballot(want);
        if (m) {
            const int leader = __builtin_ctzll(m);
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&bin.counters->n_binned, (uint32_t)__popcll(m));
            base = bcast(base, leader);
            const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (want && slot < bin.capacity) {
                float4* dst = reinterpret_cast<float4*>(static_cast<BinnedGPS*>(bin.binned) + slot);
                dst[0] = make_float4(__uint_as_float((uint32_t)frag), __uint_as_float((uint32_t)(frag >> 32)), mx, my);
                dst[1] = make_float4(L0, L1, L2, pt);
                dst[2] = make_float4(cx, cy, cz, o_rcp);
                dst[3] = make_float4(o_dilog, __uint_as_float(count), __uint_as_float(seed_x), __uint_as_float(seed_y));
                active = false;
                n_binned = 1;
            }
        }
    }

    if (active && count <= seq) {
        for (uint32_t k = 0; k < count; k++) {
            uint32_t pix;
            if (gps_point_at(k, seed_x, seed_y, mx, my, L0, L1, L2, cx, cy, cz, pt, o_rcp, o_dilog, cam, pix)) {
                atomicMin(reinterpret_cast<unsigned long long*>(&fb[pix]), (unsigned long long)frag);
                if constexpr (kCount) n_accepted++;
            }
            if constexpr (kCount) n_tested++;
        }
    }

    uint64_t mask = __ballot(active && count > seq);
    while (mask) {
        const int l = __builtin_ctzll(mask);
        mask &= mask - 1ull;
        const uint32_t cn = bcast(count, l), csx = bcast(seed_x, l), csy = bcast(seed_y, l);
        const float cmx = bcast(mx, l), cmy = bcast(my, l);
        const float cL0 = bcast(L0, l), cL1 = bcast(L1, l), cL2 = bcast(L2, l);
        const float ccx = bcast(cx, l), ccy = bcast(cy, l), ccz = bcast(cz, l), cpt = bcast(pt, l);
        const float corc = bcast(o_rcp, l), codl = bcast(o_dilog, l);
        const uint64_t cfrag = bcast(frag, l);
        for (uint32_t k = lane; k < cn; k += 64u) {
            uint32_t pix;
            if (gps_point_at(k, csx, csy, cmx, cmy, cL0, cL1, cL2, ccx, ccy, ccz, cpt, corc, codl, cam, pix)) {
                atomicMin(reinterpret_cast<unsigned long long*>(&fb[pix]), (unsigned long long)cfrag);
                if constexpr (kCount) n_accepted++;
            }
            if constexpr (kCount) n_tested++;
        }
    }

    if constexpr (kCount) {
        wave_count(&counters[0], n_tested);
        wave_count(&counters[1], n_accepted);
        wave_count(&counters[2], n_projected);
        wave_count(&counters[3], n_binned);
    }
}

hipError_t launch_gps_scatter(const SceneArgs& sc, const float* w2c, const FbCamera& cam, uint64_t* fb, uint32_t frame,
                              uint32_t max_points, uint32_t seq, const BinArgs& bin, uint64_t* counters,
                              hipStream_t stream)
{
    if (sc.n == 0) return hipSuccess;
    const unsigned grid = (sc.n + kBlock - 1u) / kBlock;
    const float4* w = reinterpret_cast<const float4*>(w2c);
    const bool binning = bin.binned != nullptr, count = counters != nullptr;
    if (binning) {
        if (count)
            hipLaunchKernelGGL((gps_scatter_kernel<true, true>), dim3(grid), dim3(kBlock), 0, stream, sc, w, cam, fb,
                               frame, max_points, seq, bin, counters);
        else
            hipLaunchKernelGGL((gps_scatter_kernel<true, false>), dim3(grid), dim3(kBlock), 0, stream, sc, w, cam, fb,
                               frame, max_points, seq, bin, counters);
    } else {
        if (count)
            hipLaunchKernelGGL((gps_scatter_kernel<false, true>), dim3(grid), dim3(kBlock), 0, stream, sc, w, cam, fb,
                               frame, max_points, seq, bin, counters);
        else
            hipLaunchKernelGGL((gps_scatter_kernel<false, false>), dim3(grid), dim3(kBlock), 0, stream, sc, w, cam, fb,
                               frame, max_points, seq, bin, counters);
    }
    return hipGetLastError();
}

}  // namespace sgs
