// Hybrid scheduling of large Gaussians, with no host synchronisation.
//
// Stochastic transparency (restates accumulate_hybrid.cu:34-321 with a different schedule):
//   1. the scatter kernel appends Gaussians whose AABB exceeds binning_pixel_threshold to a
//      binned list (wave-aggregated atomicAdd) and scatters the rest;
//   2. tile_count: for every binned Gaussian and every tile row, the x-span of its
//      power-threshold ellipse inside the row's pixel strip (the per-row spans of AccuTile,
//      arXiv:2412.00578) gives the covered tiles; one atomicAdd per (tile, Gaussian);
//   3. a single-workgroup exclusive scan turns tile counts into offsets;
//   4. tile_fill writes the binned index of every instance at offset + atomic cursor
//      (a counting sort by tile; the order inside a tile is irrelevant because min is);
//   5. tile_raster: one workgroup per TS x TS tile stages the tile's framebuffer words in
//      LDS, enumerates each instance's AABB-clipped rectangle across all 256 lanes, rejects
//      fragments behind the current LDS value before evaluating them, resolves accepted ones
//      with ds_min_u64 and writes the changed words back once.
//   If the instance list overflows its capacity, steps 4-5 do nothing and every binned
//   Gaussian is scattered by the wave-cooperative fallback instead.
// The tested pixels, alpha and thresholds are those of the scatter path, so the framebuffer
// is bit-identical to it.
//
// Gaussian point splatting (restates accumulate_hybrid_gps.cu:27-181): Gaussians with more
// than binning_point_threshold points are binned; a single-workgroup scan of their chunk
// counts (256 points per chunk) feeds a persistent kernel that processes one chunk per
// workgroup iteration, one point per lane. Every point uses the per-point counter seed of
// the scatter path, so the result is bit-identical to it.
#include "sgs_internal.h"
#include "sgs_scan.h"

namespace sgs {

constexpr uint32_t kRasterThreads = 256;
constexpr uint32_t kChunk = 256;
constexpr uint32_t kPersistentBlocks = 2048;

// ------------------------------------------------------------------------------------------
// scratch layout
// ------------------------------------------------------------------------------------------
struct StLayout {
    size_t counters, binned, tile_count, tile_cursor, tile_offset, instances, total;
    uint32_t grid_w, grid_h, n_tiles;
};

static StLayout st_layout(uint32_t fb_w, uint32_t fb_h, const Tuning& t)
{
    StLayout L;
    const uint32_t ts = t.tile_size;
    L.grid_w = (fb_w + ts - 1u) / ts;
    L.grid_h = (fb_h + ts - 1u) / ts;
    L.n_tiles = L.grid_w * L.grid_h;
    Arena a;
    L.counters = a.take(sizeof(HybridCounters));
    // tile_count and tile_cursor follow the counters so one zero-fill clears all three
    L.tile_count = a.offset;
    a.offset += 4u * (size_t)L.n_tiles;
    L.tile_cursor = a.offset;
    a.offset += 4u * (size_t)L.n_tiles;
    a.offset = (a.offset + 255u) & ~size_t(255u);
    L.tile_offset = a.take(4u * ((size_t)L.n_tiles + 1u));
    L.binned = a.take(sizeof(BinnedST) * (size_t)t.max_binned);
    L.instances = a.take(4u * (size_t)t.max_tile_instances);
    L.total = a.offset;
    return L;
}

struct GpsLayout {
    size_t counters, binned, chunk_offsets, total;
};

static GpsLayout gps_layout(const Tuning& t)
{
    GpsLayout L;
    Arena a;
    L.counters = a.take(sizeof(HybridCounters));
    L.binned = a.take(sizeof(BinnedGPS) * (size_t)t.max_binned);
    L.chunk_offsets = a.take(4u * ((size_t)t.max_binned + 1u));
    L.total = a.offset;
    return L;
}

size_t hybrid_st_scratch_bytes(uint32_t fb_w, uint32_t fb_h, const Tuning& t) { return st_layout(fb_w, fb_h, t).total; }
size_t hybrid_gps_scratch_bytes(const Tuning& t) { return gps_layout(t).total; }

struct ChunkInput {
    const BinnedGPS* binned;
    __device__ uint32_t operator()(uint32_t i) const { return (binned[i].n_points + kChunk - 1u) / kChunk; }
};

// ------------------------------------------------------------------------------------------
// tile spans of a binned ST Gaussian
// ------------------------------------------------------------------------------------------
struct BinnedView {
    uint64_t fragment;
    float msx, msy, cx, cy, cz, opacity, pt;
    uint32_t x0, x1, y0, y1;
};

__device__ __forceinline__ BinnedView load_binned(const BinnedST* __restrict__ b, uint32_t i)
{
    const float4* p = reinterpret_cast<const float4*>(b + i);
    const float4 a = p[0], c = p[1], d = p[2];
    BinnedView v;
    v.fragment = ((uint64_t)__float_as_uint(a.y) << 32) | __float_as_uint(a.x);
    v.msx = a.z;
    v.msy = a.w;
    v.cx = c.x;
    v.cy = c.y;
    v.cz = c.z;
    v.opacity = c.w;
    v.pt = d.x;
    const uint32_t xx = __float_as_uint(d.y), yy = __float_as_uint(d.z);
    v.x0 = xx & 0xffffu;
    v.x1 = xx >> 16;
    v.y0 = yy & 0xffffu;
    v.y1 = yy >> 16;
    return v;
}

// Pixel columns [c0, c1] of tile row pixel rows [ya, yb] that may hold an accepted fragment:
// the x-extent (FP64, widened by a margin) of {power <= pt} over the strip, clipped to the
// AABB. Returns false if the strip misses the ellipse.
__device__ __forceinline__ bool strip_columns(const BinnedView& v, uint32_t ya, uint32_t yb, uint32_t& c0, uint32_t& c1)
{
    const double A = v.cx, B = v.cy, C = v.cz;
    const double two_pt = 2.0 * ((double)v.pt + 0.01 + 1e-3 * fabs((double)v.pt));
    const double det = A * C - B * B;
    if (!(det > 0.0) || !(A > 0.0) || !(C > 0.0) || !(two_pt > 0.0)) {
        // degenerate: keep the whole AABB row span
        c0 = v.x0;
        c1 = v.x1 - 1u;
        return true;
    }
    const double dy_max = sqrt(two_pt * A / det);
    double d0 = (double)ya - (double)v.msy, d1 = (double)yb - (double)v.msy;
    if (d1 < -dy_max || d0 > dy_max) return false;
    d0 = fmax(d0, -dy_max);
    d1 = fmin(d1, dy_max);
    auto half = [&](double dy) { return sqrt(fmax(two_pt * A - det * dy * dy, 0.0)) / A; };
    auto centre = [&](double dy) { return -B * dy / A; };
    double lo = fmin(centre(d0) - half(d0), centre(d1) - half(d1));
    double hi = fmax(centre(d0) + half(d0), centre(d1) + half(d1));
    const double x_ext = sqrt(two_pt * C / det);
    const double dy_r = -B * x_ext / C, dy_l = B * x_ext / C;  // rightmost / leftmost points
    if (dy_r >= d0 && dy_r <= d1) hi = x_ext;
    if (dy_l >= d0 && dy_l <= d1) lo = -x_ext;
    const double plo = ceil((double)v.msx + lo - 1e-3), phi = floor((double)v.msx + hi + 1e-3);
    const double a0 = fmax(plo, (double)v.x0), a1 = fmin(phi, (double)v.x1 - 1.0);
    if (a1 < a0) return false;
    c0 = (uint32_t)a0;
    c1 = (uint32_t)a1;
    return true;
}

// visits every (tile, binned Gaussian) instance; lane per binned Gaussian, persistent grid
template <bool kFill>
__global__ __launch_bounds__(256) void tile_instances_kernel(const BinnedST* __restrict__ binned,
                                                             const HybridCounters* __restrict__ counters,
                                                             uint32_t capacity, uint32_t ts, uint32_t grid_w,
                                                             uint32_t* __restrict__ tile_count,
                                                             uint32_t* __restrict__ tile_cursor,
                                                             const uint32_t* __restrict__ tile_offset,
                                                             uint32_t* __restrict__ instances)
{
    if constexpr (kFill) {
        if (counters->overflow) return;
    }
    uint32_t n = counters->n_binned;
    if (n > capacity) n = capacity;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const BinnedView v = load_binned(binned, i);
        const uint32_t ty0 = v.y0 / ts, ty1 = (v.y1 - 1u) / ts;
        for (uint32_t ty = ty0; ty <= ty1; ty++) {
            const uint32_t ya = max(ty * ts, v.y0), yb = min(ty * ts + ts - 1u, v.y1 - 1u);
            uint32_t c0, c1;
            if (!strip_columns(v, ya, yb, c0, c1)) continue;
            for (uint32_t tx = c0 / ts; tx <= c1 / ts; tx++) {
                const uint32_t tile = ty * grid_w + tx;
                if constexpr (kFill) {
                    const uint32_t slot = atomicAdd(&tile_cursor[tile], 1u);
                    instances[tile_offset[tile] + slot] = i;
                } else {
                    atomicAdd(&tile_count[tile], 1u);
                }
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// per-tile rasterisation with the framebuffer tile in LDS
// ------------------------------------------------------------------------------------------
template <uint32_t TS, bool kBlue>
__global__ __launch_bounds__(kRasterThreads) void tile_raster_kernel(const BinnedST* __restrict__ binned,
                                                                     const HybridCounters* __restrict__ counters,
                                                                     const uint32_t* __restrict__ tile_offset,
                                                                     const uint32_t* __restrict__ instances,
                                                                     uint64_t* __restrict__ fb, uint32_t fb_w,
                                                                     uint32_t fb_h, uint32_t grid_w, uint32_t frame,
                                                                     const uint8_t* __restrict__ bn, uint32_t bn_rot)
{
    constexpr uint32_t kPix = TS * TS;
    constexpr uint32_t kPerThread = kPix / kRasterThreads;
    __shared__ unsigned long long tile_fb[kPix];
    __shared__ uint64_t s_frag[kRasterThreads];
    __shared__ float s_msx[kRasterThreads], s_msy[kRasterThreads], s_o[kRasterThreads];
    __shared__ float s_ca[kRasterThreads], s_cb[kRasterThreads], s_cc[kRasterThreads];
    __shared__ uint32_t s_rect[kRasterThreads];   // rx0 | ry0 << 8 | rw << 16 | rh << 24 (tile-local)
    __shared__ uint32_t s_magic[kRasterThreads];  // ceil(2^16 / rw)

    if (counters->overflow) return;
    const uint32_t tile = blockIdx.x;
    const uint32_t begin = tile_offset[tile], end = tile_offset[tile + 1u];
    if (begin == end) return;
    const uint32_t tx0 = (tile % grid_w) * TS, ty0 = (tile / grid_w) * TS;
    const uint32_t tid = threadIdx.x;

    uint64_t orig[kPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kPerThread; k++) {
        const uint32_t l = tid + k * kRasterThreads;
        const uint32_t px = tx0 + l % TS, py = ty0 + l / TS;
        // pixels outside the image start at 0, which no fragment beats
        orig[k] = (px < fb_w && py < fb_h) ? fb[(size_t)py * fb_w + px] : 0ull;
        tile_fb[l] = orig[k];
    }

    for (uint32_t batch = begin; batch < end; batch += kRasterThreads) {
        __syncthreads();
        const uint32_t nb = min(kRasterThreads, end - batch);
        if (tid < nb) {
            const BinnedView v = load_binned(binned, instances[batch + tid]);
            const ConicExp2 ce = conic_exp2(v.cx, v.cy, v.cz);
            const bool overlap = v.x1 > tx0 && v.x0 < tx0 + TS && v.y1 > ty0 && v.y0 < ty0 + TS;
            const uint32_t rx0 = overlap ? max(v.x0, tx0) - tx0 : 0u, rx1 = overlap ? min(v.x1, tx0 + TS) - tx0 : 0u;
            const uint32_t ry0 = overlap ? max(v.y0, ty0) - ty0 : 0u, ry1 = overlap ? min(v.y1, ty0 + TS) - ty0 : 0u;
            const uint32_t rw = rx1 - rx0, rh = ry1 - ry0;
            s_frag[tid] = v.fragment;
            s_msx[tid] = v.msx;
            s_msy[tid] = v.msy;
            s_o[tid] = v.opacity;
            s_ca[tid] = ce.a;
            s_cb[tid] = ce.b;
            s_cc[tid] = ce.c;
            s_rect[tid] = rx0 | (ry0 << 8) | (rw << 16) | (rh << 24);
            s_magic[tid] = rw ? (65536u + rw - 1u) / rw : 0u;
        }
        __syncthreads();
        for (uint32_t j = 0; j < nb; j++) {
            const uint32_t rect = s_rect[j];
            const uint32_t rx0 = rect & 0xffu, ry0 = (rect >> 8) & 0xffu, rw = (rect >> 16) & 0xffu,
                           rh = rect >> 24;
            const uint32_t cnt = rw * rh;
            if (tid >= cnt) continue;
            const uint64_t frag = s_frag[j];
            const uint32_t g = (uint32_t)frag;
            const float msx = s_msx[j], msy = s_msy[j], o = s_o[j];
            const ConicExp2 ce = { s_ca[j], s_cb[j], s_cc[j] };
            // row-major walk of the clipped rectangle, 256 fragments per step
            const uint32_t magic = s_magic[j];
            uint32_t q = (tid * magic) >> 16;
            uint32_t r = tid - q * rw;
            const uint32_t qs = (kRasterThreads * magic) >> 16;
            uint32_t rs = kRasterThreads - qs * rw;
            uint32_t qs2 = qs;
            if (rs >= rw) {  // guard the magic quotient of 256 / rw
                rs -= rw;
                qs2++;
            }
            for (uint32_t f = tid; f < cnt; f += kRasterThreads) {
                const uint32_t lx = rx0 + r, ly = ry0 + q;
                const uint32_t l = ly * TS + lx;
                // the LDS word only decreases, so a stale read never rejects a winner
                if (frag < __hip_atomic_load(&tile_fb[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
                    const uint32_t px = tx0 + lx, py = ty0 + ly;
                    const float dx = (float)px - msx, dy = (float)py - msy;
                    const float alpha = fragment_alpha(ce, o, dx, dy);
                    if (!(alpha < kMinAlpha)) {
                        const float u = kBlue ? blue_noise(bn, bn_rot, g, px, py) : rand_uniform(frame, g, px, py);
                        if (!(u >= alpha)) atomicMin(&tile_fb[l], (unsigned long long)frag);
                    }
                }
                r += rs;
                q += qs2;
                if (r >= rw) {
                    r -= rw;
                    q++;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kPerThread; k++) {
        const uint32_t l = tid + k * kRasterThreads;
        const uint32_t px = tx0 + l % TS, py = ty0 + l / TS;
        const uint64_t v = tile_fb[l];
        if (v != orig[k] && px < fb_w && py < fb_h) fb[(size_t)py * fb_w + px] = v;
    }
}

// wave per binned Gaussian: scatter its whole AABB (instance-list overflow)
template <bool kBlue>
__global__ __launch_bounds__(256) void binned_fallback_kernel(const BinnedST* __restrict__ binned,
                                                              const HybridCounters* __restrict__ counters,
                                                              uint32_t capacity, uint64_t* __restrict__ fb,
                                                              uint32_t fb_w, uint32_t frame,
                                                              const uint8_t* __restrict__ bn, uint32_t bn_rot)
{
    if (!counters->overflow) return;
    uint32_t n = counters->n_binned;
    if (n > capacity) n = capacity;
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x / 64u);
    for (uint32_t i = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u; i < n; i += waves) {
        const BinnedView v = load_binned(binned, i);
        const ConicExp2 ce = conic_exp2(v.cx, v.cy, v.cz);
        const uint32_t w = v.x1 - v.x0, cnt = w * (v.y1 - v.y0);
        const uint32_t g = (uint32_t)v.fragment;
        for (uint32_t f = lane; f < cnt; f += 64u) {
            const uint32_t px = v.x0 + f % w, py = v.y0 + f / w;
            const float dx = (float)px - v.msx, dy = (float)py - v.msy;
            const float alpha = fragment_alpha(ce, v.opacity, dx, dy);
            if (alpha < kMinAlpha) continue;
            const float u = kBlue ? blue_noise(bn, bn_rot, g, px, py) : rand_uniform(frame, g, px, py);
            if (u >= alpha) continue;
            atomicMin(reinterpret_cast<unsigned long long*>(&fb[(size_t)py * fb_w + px]), (unsigned long long)v.fragment);
        }
    }
}

template <uint32_t TS>
static void launch_raster(bool blue, uint32_t n_tiles, hipStream_t stream, const BinnedST* binned,
                          const HybridCounters* counters, const uint32_t* tile_offset, const uint32_t* instances,
                          uint64_t* fb, uint32_t fb_w, uint32_t fb_h, uint32_t grid_w, uint32_t frame,
                          const uint8_t* bn, uint32_t rot)
{
    if (blue)
        hipLaunchKernelGGL((tile_raster_kernel<TS, true>), dim3(n_tiles), dim3(kRasterThreads), 0, stream, binned,
                           counters, tile_offset, instances, fb, fb_w, fb_h, grid_w, frame, bn, rot);
    else
        hipLaunchKernelGGL((tile_raster_kernel<TS, false>), dim3(n_tiles), dim3(kRasterThreads), 0, stream, binned,
                           counters, tile_offset, instances, fb, fb_w, fb_h, grid_w, frame, bn, rot);
}

static unsigned persistent_grid(uint32_t capacity, uint32_t per_block)
{
    const uint32_t g = (capacity + per_block - 1u) / per_block;
    return g < 1u ? 1u : (g > kPersistentBlocks ? kPersistentBlocks : g);
}

hipError_t run_hybrid_st(const SceneArgs& sc, const float* w2c, const FbCamera& cam, uint64_t* fb, uint32_t frame,
                         const uint8_t* bn, uint32_t rot, const Tuning& t, void* scratch, uint64_t* counters_out,
                         hipStream_t stream)
{
    const StLayout L = st_layout(cam.width_i, cam.height_i, t);
    char* base = static_cast<char*>(scratch);
    HybridCounters* counters = reinterpret_cast<HybridCounters*>(base + L.counters);
    BinnedST* binned = reinterpret_cast<BinnedST*>(base + L.binned);
    uint32_t* tile_count = reinterpret_cast<uint32_t*>(base + L.tile_count);
    uint32_t* tile_cursor = reinterpret_cast<uint32_t*>(base + L.tile_cursor);
    uint32_t* tile_offset = reinterpret_cast<uint32_t*>(base + L.tile_offset);
    uint32_t* instances = reinterpret_cast<uint32_t*>(base + L.instances);

    // counters, tile counts and cursors are contiguous
    hipError_t e = launch_zero_u32(reinterpret_cast<uint32_t*>(counters),
                                   (L.tile_cursor + 4u * (size_t)L.n_tiles - L.counters) / 4u, stream);
    if (e != hipSuccess) return e;
    const BinArgs bin = { counters, binned, t.max_binned, t.binning_pixel_threshold };
    e = launch_st_scatter(sc, w2c, cam, fb, frame, bn, rot, t.sequential_threshold, bin, counters_out, stream);
    if (e != hipSuccess) return e;
    const unsigned grid = persistent_grid(t.max_binned, 256u);
    hipLaunchKernelGGL((tile_instances_kernel<false>), dim3(grid), dim3(256), 0, stream, binned, counters,
                       t.max_binned, t.tile_size, L.grid_w, tile_count, tile_cursor, tile_offset, instances);
    hipLaunchKernelGGL((exclusive_scan_kernel<CountInput>), dim3(1), dim3(kScanThreads), 0, stream,
                       CountInput{ tile_count }, tile_offset, nullptr, L.n_tiles, L.n_tiles, &counters->n_instances,
                       &counters->overflow, t.max_tile_instances, nullptr);
    hipLaunchKernelGGL((tile_instances_kernel<true>), dim3(grid), dim3(256), 0, stream, binned, counters,
                       t.max_binned, t.tile_size, L.grid_w, tile_count, tile_cursor, tile_offset, instances);
    const bool blue = bn != nullptr;
    switch (t.tile_size) {
        case 16: launch_raster<16>(blue, L.n_tiles, stream, binned, counters, tile_offset, instances, fb, cam.width_i, cam.height_i, L.grid_w, frame, bn, rot); break;
        case 64: launch_raster<64>(blue, L.n_tiles, stream, binned, counters, tile_offset, instances, fb, cam.width_i, cam.height_i, L.grid_w, frame, bn, rot); break;
        default: launch_raster<32>(blue, L.n_tiles, stream, binned, counters, tile_offset, instances, fb, cam.width_i, cam.height_i, L.grid_w, frame, bn, rot); break;
    }
    const unsigned fgrid = persistent_grid(t.max_binned, 4u);
    if (blue)
        hipLaunchKernelGGL((binned_fallback_kernel<true>), dim3(fgrid), dim3(256), 0, stream, binned, counters,
                           t.max_binned, fb, cam.width_i, frame, bn, rot);
    else
        hipLaunchKernelGGL((binned_fallback_kernel<false>), dim3(fgrid), dim3(256), 0, stream, binned, counters,
                           t.max_binned, fb, cam.width_i, frame, bn, rot);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// GPS chunks
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kChunk) void gps_chunks_kernel(const BinnedGPS* __restrict__ binned,
                                                            const HybridCounters* __restrict__ counters,
                                                            uint32_t capacity, const uint32_t* __restrict__ offsets,
                                                            FbCamera cam, uint64_t* __restrict__ fb)
{
    uint32_t n = counters->n_binned;
    if (n > capacity) n = capacity;
    if (n == 0) return;
    const uint32_t total = offsets[n];
    for (uint32_t chunk = blockIdx.x; chunk < total; chunk += gridDim.x) {
        // last binned Gaussian whose first chunk is <= chunk
        uint32_t lo = 0, hi = n - 1u;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1u) >> 1;
            if (offsets[mid] <= chunk) lo = mid;
            else hi = mid - 1u;
        }
        const float4* p = reinterpret_cast<const float4*>(binned + lo);
        const float4 a = p[0], b = p[1], c = p[2], d = p[3];
        const uint32_t n_points = __float_as_uint(d.y);
        const uint32_t k = (chunk - offsets[lo]) * kChunk + threadIdx.x;
        if (k >= n_points) continue;
        const uint64_t frag = ((uint64_t)__float_as_uint(a.y) << 32) | __float_as_uint(a.x);
        uint32_t pix;
        if (gps_point_at(k, __float_as_uint(d.z), __float_as_uint(d.w), a.z, a.w, b.x, b.y, b.z, c.x, c.y, c.z, b.w,
                         c.w, d.x, cam, pix))
            atomicMin(reinterpret_cast<unsigned long long*>(&fb[pix]), (unsigned long long)frag);
    }
}

hipError_t run_hybrid_gps(const SceneArgs& sc, const float* w2c, const FbCamera& cam, uint64_t* fb, uint32_t frame,
                          uint32_t max_points, const Tuning& t, void* scratch, uint64_t* counters_out,
                          hipStream_t stream)
{
    const GpsLayout L = gps_layout(t);
    char* base = static_cast<char*>(scratch);
    HybridCounters* counters = reinterpret_cast<HybridCounters*>(base + L.counters);
    BinnedGPS* binned = reinterpret_cast<BinnedGPS*>(base + L.binned);
    uint32_t* offsets = reinterpret_cast<uint32_t*>(base + L.chunk_offsets);
    hipError_t e = launch_zero_u32(reinterpret_cast<uint32_t*>(counters), sizeof(HybridCounters) / 4u, stream);
    if (e != hipSuccess) return e;
    const BinArgs bin = { counters, binned, t.max_binned, t.binning_point_threshold };
    e = launch_gps_scatter(sc, w2c, cam, fb, frame, max_points, t.sequential_threshold, bin, counters_out, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((exclusive_scan_kernel<ChunkInput>), dim3(1), dim3(kScanThreads), 0, stream,
                       ChunkInput{ binned }, offsets, &counters->n_binned, 0u, t.max_binned, nullptr, nullptr, 0u, nullptr);
    hipLaunchKernelGGL(gps_chunks_kernel, dim3(kPersistentBlocks), dim3(kChunk), 0, stream, binned, counters,
                       t.max_binned, offsets, cam, fb);
    return hipGetLastError();
}

}  // namespace sgs
