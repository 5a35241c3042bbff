// Batched multi-view accumulation in one persistent launch.
//
// Phase A work items are (chunk of 256 Gaussians, view) pairs claimed with one atomic per
// workgroup. A lane projects its Gaussian once for the view, and every fragment's alpha is
// evaluated once and then tested against all spp samples of the view: sample s draws
// u_s = rand(frame0 + s, g, x, y) and is accepted iff u_s < alpha, done exactly in integers
// as (pcg_hash(...) >> 9) < ceil(alpha * 2^23), then atomicMin into that sample's own
// framebuffer. This is the plain path's per-sample test (bit-identical winners), with the
// projection and the alpha evaluation shared by the samples.
//
// Gaussians with more than big_threshold fragments are not scattered by their wave: they are
// split into entries of entry_fragments fragments on a device queue that phase B waves drain
// in the same launch, so one giant Gaussian no longer stalls a wave while the rest of the GPU
// idles. Phase B waits only on phase A items that are already running (a workgroup enters
// phase B after the item counter is exhausted), so the protocol cannot deadlock whatever
// the residency of the grid.
#include "sgs_internal.h"
#include "sgs_scan.h"

namespace sgs {

constexpr uint32_t kBatchBlock = 256;
constexpr uint32_t kSlotQueued = 1u, kSlotSkipped = 2u, kSlotExit = 3u;

// u < alpha  <=>  (h >> 9) < ceil(alpha * 2^23)  for u = (h >> 9) * 2^-23; NaN never accepts
__device__ __forceinline__ uint32_t alpha_threshold(float alpha)
{
    if (!(alpha == alpha)) return 0u;
    return (uint32_t)ceilf(alpha * 8388608.0f);
}

struct ViewRun {
    uint64_t* fbv;
    size_t P;
    uint32_t W, spp, fs0;
};

__device__ __forceinline__ ViewRun view_run(const ViewDesc& vd, uint64_t* fb)
{
    ViewRun r;
    r.fbv = fb + vd.fb_offset;
    r.W = vd.cam.width_i;
    r.P = (size_t)vd.cam.width_i * vd.cam.height_i;
    r.spp = vd.spp;
    r.fs0 = vd.frame0 * 1973u;
    return r;
}

// every sample of one fragment that passed the alpha threshold
__device__ __forceinline__ void test_samples(uint32_t pix, uint32_t base, uint32_t T, uint64_t frag, const ViewRun& vr)
{
    unsigned long long* p = reinterpret_cast<unsigned long long*>(vr.fbv + pix);
    uint32_t fs = vr.fs0;
    for (uint32_t s = 0; s < vr.spp; s++) {
        const uint32_t h = pcg_hash(base ^ fs);
        if ((h >> 9u) < T) atomicMin(p, (unsigned long long)frag);
        fs += 1973u;
        p += vr.P;
    }
}

// Stream compaction of a wave's passing fragments through LDS: the sample loop only runs
// over full batches of 64 passing fragments, so no lane idles on a rejected fragment while
// the others test their spp samples. kQueue entries per wave; count stays below 128.
constexpr uint32_t kQueue = 128;

struct WaveQueue {
    uint32_t* pix;
    uint32_t* base;
    uint32_t* thr;
    uint64_t* frag;
    uint32_t count;  // wave-uniform
};

__device__ __forceinline__ void queue_flush(WaveQueue& q, uint32_t n, uint32_t lane, const ViewRun& vr)
{
    __builtin_amdgcn_wave_barrier();
    if (lane < n) test_samples(q.pix[lane], q.base[lane], q.thr[lane], q.frag[lane], vr);
}

// fragment (px, py) of Gaussian g: evaluates alpha once; samples run when a batch is full
__device__ __forceinline__ void queue_fragment(WaveQueue& q, bool valid, uint32_t px, uint32_t py, float msx, float msy,
                                               const ConicExp2 ce, float o, uint32_t g, uint64_t frag, uint32_t lane,
                                               const ViewRun& vr)
{
    bool pass = false;
    uint32_t T = 0;
    if (valid) {
        const float alpha = fragment_alpha(ce, o, (float)px - msx, (float)py - msy);
        if (!(alpha < kMinAlpha)) {
            T = alpha_threshold(alpha);
            pass = true;
        }
    }
    const uint64_t m = __ballot(pass);
    if (!m) return;
    const uint32_t slot = q.count + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (pass) {
        q.pix[slot] = py * vr.W + px;
        q.base[slot] = g ^ (px * 9277u) ^ (py * 2699u);
        q.thr[slot] = T;
        q.frag[slot] = frag;
    }
    q.count += (uint32_t)__popcll(m);
    if (q.count >= 64u) {
        queue_flush(q, 64u, lane, vr);
        const uint32_t rest = q.count - 64u;
        uint32_t a = 0, b = 0, c = 0;
        uint64_t d = 0;
        __builtin_amdgcn_wave_barrier();
        if (lane < rest) {
            a = q.pix[64u + lane];
            b = q.base[64u + lane];
            c = q.thr[64u + lane];
            d = q.frag[64u + lane];
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < rest) {
            q.pix[lane] = a;
            q.base[lane] = b;
            q.thr[lane] = c;
            q.frag[lane] = d;
        }
        __builtin_amdgcn_wave_barrier();
        q.count = rest;
    }
}

// fragments [f_begin, f_end) of a rectangle of width w at (x0, y0), lanes striding 64
__device__ __forceinline__ void wave_fragments(WaveQueue& q, uint32_t f_begin, uint32_t f_end, uint32_t x0, uint32_t y0,
                                               uint32_t w, float msx, float msy, const ConicExp2 ce, float o,
                                               uint32_t g, uint64_t frag, const ViewRun& vr, uint32_t lane)
{
    uint32_t f = f_begin + lane;
    uint32_t qy = f / w, r = f - qy * w;
    uint32_t qs, rs;
    if (w >= 64u) {
        qs = 0;
        rs = 64u;
    } else {
        qs = 64u / w;
        rs = 64u - qs * w;
    }
    const uint32_t n_iter = (f_end - f_begin + 63u) / 64u;
    for (uint32_t it = 0; it < n_iter; it++, f += 64u) {
        queue_fragment(q, f < f_end, x0 + r, y0 + qy, msx, msy, ce, o, g, frag, lane, vr);
        r += rs;
        qy += qs;
        if (r >= w) {
            r -= w;
            qy++;
        }
    }
}

__global__ __launch_bounds__(kBatchBlock) void batch_accumulate_kernel(SceneArgs sc, const ViewDesc* __restrict__ views,
                                                                       uint32_t n_views, uint32_t n_items,
                                                                       uint32_t view_major, uint64_t* __restrict__ fb,
                                                                       BatchCounters* __restrict__ ctr,
                                                                       BigEntry* __restrict__ queue,
                                                                       uint32_t* __restrict__ ready, uint32_t queue_cap,
                                                                       uint32_t seq, uint32_t big_threshold,
                                                                       uint32_t entry_frags,
                                                                       const uint32_t* __restrict__ run_if_set)
{
    if (run_if_set && !*run_if_set) return;
    __shared__ uint32_t s_item;
    __shared__ uint32_t q_pix[kBatchBlock / 64][kQueue], q_base[kBatchBlock / 64][kQueue], q_thr[kBatchBlock / 64][kQueue];
    __shared__ uint64_t q_frag[kBatchBlock / 64][kQueue];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t wave = tid / 64u;
    WaveQueue q = { q_pix[wave], q_base[wave], q_thr[wave], q_frag[wave], 0u };
    const uint32_t n_chunks = (sc.n + kBatchBlock - 1u) / kBatchBlock;

    // ---------------- phase A: (Gaussian chunk, view) items ----------------
    for (;;) {
        if (tid == 0) s_item = atomicAdd(&ctr->next_item, 1u);
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= n_items) break;
        uint32_t chunk, v;
        if (view_major) {
            v = item / n_chunks;
            chunk = item - v * n_chunks;
        } else {
            chunk = item / n_views;
            v = item - chunk * n_views;
        }
        const ViewDesc& vd = views[v];
        if (vd.route) continue;  // accumulated by the single-view scatter kernels
        const ViewRun vr = view_run(vd, fb);
        const uint32_t g = chunk * kBatchBlock + tid;
        Projected p;
        bool active =
            g < sc.n && project_scene(sc, g, vd.r1, vd.r2, vd.r3,
                                         vd.cam, p);
        uint32_t x0 = 0, y0 = 0, w = 0, count = 0;
        float msx = 0.0f, msy = 0.0f, o = 0.0f;
        ConicExp2 ce = { 0.0f, 0.0f, 0.0f };
        uint64_t frag = 0;
        if (active) {
            const Aabb b = pixel_aabb(p, vd.cam);
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
        // large Gaussians go to the queue in fragment ranges
        if (active && count > big_threshold) {
            const uint32_t n_e = (count + entry_frags - 1u) / entry_frags;
            const uint32_t base = atomicAdd(&ctr->produced, n_e);
            const bool queued = base + n_e <= queue_cap;
            for (uint32_t e = 0; e < n_e; e++) {
                const uint32_t k = base + e;
                if (k >= queue_cap) break;
                if (queued) {
                    const uint32_t f0 = e * entry_frags;
                    const uint32_t f1 = min(count, f0 + entry_frags);
                    float4* dst = reinterpret_cast<float4*>(queue + k);
                    dst[0] = make_float4(__uint_as_float((uint32_t)frag), __uint_as_float((uint32_t)(frag >> 32)), msx,
                                         msy);
                    dst[1] = make_float4(ce.a, ce.b, ce.c, o);
                    dst[2] = make_float4(__uint_as_float(x0), __uint_as_float(y0), __uint_as_float(w),
                                         __uint_as_float(v));
                    dst[3] = make_float4(__uint_as_float(f0), __uint_as_float(f1), 0.0f, 0.0f);
                }
                __hip_atomic_store(&ready[k], queued ? kSlotQueued : kSlotSkipped, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            if (queued) active = false;
        }
        // small Gaussians: the owning lane walks its AABB, in a wave-uniform loop
        const uint32_t my_small = (active && count <= seq) ? count : 0u;
        uint32_t n_small = my_small;
        for (int off = 32; off > 0; off >>= 1) n_small = max(n_small, (uint32_t)__shfl_xor((int)n_small, off, 64));
        n_small = __builtin_amdgcn_readfirstlane(n_small);
        {
            uint32_t r = 0, qy = 0;
            for (uint32_t f = 0; f < n_small; f++) {
                queue_fragment(q, f < my_small, x0 + r, y0 + qy, msx, msy, ce, o, g, frag, lane, vr);
                if (++r == w) {
                    r = 0;
                    qy++;
                }
            }
        }
        // the rest: the whole wave, one Gaussian at a time (parameters through SGPRs)
        uint64_t mask = __ballot(active && count > seq);
        while (mask) {
            const int l = __builtin_ctzll(mask);
            mask &= mask - 1ull;
            const ConicExp2 cce = { bcast(ce.a, l), bcast(ce.b, l), bcast(ce.c, l) };
            wave_fragments(q, 0u, bcast(count, l), bcast(x0, l), bcast(y0, l), bcast(w, l), bcast(msx, l),
                           bcast(msy, l), cce, bcast(o, l), bcast(g, l), bcast(frag, l), vr, lane);
        }
        queue_flush(q, q.count, lane, vr);
        q.count = 0;
        // every wave's queue writes are complete device-wide before the item counts as done
        __threadfence();
        __syncthreads();
        if (tid == 0) __hip_atomic_fetch_add(&ctr->done_items, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }

    // ---------------- phase B: queue entries, one per wave ----------------
    for (;;) {
        uint32_t k = 0;
        if (lane == 0u) k = atomicAdd(&ctr->consumed, 1u);
        k = bcast(k, 0);
        if (k >= queue_cap) break;
        uint32_t state = 0;
        if (lane == 0u) {
            for (;;) {
                state = __hip_atomic_load(&ready[k], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                if (state) break;
                const uint32_t done = __hip_atomic_load(&ctr->done_items, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                if (done >= n_items) {
                    // every reservation is final and published: k is either set or unused
                    state = __hip_atomic_load(&ready[k], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                    if (!state) state = kSlotExit;
                    break;
                }
                __builtin_amdgcn_s_sleep(4);
            }
        }
        state = bcast(state, 0);
        if (state == kSlotExit) break;
        if (state != kSlotQueued) continue;
        const float4* src = reinterpret_cast<const float4*>(queue + k);
        const float4 a = src[0], b = src[1], c = src[2], d = src[3];
        const uint64_t frag = ((uint64_t)__float_as_uint(a.y) << 32) | __float_as_uint(a.x);
        const uint32_t v = __float_as_uint(c.w);
        const ViewRun vr = view_run(views[v], fb);
        const ConicExp2 ce = { b.x, b.y, b.z };
        wave_fragments(q, __float_as_uint(d.x), __float_as_uint(d.y), __float_as_uint(c.x), __float_as_uint(c.y),
                       __float_as_uint(c.z), a.z, a.w, ce, b.w, (uint32_t)frag, frag, vr, lane);
        queue_flush(q, q.count, lane, vr);
        q.count = 0;
    }
}

hipError_t launch_batch_accumulate(const SceneArgs& sc, const ViewDesc* views, uint32_t n_views, uint64_t* fb,
                                   BatchCounters* counters, BigEntry* queue, uint32_t* ready, const BatchTuning& t,
                                   const uint32_t* run_if_set, hipStream_t stream)
{
    if (sc.n == 0 || n_views == 0) return hipSuccess;
    const uint32_t n_chunks = (sc.n + kBatchBlock - 1u) / kBatchBlock;
    const uint64_t n_items64 = (uint64_t)n_chunks * n_views;
    if (n_items64 >= 0xffffffffull) return hipErrorInvalidValue;
    int device = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(batch_accumulate_kernel),
                                                     kBatchBlock, 0);
    if (e != hipSuccess) return e;
    if (t.blocks_per_cu && (int)t.blocks_per_cu < per_cu) per_cu = (int)t.blocks_per_cu;
    if (per_cu < 1) per_cu = 1;
    uint64_t grid = (uint64_t)cus * per_cu;
    if (grid > n_items64) grid = n_items64;
    hipLaunchKernelGGL(batch_accumulate_kernel, dim3((unsigned)grid), dim3(kBatchBlock), 0, stream, sc, views, n_views,
                       (uint32_t)n_items64, t.view_major, fb, counters, queue, ready, t.queue_capacity,
                       t.sequential_threshold, t.big_threshold, t.entry_fragments, run_if_set);
    return hipGetLastError();
}

// ============================================================================================
// Tile schedule: bin every (view, Gaussian) into the view's framebuffer bins (16, 32 or 64
// pixels) its AABB overlaps, sort each bin's list by depth bucket, then rasterise each
// 16x16 tile of every bin in one workgroup that keeps the tile's framebuffer words of up to 16
// samples in LDS. Each fragment's alpha is evaluated once; a sample is hashed only if
// the fragment is closer than that sample's current LDS word, and accepted fragments land with
// ds_min_u64; the tile is written out once with plain stores. No framebuffer clear is needed
// because every tile writes all of its pixels. The tested pixels, alpha and thresholds are
// those of the scatter path, so the framebuffers are bit-identical to it.
// ============================================================================================
constexpr uint32_t kTile = kRasterPixels;
constexpr uint32_t kWindow = 512;    // tiles of a workgroup's LDS histogram window
// Each window cell has kHistCopies adjacent counters (distinct LDS banks); lane l counts into
// copy l % kHistCopies, so the many lanes of a spatially compact chunk that hit one bin
// spread over that many addresses instead of serialising on one. The counting pass (a plain
// add per instance) measured fastest with four copies, the fill pass (an add, then a
// subtract with return, per instance) with two.
template <bool kFill>
constexpr uint32_t kHistCopiesOf = kFill ? 2u : 4u;

struct TileRange {
    uint32_t tx0, tx1, ty0, ty1;  // inclusive
};

// tile range and depth key (float bits of the view depth) in view vd of a Gaussian whose mean
// and cached view-independent terms are in registers (the operations of project_cached, so
// the same bits); false if culled or its AABB is empty
__device__ __forceinline__ bool gaussian_tiles_reg(float mx, float my, float mz, const Cov3d& cv, bool valid,
                                                   const ViewDesc& vd, uint32_t bin_shift, TileRange& t, uint32_t& key)
{
    if (!valid) return false;
    const float4 r3 = vd.r3;
    const float depth = r3.x * mx + r3.y * my + r3.z * mz + r3.w;
    if (!(depth >= vd.cam.near_plane && depth <= vd.cam.far_plane)) return false;
    Projected p;
    if (!project_cov3d(mx, my, mz, depth, cv, vd.r1, vd.r2, r3, vd.cam, p)) return false;
    key = f2u(p.depth);
    const Aabb b = pixel_aabb(p, vd.cam);
    if (b.x1 <= b.x0 || b.y1 <= b.y0) return false;
    t.tx0 = b.x0 >> bin_shift;
    t.tx1 = (b.x1 - 1u) >> bin_shift;
    t.ty0 = b.y0 >> bin_shift;
    t.ty1 = (b.y1 - 1u) >> bin_shift;
    return true;
}

// A superset of the tile range gaussian_tiles_reg gives, at about half its cost, for the
// counting pass. The near/far test and the numerators of the projected mean are the exact
// operations; the divisions by the depth go through the hardware reciprocal (within 1 ulp),
// the diagonal of the 2D covariance is evaluated with fused multiply-adds, the determinant and
// antialiasing culls are skipped (a culled Gaussian only costs instances), and the AABB is
// widened by bounds on the distance of every approximate value from the exact one:
// * x = X / depth: the reciprocal and the product are within 2^-21 |x| of the exact quotient.
// * cov_xx = a^T C a with a = j11 (r1 - xc r3): both evaluations, from Jacobians within 2^-20
//   |j11| (|r1| + |xc| |r3|) + |j11| |r3| dx of each other, differ by at most
//   tr(C) j11^2 S (2^-18 S + 2 dx |r3|) with S >= |r1| + (|xc| + dx) |r3|: by Cauchy-Schwarz,
//   |u^T C v| <= |u| |v| tr(C) for the positive semidefinite C, and the roundings of the two
//   six-operation dot products add at most 12 2^-24 |a|^2 tr(C).
// * the cutoff: the cached log(255 opacity) is at least the per-view value (antialiasing only
//   lowers the opacity); the half-width sqrt((cov_xx + kernel) cutoff) is scaled by 1 + 2^-16.
// * the mean and the edges: 2^-20 (|x fx| + |mean| + half-width) covers the quotient's error
//   through the mean and every later rounding.
// Row norms are bounded by their L1 norms. A Gaussian whose bounds are not finite takes the
// exact path. The rasteriser projects exactly and clips every instance to its exact AABB, so
// an extra tile changes no result.
// the per-view terms of gaussian_tiles_bound, computed once per view and wave and kept in
// scalar registers
struct BoundView {
    float n1, n2, n3;  // L1 bounds of the rotation rows' Euclidean norms, rounded up
    float kernel;      // the antialiasing kernel added to the covariance diagonal
};

__device__ __forceinline__ float uniform_f32(float x)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

__device__ __forceinline__ BoundView bound_view(const ViewDesc& vd)
{
    const float4 r1 = vd.r1, r2 = vd.r2, r3 = vd.r3;
    return { uniform_f32((fabsf(r1.x) + fabsf(r1.y) + fabsf(r1.z)) * (1.0f + 0x1p-20f)),
             uniform_f32((fabsf(r2.x) + fabsf(r2.y) + fabsf(r2.z)) * (1.0f + 0x1p-20f)),
             uniform_f32((fabsf(r3.x) + fabsf(r3.y) + fabsf(r3.z)) * (1.0f + 0x1p-20f)),
             uniform_f32((vd.cam.proper_antialiasing ? 0.1f : 0.3f) * vd.cam.ss * vd.cam.ss) };
}

// Every lane evaluates the whole bound (nearly every pair is in front of the camera), and the
// tests combine into one condition, so the conversions to tile indices run under a single
// branch.
__device__ __forceinline__ bool gaussian_tiles_bound(float mx, float my, float mz, const Cov3d& cv, float trace,
                                                     bool valid, const ViewDesc& vd, const BoundView& bv,
                                                     uint32_t shift, TileRange& t, bool& exact)
{
    const float4 r1 = vd.r1, r2 = vd.r2, r3 = vd.r3;
    const FbCamera& cam = vd.cam;
    const float depth = r3.x * mx + r3.y * my + r3.z * mz + r3.w;
    const bool in_range = valid && depth >= cam.near_plane && depth <= cam.far_plane;
    const float X = r1.x * mx + r1.y * my + r1.z * mz + r1.w;
    const float Y = r2.x * mx + r2.y * my + r2.z * mz + r2.w;
    const float rz = __builtin_amdgcn_rcpf(depth);
    const float x = X * rz, y = Y * rz;
    const float dx = fabsf(x) * 0x1p-21f, dy = fabsf(y) * 0x1p-21f;
    // the projection's clamp (clip_left <= clip_right): the median of three for any x that is
    // not NaN, and a NaN x makes the mean below not finite
    const float xc = __builtin_amdgcn_fmed3f(x, cam.clip_left, cam.clip_right);
    const float yc = __builtin_amdgcn_fmed3f(y, cam.clip_top, cam.clip_bottom);
    const float j11 = cam.focal_x * rz, j22 = cam.focal_y * rz;
    const float j13 = -j11 * xc, j23 = -j22 * yc;
    const float a1 = fmaf(j13, r3.x, j11 * r1.x), a2 = fmaf(j13, r3.y, j11 * r1.y), a3 = fmaf(j13, r3.z, j11 * r1.z);
    const float b1 = fmaf(j23, r3.x, j22 * r2.x), b2 = fmaf(j23, r3.y, j22 * r2.y), b3 = fmaf(j23, r3.z, j22 * r2.z);
    const float p1 = fmaf(a3, cv.c13, fmaf(a2, cv.c12, a1 * cv.c11));
    const float p2 = fmaf(a3, cv.c23, fmaf(a2, cv.c22, a1 * cv.c12));
    const float p3 = fmaf(a3, cv.c33, fmaf(a2, cv.c23, a1 * cv.c13));
    const float q1 = fmaf(b3, cv.c13, fmaf(b2, cv.c12, b1 * cv.c11));
    const float q2 = fmaf(b3, cv.c23, fmaf(b2, cv.c22, b1 * cv.c12));
    const float q3 = fmaf(b3, cv.c33, fmaf(b2, cv.c23, b1 * cv.c13));
    const float cxx = fmaf(p3, a3, fmaf(p2, a2, p1 * a1));
    const float cyy = fmaf(q3, b3, fmaf(q2, b2, q1 * b1));
    const float n1 = bv.n1, n2 = bv.n2, n3 = bv.n3;
    const float sx = fmaf(fabsf(xc) + dx, n3, n1), sy = fmaf(fabsf(yc) + dy, n3, n2);
    const float tr = trace * (1.0f + 0x1p-16f);
    const float exx = tr * (j11 * j11) * sx * fmaf(sx, 0x1p-18f, 2.0f * dx * n3);
    const float eyy = tr * (j22 * j22) * sy * fmaf(sy, 0x1p-18f, 2.0f * dy * n3);
    const float kernel = bv.kernel;
    const float cut = fmaf(fmaxf(2.0f * cv.pt, 0.0f), 1.0f + 0x1p-16f, 0x1p-20f);
    const float ex = __builtin_amdgcn_sqrtf(((cxx + exx) + kernel) * cut) * (1.0f + 0x1p-16f);
    const float ey = __builtin_amdgcn_sqrtf(((cyy + eyy) + kernel) * cut) * (1.0f + 0x1p-16f);
    const float mean_x = x * cam.focal_x + cam.center_x, mean_y = y * cam.focal_y + cam.center_y;
    const float wx = ex + fmaf(fabsf(x * cam.focal_x) + fabsf(mean_x) + ex, 0x1p-20f, 0x1p-20f);
    const float wy = ey + fmaf(fabsf(y * cam.focal_y) + fabsf(mean_y) + ey, 0x1p-20f, 0x1p-20f);
    const float bx0 = mean_x - wx, bx1 = mean_x + wx, by0 = mean_y - wy, by1 = mean_y + wy;
    const bool finite = wx <= 3.0e38f && wy <= 3.0e38f && fabsf(mean_x) <= 3.0e38f && fabsf(mean_y) <= 3.0e38f;
    exact = in_range && !finite;
    bool has = in_range && finite && bx0 < cam.width && bx1 >= 0.0f && by0 < cam.height && by1 >= 0.0f;
    t = { 0u, 0u, 0u, 0u };
    if (has) {
        // finite edges with bx0 < width and bx1 >= 0: the medians are max(bx0, 0) and min(bx1, width)
        const uint32_t x0 = (uint32_t)floorf(__builtin_amdgcn_fmed3f(bx0, 0.0f, cam.width));
        const uint32_t x1 = (uint32_t)ceilf(__builtin_amdgcn_fmed3f(bx1, 0.0f, cam.width));
        const uint32_t y0 = (uint32_t)floorf(__builtin_amdgcn_fmed3f(by0, 0.0f, cam.height));
        const uint32_t y1 = (uint32_t)ceilf(__builtin_amdgcn_fmed3f(by1, 0.0f, cam.height));
        has = x1 > x0 && y1 > y0;
        t = { x0 >> shift, (x1 - 1u) >> shift, y0 >> shift, (y1 - 1u) >> shift };
    }
    return has;
}

// Depth buckets of the per-bin sort (tile_sort_kernel): the fill pass stores each instance's
// bucket, the sort counts and places by it. One bucket below the view's depth range, 256
// across it, the rest above (measured against 1024 buckets with 512 across: the sort 0.23 ms
// faster, the rasteriser's front-to-back order costing it 0.08 ms).
constexpr uint32_t kSortBins = 512;
constexpr uint32_t kSortRangeLog2 = 8;  // buckets across the range: 2^kSortRangeLog2
constexpr uint32_t kSortRegKeys = 32;  // entries per lane kept in registers for tiles of <= 8192
constexpr uint32_t kSortInFlight = 16;  // entries per lane in flight beyond the registers' entries
constexpr uint32_t kSortCountInFlight = 8;  // the same while counting, beside the registers' entries
constexpr uint32_t kSortStage = 12288;  // buckets held in LDS beyond the registers' entries (24 KB)
static_assert(kSortStage % (kSortInFlight * 256u) == 0u && kSortStage % (kSortCountInFlight * 256u) == 0u,
              "staged groups are whole");

__device__ __forceinline__ uint32_t sort_bucket(uint32_t key, uint2 vk)
{
    if (key < vk.x) return 0u;
    return 1u + min((key - vk.x) >> vk.y, kSortBins - 2u);
}

__device__ __forceinline__ uint32_t bucket_floor(uint32_t b, uint2 vk)
{
    if (b == 0u) return 0u;
    const uint64_t f = (uint64_t)vk.x + ((uint64_t)(b - 1u) << vk.y);
    return f > 0xffffffffull ? 0xffffffffu : (uint32_t)f;
}

// Gaussians touching at most kSmallTiles tiles are counted through the workgroup's LDS
// histogram (one global atomic per touched tile and workgroup); larger ones are spread over
// their wave's lanes with one global atomic per tile.
constexpr uint32_t kSmallTiles = 32;

// Bins of 2^b x 2^b tiles: an instance word carries, above the Gaussian index, the
// sub-rectangle [x0, x1] x [y0, y1] of its bin's tiles that the pair's tile rectangle touches
// (4b bits), so each tile's rasteriser projects only the instances that reach it. Stored when
// the index leaves the room (n <= 2^(32 - 4b)); returns the bit position (32: not stored).
__host__ __device__ __forceinline__ uint32_t subrect_shift(uint32_t bin_log, uint32_t n)
{
    if (bin_log == 0u) return 32u;
    const uint32_t s = 32u - 4u * bin_log;
    return (uint64_t)n <= (1ull << s) ? s : 32u;
}

// instance word of pair (t16: its rectangle of 16x16 tiles) in bin (bx, by)
__device__ __forceinline__ uint32_t instance_word(uint32_t g, uint32_t bx, uint32_t by, const TileRange& t16,
                                                  uint32_t bin_log, uint32_t shift)
{
    if (shift >= 32u) return g;
    const uint32_t m = (1u << bin_log) - 1u;
    const uint32_t x0 = max(t16.tx0, bx << bin_log) & m, x1 = min(t16.tx1, (bx << bin_log) | m) & m;
    const uint32_t y0 = max(t16.ty0, by << bin_log) & m, y1 = min(t16.ty1, (by << bin_log) | m) & m;
    return g | ((x0 | (x1 << bin_log) | (y0 << (2u * bin_log)) | (y1 << (3u * bin_log))) << shift);
}

template <bool kFill>
__device__ __forceinline__ void bin_direct(uint32_t gt, uint32_t word, uint32_t bucket, uint32_t* tile_count,
                                           uint32_t* tile_cursor, const uint32_t* offsets, const EntryArrays& entries)
{
    if constexpr (kFill) {
        const uint32_t slot = offsets[gt] + atomicAdd(&tile_cursor[gt], 1u);
        entries.store(slot, bucket, word);
    } else {
        atomicAdd(&tile_count[gt], 1u);
    }
}

// Wave reductions through DPP lane shifts (VALU) rather than ds_bpermute, which would share
// the LDS issue queue with the binning histograms: row_shr 1, 2, 4, 8 reduce each row of 16
// lanes into its lane 15 (lanes without a source keep the identity), row_bcast 15 and 31 carry
// rows 0-2 into lane 63, whose value is returned to every lane.
template <typename Op>
__device__ __forceinline__ uint32_t wave_reduce_dpp(uint32_t x, uint32_t identity, Op op)
{
    const int id = (int)identity;
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x118, 0xf, 0xf, false));  // row_shr:8
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x)
{
    return wave_reduce_dpp(x, 0xffffffffu, [](uint32_t a, uint32_t b) { return min(a, b); });
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x)
{
    return wave_reduce_dpp(x, 0u, [](uint32_t a, uint32_t b) { return max(a, b); });
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x)
{
    return wave_reduce_dpp(x, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}

// view-independent part of every Gaussian's projection, once per render
__global__ __launch_bounds__(kBatchBlock) void cov3d_kernel(SceneArgs sc, float4* __restrict__ cache)
{
    const uint32_t g = blockIdx.x * kBatchBlock + threadIdx.x;
    if (g >= sc.n) return;
    Cov3d c = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
    const bool valid = gaussian_cov3d(sc.scales, sc.rotations, sc.opacities, g, c);
    store_cov3d(cache, g, valid, c);
}

hipError_t launch_cov3d(const SceneArgs& sc, float4* cache, hipStream_t stream)
{
    if (sc.n == 0) return hipSuccess;
    hipLaunchKernelGGL(cov3d_kernel, dim3((sc.n + kBatchBlock - 1u) / kBatchBlock), dim3(kBatchBlock), 0, stream, sc,
                       cache);
    return hipGetLastError();
}

// tile coordinates fit 16 bits (framebuffers are at most 65535 pixels wide); a pair with no
// tiles is marked by all ones
constexpr uint32_t kNoTiles = 0xffffffffu;

// The counting pass's record of each (view, Gaussian) pair, read back by the fill pass: its
// rectangle of 16x16 tiles, 8 bytes (16-bit corners), or 4 bytes (8-bit corners) when every
// view's tile grid has at most 255 columns and rows (framebuffers up to 4080 pixels a side),
// so that all ones, the mark of a pair without tiles, is no rectangle.
struct PairRecords {
    void* base;
    uint32_t compact;
    __device__ __forceinline__ uint2 load(size_t i) const
    {
        if (compact) return make_uint2(static_cast<const uint32_t*>(base)[i], 0u);
        return static_cast<const uint2*>(base)[i];
    }
    __device__ __forceinline__ void store(size_t i, bool has, const TileRange& t) const
    {
        if (compact)
            static_cast<uint32_t*>(base)[i] = has ? t.tx0 | (t.tx1 << 8) | (t.ty0 << 16) | (t.ty1 << 24) : kNoTiles;
        else
            static_cast<uint2*>(base)[i] = has ? make_uint2(t.tx0 | (t.ty0 << 16), t.tx1 | (t.ty1 << 16))
                                               : make_uint2(kNoTiles, kNoTiles);
    }
    // the rectangle of a loaded record; false for a pair without tiles
    __device__ __forceinline__ bool decode(uint2 r, TileRange& t) const
    {
        if (compact) t = { r.x & 0xffu, (r.x >> 8) & 0xffu, (r.x >> 16) & 0xffu, r.x >> 24 };
        else t = { r.x & 0xffffu, r.y & 0xffffu, r.x >> 16, r.y >> 16 };
        return r.x != kNoTiles;
    }
};

// Gaussians per lane in the binning passes (counting, fill): a workgroup bins chunks of
// kBinPer x 256 Gaussians per view step, so the per-step work shared by the workgroup (window
// reductions, barriers, the histogram's flush with one global atomic per touched bin) is paid
// once per kBinPer x 256 pairs.
template <bool kFill>
constexpr uint32_t kBinPer = kFill ? 1u : 4u;  // measured: count 2.85 ms with 2, 2.96 with 3, 2.73 with 4
template <bool kFill>
constexpr uint32_t kBinChunk = kBinPer<kFill> * kBatchBlock;

template <bool kFill>
__global__ __launch_bounds__(kBatchBlock, kFill ? 1 : 4) void batch_bin_kernel(SceneArgs sc, const float4* __restrict__ cov3d,
                                                                const ViewDesc* __restrict__ views,
                                                                uint32_t n_views,
                                                                uint32_t* __restrict__ tile_count,
                                                                uint32_t* __restrict__ tile_cursor,
                                                                const uint32_t* __restrict__ offsets,
                                                                EntryArrays entries,
                                                                const uint2* __restrict__ view_keys,
                                                                PairRecords pair_rect,
                                                                BatchCounters* __restrict__ ctr)
{
    constexpr uint32_t kHistCopies = kHistCopiesOf<kFill>;
    // window of the small Gaussians (x0, x1, y0, y1): one buffer reset at each view start in
    // the counting pass; in the fill pass two, a view's buffer being reset while the next one
    // is in use, so no barrier precedes a view's window atomics (each layout measured the
    // faster one for its pass)
    __shared__ uint32_t s_wx0, s_wx1, s_wy0, s_wy1;
    __shared__ uint32_t s_win[2][4];
    __shared__ unsigned long long s_total, s_visible;
    // Counters of the window's bins. The counting pass clears them at each view; in the fill pass
    // they are zero between views, as each bin's slots are counted back down to zero.
    __shared__ uint32_t hist[kWindow * kHistCopies];
    __shared__ uint32_t hbase[kWindow * kHistCopies];
    constexpr uint32_t kPer = kBinPer<kFill>;
    if constexpr (kFill) {
        if (ctr->overflow) return;
    }
    const uint32_t tid = threadIdx.x, lane = lane_id();
    const uint32_t hc = lane & (kHistCopies - 1u);
    const uint32_t n_chunks = (sc.n + kBinChunk<kFill> - 1u) / kBinChunk<kFill>;
    if (tid == 0) {
        s_total = 0ull;
        s_visible = 0ull;
    }
    if constexpr (kFill) {
        if (tid < 8u) s_win[tid >> 2][tid & 3u] = (tid & 1u) ? 0u : 0xffffffffu;
        for (uint32_t i = tid; i < kWindow * kHistCopies; i += kBatchBlock) hist[i] = 0u;
        __syncthreads();
    }
    uint32_t wb = 0;  // window buffer of the next binned view
    // Chunks are dealt round-robin (a shared claim counter would serialise one atomic per
    // chunk on one address; the per-chunk work is nearly uniform). A lane loads its
    // Gaussians' parameters once and bins them for every view in turn, so the parameter
    // stream is read from memory once per pass instead of once per view. The counting pass
    // records each (view, Gaussian) pair's rectangle of 16x16 tiles (a superset of the exact
    // one, from gaussian_tiles_bound); the fill pass reads it back instead of projecting again,
    // with the depth key recomputed from the mean. Both bin the pair into the bins holding
    // those tiles. Lane l holds Gaussians chunk * kPer * 256 + k * 256 + l (coalesced loads).
    for (uint32_t chunk = blockIdx.x; chunk < n_chunks; chunk += gridDim.x) {
        uint32_t g[kPer];
        float mx[kPer], my[kPer], mz[kPer];
        Cov3d cv[kPer];
        uint2 next_rect[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            g[k] = chunk * kBinChunk<kFill> + k * kBatchBlock + tid;
            mx[k] = my[k] = mz[k] = 0.0f;
            cv[k] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, -1.0f, 0.0f };
            if (g[k] < sc.n) {
                mx[k] = sc.means[3u * g[k] + 0u];
                my[k] = sc.means[3u * g[k] + 1u];
                mz[k] = sc.means[3u * g[k] + 2u];
            }
            if (!kFill && g[k] < sc.n) {
                const float4 a = cov3d[2u * g[k]], b = cov3d[2u * g[k] + 1u];
                cv[k] = { a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w };
            }
            // the fill pass reads each view's pair record one view ahead, so the load of view
            // v + 1 is in flight while view v is binned
            next_rect[k] = make_uint2(kNoTiles, kNoTiles);
            if (kFill && g[k] < sc.n) next_rect[k] = pair_rect.load(g[k]);
        }
        for (uint32_t v = 0; v < n_views; v++) {
            if (views[v].route) {
                // a view without bins: only keep the fill pass's one-view-ahead record read going
                if (kFill && v + 1u < n_views)
#pragma unroll
                    for (uint32_t k = 0; k < kPer; k++)
                        if (g[k] < sc.n) next_rect[k] = pair_rect.load((size_t)(v + 1u) * sc.n + g[k]);
                continue;
            }
            if constexpr (!kFill) {
                if (tid == 0) {
                    s_wx0 = 0xffffffffu;
                    s_wy0 = 0xffffffffu;
                    s_wx1 = 0u;
                    s_wy1 = 0u;
                }
                __syncthreads();
            }
            const ViewDesc& vd = views[v];
            const uint32_t bin0 = vd.tile_begin, gw = vd.grid_w, bl = vd.bin_log;
            const uint32_t ms = subrect_shift(bl, sc.n);
            BoundView bv = {};
            if constexpr (!kFill) bv = bound_view(vd);
            TileRange t16[kPer], tr[kPer];
            uint32_t key[kPer], tw[kPer], n_t[kPer];
            bool has[kPer], small[kPer];
            uint32_t wx0l = 0xffffffffu, wy0l = 0xffffffffu, wx1l = 0u, wy1l = 0u, sum = 0u, vis = 0u;
#pragma unroll
            for (uint32_t k = 0; k < kPer; k++) {
                t16[k] = { 0u, 0u, 0u, 0u };
                key[k] = 0u;
                const size_t pair = (size_t)v * sc.n + g[k];
                if constexpr (kFill) {
                    const uint2 r = next_rect[k];
                    if (g[k] < sc.n && v + 1u < n_views) next_rect[k] = pair_rect.load(pair + sc.n);
                    has[k] = pair_rect.decode(r, t16[k]);
                    // the depth key again from the mean in registers (the operations of
                    // gaussian_tiles_reg, so the same bits), stored as its sort bucket
                    const float4 r3 = vd.r3;
                    key[k] = sort_bucket(f2u(r3.x * mx[k] + r3.y * my[k] + r3.z * mz[k] + r3.w), view_keys[v]);
                } else {
                    const bool valid = g[k] < sc.n && !(cv[k].opacity < 0.0f);
                    const float trace = cv[k].c11 + cv[k].c22 + cv[k].c33;
                    bool exact = false;
                    has[k] = gaussian_tiles_bound(mx[k], my[k], mz[k], cv[k], trace, valid, vd, bv, kRasterLog2,
                                                  t16[k], exact);
                    if (exact)
                        has[k] = gaussian_tiles_reg(mx[k], my[k], mz[k], cv[k], valid, vd, kRasterLog2, t16[k], key[k]);
                    if (g[k] < sc.n) pair_rect.store(pair, has[k], t16[k]);
                }
                tr[k] = { t16[k].tx0 >> bl, t16[k].tx1 >> bl, t16[k].ty0 >> bl, t16[k].ty1 >> bl };  // bins
                tw[k] = has[k] ? tr[k].tx1 - tr[k].tx0 + 1u : 0u;
                // bin rows and columns are below 2^16, so the 24-bit multiply is the product
                n_t[k] = has[k] ? __umul24(tw[k], tr[k].ty1 - tr[k].ty0 + 1u) : 0u;
                small[k] = has[k] && n_t[k] <= kSmallTiles;
                wx0l = min(wx0l, small[k] ? tr[k].tx0 : 0xffffffffu);
                wy0l = min(wy0l, small[k] ? tr[k].ty0 : 0xffffffffu);
                wx1l = max(wx1l, small[k] ? tr[k].tx1 : 0u);
                wy1l = max(wy1l, small[k] ? tr[k].ty1 : 0u);
                sum += n_t[k];
                vis += has[k] ? 1u : 0u;
            }
            {
                // window of the small Gaussians: wave reductions, one LDS atomic per wave
                const uint32_t mx0 = wave_min_u32(wx0l), my0 = wave_min_u32(wy0l);
                const uint32_t mx1 = wave_max_u32(wx1l), my1 = wave_max_u32(wy1l);
                if (lane == 0u && mx0 <= mx1) {
                    atomicMin(kFill ? &s_win[wb][0] : &s_wx0, mx0);
                    atomicMax(kFill ? &s_win[wb][1] : &s_wx1, mx1);
                    atomicMin(kFill ? &s_win[wb][2] : &s_wy0, my0);
                    atomicMax(kFill ? &s_win[wb][3] : &s_wy1, my1);
                }
            }
            if constexpr (!kFill) {
                // a Gaussian touches < 2^24 tiles, so a wave's sum fits 32 bits
                const uint32_t wsum = wave_sum_u32(sum), wvis = wave_sum_u32(vis);
                if (lane == 0u && wsum) atomicAdd(&s_total, (unsigned long long)wsum);
                if (lane == 0u && wvis) atomicAdd(&s_visible, (unsigned long long)wvis);
            }
            // large Gaussians: the wave walks each one's tiles
#pragma unroll
            for (uint32_t k = 0; k < kPer; k++) {
                uint64_t big = __ballot(has[k] && !small[k]);
                while (big) {
                    const int l = __builtin_ctzll(big);
                    big &= big - 1ull;
                    const uint32_t bx0 = bcast(tr[k].tx0, l), by0 = bcast(tr[k].ty0, l), bw = bcast(tw[k], l);
                    const uint32_t bn = bcast(n_t[k], l), bg = bcast(g[k], l), bkey = bcast(key[k], l);
                    const TileRange b16 = { bcast(t16[k].tx0, l), bcast(t16[k].tx1, l), bcast(t16[k].ty0, l),
                                            bcast(t16[k].ty1, l) };
                    for (uint32_t kk = lane; kk < bn; kk += 64u) {
                        const uint32_t ty = by0 + kk / bw, tx = bx0 + kk % bw;
                        bin_direct<kFill>(bin0 + ty * gw + tx, instance_word(bg, tx, ty, b16, bl, ms), bkey,
                                          tile_count, tile_cursor, offsets, entries);
                    }
                }
            }
            __syncthreads();
            const uint32_t wx0 = kFill ? s_win[wb][0] : s_wx0, wx1 = kFill ? s_win[wb][1] : s_wx1;
            const uint32_t wy0 = kFill ? s_win[wb][2] : s_wy0, wy1 = kFill ? s_win[wb][3] : s_wy1;
            // every wave has read the previous view's window (before this barrier): reset it for
            // the next view (ordered before that view's atomics by the barriers below)
            if constexpr (kFill) {
                if (tid < 4u) s_win[wb ^ 1u][tid] = (tid & 1u) ? 0u : 0xffffffffu;
                wb ^= 1u;
            }
            if (wx0 <= wx1) {
                const uint32_t wt = wx1 - wx0 + 1u, ht = wy1 - wy0 + 1u;
                if (wt * ht <= kWindow) {
                    const uint32_t nw = wt * ht;
                    if constexpr (!kFill) {
                        for (uint32_t i = tid; i < nw * kHistCopies; i += kBatchBlock) hist[i] = 0u;
                        __syncthreads();
                    }
#pragma unroll
                    for (uint32_t k = 0; k < kPer; k++)
                        if (small[k])
                            for (uint32_t ty = tr[k].ty0; ty <= tr[k].ty1; ty++)
                                for (uint32_t tx = tr[k].tx0; tx <= tr[k].tx1; tx++)
                                    atomicAdd(&hist[((ty - wy0) * wt + tx - wx0) * kHistCopies + hc], 1u);
                    __syncthreads();
                    for (uint32_t i = tid; i < nw; i += kBatchBlock) {
                        uint32_t cc[kHistCopies], c = 0;
#pragma unroll
                        for (uint32_t q = 0; q < kHistCopies; q++) {
                            cc[q] = hist[i * kHistCopies + q];
                            c += cc[q];
                        }
                        if (!c) continue;
                        const uint32_t gt = bin0 + (wy0 + i / wt) * gw + wx0 + i % wt;
                        if constexpr (kFill) {
                            // copy q's instances take the slots after copies 0 .. q-1
                            uint32_t b = offsets[gt] + atomicAdd(&tile_cursor[gt], c);
#pragma unroll
                            for (uint32_t q = 0; q < kHistCopies; q++) {
                                hbase[i * kHistCopies + q] = b;
                                b += cc[q];
                            }
                        } else {
                            atomicAdd(&tile_count[gt], c);
                        }
                    }
                    if constexpr (kFill) {
                        __syncthreads();
#pragma unroll
                        for (uint32_t k = 0; k < kPer; k++)
                            if (small[k])
                                for (uint32_t ty = tr[k].ty0; ty <= tr[k].ty1; ty++)
                                    for (uint32_t tx = tr[k].tx0; tx <= tr[k].tx1; tx++) {
                                        // the copy's count down to zero: slots hbase .. hbase + count - 1
                                        const uint32_t i = ((ty - wy0) * wt + tx - wx0) * kHistCopies + hc;
                                        const uint32_t slot = hbase[i] + atomicSub(&hist[i], 1u) - 1u;
                                        entries.store(slot, key[k], instance_word(g[k], tx, ty, t16[k], bl, ms));
                                    }
                    }
                } else {
                    // spread-out chunk: one global atomic per instance
#pragma unroll
                    for (uint32_t k = 0; k < kPer; k++)
                        if (small[k])
                            for (uint32_t ty = tr[k].ty0; ty <= tr[k].ty1; ty++)
                                for (uint32_t tx = tr[k].tx0; tx <= tr[k].tx1; tx++)
                                    bin_direct<kFill>(bin0 + ty * gw + tx, instance_word(g[k], tx, ty, t16[k], bl, ms),
                                                      key[k], tile_count, tile_cursor, offsets, entries);
                }
            }
            __syncthreads();
        }
    }
    if constexpr (!kFill) {
        if (tid == 0 && s_total) atomicAdd(&ctr->total_instances, s_total);
        if (tid == 0 && s_visible) atomicAdd(&ctr->visible_pairs, s_visible);
    }
}

// ---- depth range of each view (chunk centroids) ----
__global__ __launch_bounds__(kBatchBlock) void chunk_centroid_kernel(SceneArgs sc, float4* __restrict__ centroids)
{
    __shared__ float s_sum[kBatchBlock / 64u][4];
    const uint32_t g = blockIdx.x * kBatchBlock + threadIdx.x;
    float x = 0.0f, y = 0.0f, z = 0.0f, c = 0.0f;
    if (g < sc.n) {
        x = sc.means[3u * g + 0u];
        y = sc.means[3u * g + 1u];
        z = sc.means[3u * g + 2u];
        c = 1.0f;
    }
    for (int off = 32; off > 0; off >>= 1) {
        x += __shfl_xor(x, off, 64);
        y += __shfl_xor(y, off, 64);
        z += __shfl_xor(z, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    const uint32_t wave = threadIdx.x / 64u;
    if (lane_id() == 0u) {
        s_sum[wave][0] = x;
        s_sum[wave][1] = y;
        s_sum[wave][2] = z;
        s_sum[wave][3] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0u) {
        float sx = 0.0f, sy = 0.0f, sz = 0.0f, sc_ = 0.0f;
        for (uint32_t w = 0; w < kBatchBlock / 64u; w++) {
            sx += s_sum[w][0];
            sy += s_sum[w][1];
            sz += s_sum[w][2];
            sc_ += s_sum[w][3];
        }
        centroids[blockIdx.x] = make_float4(sx / sc_, sy / sc_, sz / sc_, 0.0f);
    }
}

constexpr uint32_t kRangeThreads = 1024;

// One workgroup per view: the depth-key bucketing of the per-tile sort. Keys are float bits
// of positive depths, so (key - kmin) >> shift over the view's depth range of the chunk
// centroids gives ordered buckets of equal relative depth width.
__global__ __launch_bounds__(kRangeThreads) void view_key_range_kernel(const ViewDesc* __restrict__ views,
                                                                        const float4* __restrict__ centroids,
                                                                        uint32_t n_chunks, uint2* __restrict__ view_keys)
{
    __shared__ float s_lo[kRangeThreads / 64u], s_hi[kRangeThreads / 64u];
    const ViewDesc& vd = views[blockIdx.x];
    const float4 r3 = vd.r3;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / 64u;
    float lo = 3.0e38f, hi = -3.0e38f;
    for (uint32_t c = tid; c < n_chunks; c += kRangeThreads) {
        const float4 m = centroids[c];
        const float z = r3.x * m.x + r3.y * m.y + r3.z * m.z + r3.w;
        if (z == z) {
            lo = fminf(lo, z);
            hi = fmaxf(hi, z);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off, 64));
        hi = fmaxf(hi, __shfl_xor(hi, off, 64));
    }
    if (lane == 0u) {
        s_lo[wave] = lo;
        s_hi[wave] = hi;
    }
    __syncthreads();
    if (tid == 0u) {
        float zmin = 3.0e38f, zmax = -3.0e38f;
        for (uint32_t w = 0; w < kRangeThreads / 64u; w++) {
            zmin = fminf(zmin, s_lo[w]);
            zmax = fmaxf(zmax, s_hi[w]);
        }
        const float lo_d = fmaxf(zmin, vd.cam.near_plane), hi_d = fminf(zmax, vd.cam.far_plane);
        const uint32_t kmin = f2u(lo_d), kmax = hi_d > lo_d ? f2u(hi_d) : kmin;
        const uint32_t range = kmax - kmin;
        const uint32_t bits = range ? 32u - (uint32_t)__clz(range) : 0u;
        // (range >> shift) < 2^kSortRangeLog2
        view_keys[blockIdx.x] = make_uint2(kmin, bits > kSortRangeLog2 ? bits - kSortRangeLog2 : 0u);
    }
}

// ---- order of the sort's workgroups ----
// Bins by descending list length class (bit length of the length, 2^15 entries and above in
// one class; order within a class is irrelevant), so the longest lists start first instead of
// extending the sort's tail. Written over the bin counts, dead once the scan has turned them
// into offsets. Two launches over 1024-bin blocks: class counts, then placement; a wave
// counts and places each distinct class of its bins once.
constexpr uint32_t kOrderThreads = 1024, kOrderClasses = 17;

__device__ __forceinline__ uint32_t order_class(const uint32_t* offsets, uint32_t b)
{
    const uint32_t len = offsets[b + 1u] - offsets[b];
    return min(len ? 32u - (uint32_t)__clz(len) : 0u, kOrderClasses - 1u);
}

template <bool kPlace>
__global__ __launch_bounds__(kOrderThreads) void bin_order_kernel(const uint32_t* __restrict__ offsets,
                                                                   uint32_t n_bins, uint32_t* __restrict__ order,
                                                                   BatchCounters* __restrict__ ctr)
{
    __shared__ uint32_t s_count[kOrderClasses], s_base[kOrderClasses];
    if (ctr->overflow) return;
    const uint32_t tid = threadIdx.x, lane = lane_id();
    const uint32_t b = blockIdx.x * kOrderThreads + tid;
    const uint32_t c = b < n_bins ? order_class(offsets, b) : kOrderClasses;
    if (tid < kOrderClasses) s_count[tid] = 0u;
    __syncthreads();
    // this block's bins per class
    uint32_t rank = 0;  // among this block's bins of the same class
    uint64_t todo = __ballot(b < n_bins);
    while (todo) {
        const int leader = __builtin_ctzll(todo);
        const uint32_t lc = (uint32_t)__shfl((int)c, leader, 64);
        const uint64_t same = __ballot(c == lc) & todo;
        todo &= ~same;
        uint32_t first = 0;
        if ((int)lane == leader) first = atomicAdd(&s_count[lc], (uint32_t)__popcll(same));
        first = (uint32_t)__shfl((int)first, leader, 64);
        if (c == lc) rank = first + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
    }
    __syncthreads();
    if constexpr (!kPlace) {
        if (tid < kOrderClasses && s_count[tid]) atomicAdd(&ctr->order_count[tid], s_count[tid]);
    } else {
        // the block's slots: after the longer classes and the blocks that reserved before it
        if (tid < kOrderClasses) {
            uint32_t base = 0;
            for (uint32_t k = tid + 1u; k < kOrderClasses; k++) base += ctr->order_count[k];
            s_base[tid] = base + (s_count[tid] ? atomicAdd(&ctr->order_cursor[tid], s_count[tid]) : 0u);
        }
        __syncthreads();
        if (b < n_bins) order[s_base[c] + rank] = b;
    }
}

// ---- per-tile counting sort by depth ----
// One workgroup per tile: the tile's instances are written bucket by bucket, nearest first.
// Bucket 0 takes keys below the view's kmin; bucket b >= 1 takes (key - kmin) >> shift = b - 1
// (clamped to the last bucket). keys_out[begin + 256 m] holds the floor of the bucket at
// sorted position 256 m (0, or kmin + (b - 1) << shift): a lower bound of the true keys from
// that position on, read where the rasteriser's batches of 256 start; other positions of
// keys_out are not written.

__global__ __launch_bounds__(kBatchBlock) void tile_sort_kernel(uint32_t n_views, const uint32_t* __restrict__ order,
                                                                const uint32_t* __restrict__ tile_prefix,
                                                                const uint2* __restrict__ view_keys,
                                                                const uint32_t* __restrict__ offsets,
                                                                const EntryArrays entries_in,
                                                                uint32_t* __restrict__ keys_out,
                                                                uint32_t* __restrict__ inst_out,
                                                                const BatchCounters* __restrict__ ctr)
{
    __shared__ uint32_t hist[kSortBins];
    __shared__ uint32_t s_wave[kBatchBlock / 64u];
    __shared__ uint16_t s_stage[kSortStage];
    if (ctr->overflow) return;
    const uint32_t gt = order[blockIdx.x];
    const uint32_t begin = offsets[gt], end = offsets[gt + 1u];
    if (end <= begin) return;
    uint32_t lo = 0, hi = n_views - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (tile_prefix[mid] <= gt) lo = mid;
        else hi = mid - 1u;
    }
    const uint2 vk = view_keys[lo];
    const uint32_t n = end - begin, tid = threadIdx.x, lane = lane_id(), wave = tid / 64u;
    for (uint32_t i = tid; i < kSortBins; i += kBatchBlock) hist[i] = 0u;
    __syncthreads();
    // The first kSortRegKeys * 256 entries of a bin (all of them in most bins) stay in registers
    // between the histogram and the scatter (two 16-bit buckets per register, so more
    // workgroups fit a CU); the buckets of the next kSortStage entries stay in LDS. Only the
    // buckets of entries beyond both are read twice.
    constexpr uint32_t kRegEntries = kSortRegKeys * kBatchBlock;
    const uint32_t n_regs = min(n, kRegEntries);
    uint32_t rb[kSortRegKeys / 2u], rg[kSortRegKeys];
#pragma unroll
    for (uint32_t u = 0; u < kSortRegKeys; u++) {
        const uint32_t idx = tid + u * kBatchBlock;
        const uint32_t b = idx < n_regs ? entries_in.bucket[begin + idx] : 0u;
        rb[u / 2u] = (u & 1u) ? rb[u / 2u] | (b << 16) : b;
        rg[u] = idx < n_regs ? entries_in.gaussian[begin + idx] : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < kSortRegKeys; u++)
        if (tid + u * kBatchBlock < n_regs) atomicAdd(&hist[(rb[u / 2u] >> (16u * (u & 1u))) & 0xffffu], 1u);
    // entries kRegEntries + i, kSortCountInFlight buckets in flight per lane (the registers'
    // entries are live here); i < kSortStage is uniform over each group of kSortCountInFlight
    // * 256 (kSortStage is a multiple of it)
    const uint32_t n_rest = n - n_regs;
    const uint32_t rbegin = begin + n_regs;
    {
        uint32_t i = tid;
        for (; i + (kSortCountInFlight - 1u) * kBatchBlock < n_rest; i += kSortCountInFlight * kBatchBlock) {
            uint32_t k[kSortCountInFlight];
#pragma unroll
            for (uint32_t u = 0; u < kSortCountInFlight; u++) k[u] = entries_in.bucket[rbegin + i + u * kBatchBlock];
            if (i < kSortStage)
#pragma unroll
                for (uint32_t u = 0; u < kSortCountInFlight; u++) s_stage[i + u * kBatchBlock] = (uint16_t)k[u];
#pragma unroll
            for (uint32_t u = 0; u < kSortCountInFlight; u++) atomicAdd(&hist[k[u]], 1u);
        }
        for (; i < n_rest; i += kBatchBlock) {
            const uint32_t k = entries_in.bucket[rbegin + i];
            if (i < kSortStage) s_stage[i] = (uint16_t)k;
            atomicAdd(&hist[k], 1u);
        }
    }
    __syncthreads();
    // exclusive scan of the buckets, kQ per thread
    constexpr uint32_t kQ = kSortBins / kBatchBlock;
    uint32_t v[kQ], local = 0;
#pragma unroll
    for (uint32_t q = 0; q < kQ; q++) {
        v[q] = hist[tid * kQ + q];
        local += v[q];
    }
    uint32_t incl = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, off, 64);
        if ((int)lane >= off) incl += y;
    }
    if (lane == 63u) s_wave[wave] = incl;
    __syncthreads();
    uint32_t run = incl - local;
    for (uint32_t w = 0; w < wave; w++) run += s_wave[w];
#pragma unroll
    for (uint32_t q = 0; q < kQ; q++) {
        hist[tid * kQ + q] = run;
        run += v[q];
    }
    __syncthreads();
    // The rasteriser reads the depth floor only where its batches of 256 start: the floor of
    // the bucket holding sorted position 256 m is the last bucket whose start is <= 256 m.
    for (uint32_t m = tid; m * kBatchBlock < n; m += kBatchBlock) {
        const uint32_t p = m * kBatchBlock;
        uint32_t b = 0, e = kSortBins - 1u;
        while (b < e) {
            const uint32_t mid = (b + e + 1u) >> 1;
            if (hist[mid] <= p) b = mid;
            else e = mid - 1u;
        }
        keys_out[begin + p] = bucket_floor(b, vk);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < kSortRegKeys; u++)
        if (tid + u * kBatchBlock < n_regs)
            inst_out[begin + atomicAdd(&hist[(rb[u / 2u] >> (16u * (u & 1u))) & 0xffffu], 1u)] = rg[u];
    uint32_t i = tid;
    for (; i + (kSortInFlight - 1u) * kBatchBlock < n_rest; i += kSortInFlight * kBatchBlock) {
        uint32_t k[kSortInFlight], g[kSortInFlight];
        if (i < kSortStage) {
#pragma unroll
            for (uint32_t u = 0; u < kSortInFlight; u++) {
                k[u] = s_stage[i + u * kBatchBlock];
                g[u] = entries_in.gaussian[rbegin + i + u * kBatchBlock];
            }
        } else {
#pragma unroll
            for (uint32_t u = 0; u < kSortInFlight; u++) {
                k[u] = entries_in.bucket[rbegin + i + u * kBatchBlock];
                g[u] = entries_in.gaussian[rbegin + i + u * kBatchBlock];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kSortInFlight; u++) inst_out[begin + atomicAdd(&hist[k[u]], 1u)] = g[u];
    }
    for (; i < n_rest; i += kBatchBlock) {
        const uint32_t k = i < kSortStage ? (uint32_t)s_stage[i] : (uint32_t)entries_in.bucket[rbegin + i];
        inst_out[begin + atomicAdd(&hist[k], 1u)] = entries_in.gaussian[rbegin + i];
    }
}

// Inclusive scan over the wave through DPP (VALU, no LDS round trip): row_shr 1, 2, 4, 8 scan
// each row of 16 lanes (lanes without a source add 0), row_bcast 15 adds the end of rows 0 and
// 2 to rows 1 and 3, row_bcast 31 the end of rows 0-1 to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_inclusive_scan_dpp(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// inclusive scan of one value per thread over the 256-thread workgroup; total: their sum
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t x, uint32_t* wave_sums, uint32_t& total)
{
    const uint32_t lane = lane_id(), wave = threadIdx.x / 64u;
    x = wave_inclusive_scan_dpp(x);
    if (lane == 63u) wave_sums[wave] = x;
    __syncthreads();
    total = 0u;
#pragma unroll
    for (uint32_t w = 0; w < kBatchBlock / 64u; w++) {
        const uint32_t sw = wave_sums[w];
        if (w < wave) x += sw;
        total += sw;
    }
    return x;
}

// A workgroup's pixels (16 x kRows: a 16x16 tile, or half of one) are sub-tiles of 4x4; a wave
// holds rows 4w..4w+3, i.e. sub-tiles (0..3, w). Their maxima bound the framebuffer words an
// instance can still lower.
constexpr uint32_t kSub = 4;
// One queued fragment (instance j << 8 | pixel l): alpha once, then every sample word of the
// pixel the fragment still lies in front of gets its hash test and, on acceptance, the LDS
// atomicMin. The words of the G samples of the pass are read in groups of four or eight.
template <uint32_t kLayers, uint32_t kPix>
__device__ __forceinline__ void tile_fragment(uint32_t entry, unsigned long long (*layer)[kPix],
                                              const uint64_t* s_frag, const float* s_msx, const float* s_msy,
                                              const float* s_o, const float* s_ca, const float* s_cb,
                                              const float* s_cc, uint32_t ox, uint32_t oy, uint32_t G,
                                              uint32_t fs_group)
{
    const uint32_t j = entry >> 8, l = entry & 0xffu;
    const uint32_t px = ox + (l & 0xfu), py = oy + (l >> 4);
    const ConicExp2 ce = { s_ca[j], s_cb[j], s_cc[j] };
    const float alpha = fragment_alpha(ce, s_o[j], (float)px - s_msx[j], (float)py - s_msy[j]);
    // a NaN alpha accepts no sample (alpha_threshold is 0), so it returns here as well
    if (!(alpha >= kMinAlpha)) return;
    const uint64_t fj = s_frag[j];
    const uint32_t fd = (uint32_t)(fj >> 32);
    // (h >> 9) < T  <=>  h <= 512 T - 1 for 1 <= T <= 2^23 (alpha <= 1; T = 2^23 wraps to all
    // ones, which accepts every h as (h >> 9) < 2^23 does)
    const uint32_t lim = (min(alpha_threshold(alpha), 1u << 23) << 9) - 1u;
    // px, py < 2^16, so the 24-bit multiplies give the full products of rand_uniform's seed
    const uint32_t base = (uint32_t)fj ^ __umul24(px, 9277u) ^ __umul24(py, 2699u);
    // Only the depth halves of the words are read: a sample is tested if its word's depth is not
    // in front of the fragment (equal depths are resolved by the atomicMin), and words only
    // decrease, so a stale read never drops a winner. The comparison is the branch condition
    // itself (a lane mask), and a sample no lane tests is skipped as a whole; each sample's
    // frame term is a wave-uniform constant.
    // (eight words in flight for 16 samples measured faster than four; four for 8)
    constexpr uint32_t kGroup = kLayers < 4u ? kLayers : kLayers < 16u ? 4u : 8u;
#pragma unroll
    for (uint32_t s0 = 0; s0 < kLayers; s0 += kGroup) {
        if (s0 >= G) break;
        uint32_t w[kGroup];
#pragma unroll
        for (uint32_t i = 0; i < kGroup; i++)
            w[i] = __hip_atomic_load(reinterpret_cast<uint32_t*>(&layer[s0 + i][l]) + 1, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
        for (uint32_t i = 0; i < kGroup; i++) {
            const uint32_t s = s0 + i;
            if (s >= G) break;
            if (fd <= w[i]) {
                const uint32_t h = pcg_hash(base ^ (fs_group + s * 1973u));
                if (h <= lim) atomicMin(&layer[s][l], (unsigned long long)fj);
            }
        }
    }
}

// kLayers: samples held in LDS per pass over the instances (1 for views of one sample, whose
// LDS then holds one framebuffer word per pixel). kRows: rows of the workgroup's pixels, 16 for
// a whole 16x16 tile, 8 for half of one (two workgroups per tile, the same LDS for twice the
// samples: a view of 16 samples then walks each instance's pixels and evaluates each fragment's
// alpha once instead of once per pass of 8). Only views with lo <= spp <= hi are rasterised
// here; the others return at once.
template <uint32_t kLayers, uint32_t kRows>
__global__ __launch_bounds__(kBatchBlock, kLayers == 1u ? 7 : 5) void batch_tile_raster_kernel(SceneArgs sc,
                                                                        const float4* __restrict__ cov3d,
                                                                        const ViewDesc* __restrict__ views,
                                                                        uint32_t n_views,
                                                                        const uint32_t* __restrict__ raster_prefix,
                                                                        const uint32_t* __restrict__ offsets,
                                                                        const uint32_t* __restrict__ instances,
                                                                        const uint32_t* __restrict__ keys,
                                                                        const BatchCounters* __restrict__ ctr,
                                                                        uint64_t* __restrict__ fb, uint32_t spp_lo,
                                                                        uint32_t spp_hi)
{
    static_assert(kRows == kTile || kRows * 2u == kTile, "a workgroup covers a tile or half of one");
    static_assert(kLayers * kTile * kRows <= 2048u, "16 KB of framebuffer words per workgroup");
    constexpr uint32_t kPix = kTile * kRows;
    constexpr uint32_t kSubtiles = (kTile / kSub) * (kRows / kSub);
    __shared__ unsigned long long layer[kLayers][kPix];
    __shared__ uint64_t s_frag[kBatchBlock];
    __shared__ uint32_t s_sub[kSubtiles];  // depth bits of each 4x4 sub-tile's largest word
    // depth bits of each pixel's largest word (the 17 padding words are read, and ignored, for
    // the missing pixels of a 2x2 quad at the last row and column)
    __shared__ uint32_t s_pmax[kPix + 17u];
    __shared__ uint16_t s_queue[kBatchBlock / 64u][320];  // per-wave (instance << 8 | pixel) queue
    __shared__ float s_msx[kBatchBlock], s_msy[kBatchBlock], s_o[kBatchBlock];
    __shared__ float s_thr[kBatchBlock];  // exponent below which a fragment's alpha is < 1/255
    __shared__ float s_ca[kBatchBlock], s_cb[kBatchBlock], s_cc[kBatchBlock];
    __shared__ uint32_t s_rect[kBatchBlock], s_end[kBatchBlock];
    __shared__ uint32_t s_wave[kBatchBlock / 64u];
    __shared__ uint16_t s_start[kBatchBlock];  // the instance holding each lane's first quad
    // the batch's instances that reach this tile, in s_end: the projection reads them before the
    // scan's barrier, and s_end is written after it
    uint32_t* const s_g = s_end;
    __shared__ uint32_t s_mine[kBatchBlock / 64u];   // their count per wave
    if (ctr->overflow) return;
    // Workgroups are handed to the eight XCDs round-robin (workgroup i to XCD i % 8), each XCD
    // with its own L2. Within every 8 x kPer workgroups, the kPer that share an XCD take kPer
    // consecutive slots (a 32-pixel bin's four tiles, or their eight halves, as views start at
    // multiples of 4 tiles), so a bin's list and its Gaussians are fetched into one L2: wg is the
    // workgroup in view order, blk its tile and half the half of the tile.
    constexpr uint32_t kPer = kRows == kTile ? 4u : 8u, kPerLog = kRows == kTile ? 2u : 3u;
    const uint32_t j = blockIdx.x & (8u * kPer - 1u);
    const uint32_t wg = (blockIdx.x & ~(8u * kPer - 1u)) | ((j & 7u) << kPerLog) | (j >> 3);
    const uint32_t blk = kRows == kTile ? wg : wg >> 1, half = kRows == kTile ? 0u : wg & 1u;
    if (blk >= raster_prefix[n_views]) return;  // the grid is a multiple of 8 x kPer
    // tile -> view (last v with raster_prefix[v] <= blk) -> (bin lt of the view, 16x16 tile q
    // of the bin's 2^b x 2^b)
    uint32_t lo = 0, hi = n_views - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (raster_prefix[mid] <= blk) lo = mid;
        else hi = mid - 1u;
    }
    const ViewDesc& vd = views[lo];
    if (vd.spp < spp_lo || vd.spp > spp_hi) return;  // a view of another launch
    const uint32_t bl = vd.bin_log, r = blk - raster_prefix[lo];
    const uint32_t lt = r >> (2u * bl), q = r & ((1u << (2u * bl)) - 1u);
    const uint32_t gt = vd.tile_begin + lt;
    const uint32_t gw = vd.grid_w;
    const uint32_t ox = ((lt % gw) << (kRasterLog2 + bl)) + (q & ((1u << bl) - 1u)) * kTile;
    const uint32_t oy = ((lt / gw) << (kRasterLog2 + bl)) + (q >> bl) * kTile + half * kRows;
    const uint32_t W = vd.cam.width_i, H = vd.cam.height_i;
    if (ox >= W || oy >= H) return;  // a tile of an edge bin beyond the framebuffer, or padding
    const size_t P = (size_t)W * H;
    const uint32_t begin = offsets[gt], end = offsets[gt + 1u];
    const uint32_t tid = threadIdx.x, lane = lane_id();
    // whether pixel p (row-major, 16 per row) of the workgroup lies inside the view
    const auto inside = [&](uint32_t p) { return ox + p % kTile < W && oy + p / kTile < H; };
    uint64_t* fbv = fb + vd.fb_offset;
    // the tile's position in its bin, for the instances' stored sub-rectangles
    const uint32_t ms = subrect_shift(bl, sc.n), bm = (1u << bl) - 1u;
    const uint32_t qx = q & bm, qy = q >> bl;

    for (uint32_t s0 = 0; s0 < vd.spp; s0 += kLayers) {
        const uint32_t G = min(kLayers, vd.spp - s0);
        // pixels outside the view hold 0 so they never keep an instance alive
        for (uint32_t i = tid; i < G * kPix; i += kBatchBlock) layer[i / kPix][i % kPix] = inside(i % kPix) ? kEmpty : 0ull;
        const uint32_t fs_group = (vd.frame0 + s0) * 1973u;
        for (uint32_t b0 = begin; b0 < end; b0 += kBatchBlock) {
            const uint32_t nb = min(kBatchBlock, end - b0);
            // With sub-rectangles stored, the batch's instances whose sub-rectangle holds this
            // tile are compacted to the front, so the projection below runs on as few waves as
            // they fill: a bin's tiles share its list, and about half of a list's Gaussians miss
            // any one of its tiles. Without (bins of one tile), every instance is projected.
            uint32_t gi = 0;
            bool mine = false;
            uint64_t mine_mask = 0;
            if (ms < 32u) {
                if (tid < nb) {
                    const uint32_t w = instances[b0 + tid], sr = w >> ms;
                    gi = w & ((1u << ms) - 1u);
                    mine = qx >= (sr & bm) && qx <= ((sr >> bl) & bm) && qy >= ((sr >> (2u * bl)) & bm) &&
                           qy <= (sr >> (3u * bl));
                }
                mine_mask = __ballot(mine);
                if (lane == 0u) s_mine[tid / 64u] = (uint32_t)__popcll(mine_mask);
            }
            __syncthreads();
            // pixel and sub-tile maxima of the words' depths so far (the depth half of a word
            // orders it before its index half)
            if (tid < kPix) {  // whole waves
                uint32_t m = 0;
                for (uint32_t s = 0; s < G; s++) m = max(m, (uint32_t)(layer[s][tid] >> 32));
                s_pmax[tid] = m;
                m = max(m, (uint32_t)__shfl_xor((int)m, 1, 64));
                m = max(m, (uint32_t)__shfl_xor((int)m, 2, 64));
                m = max(m, (uint32_t)__shfl_xor((int)m, 16, 64));
                m = max(m, (uint32_t)__shfl_xor((int)m, 32, 64));
                if ((lane & 51u) == 0u) s_sub[(tid / 64u) * (kTile / kSub) + (lane & 15u) / kSub] = m;
            }
            uint32_t n_mine = nb;
            if (ms < 32u) {
                uint32_t slot = __builtin_amdgcn_mbcnt_hi((uint32_t)(mine_mask >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mine_mask, 0u));
                n_mine = 0;
                for (uint32_t w = 0; w < kBatchBlock / 64u; w++) {
                    const uint32_t c = s_mine[w];
                    if (w < tid / 64u) slot += c;
                    n_mine += c;
                }
                if (mine) s_g[slot] = gi;
            }
            __syncthreads();
            // bucket floors ascend: once the nearest remaining one lies behind every word of
            // the tile, no later instance can win a sample
            uint32_t tile_max = 0;
            for (uint32_t q = 0; q < kSubtiles; q++) tile_max = max(tile_max, s_sub[q]);
            if (keys[b0] > tile_max) break;
            uint32_t cnt = 0, rect = 0;
            if (tid < n_mine) {
                const uint32_t g = ms < 32u ? s_g[tid] : instances[b0 + tid];
                Projected p;
                if (g < sc.n && project_cached(sc.means, cov3d, g, vd.r1, vd.r2, vd.r3, vd.cam, p)) {
                    const Aabb b = pixel_aabb(p, vd.cam);
                    uint32_t cx0 = max(b.x0, ox), cx1 = min(b.x1, ox + kTile);
                    uint32_t cy0 = max(b.y0, oy), cy1 = min(b.y1, oy + kRows);
                    const uint64_t frag = fragment_key(p.depth, g);
                    if (cx1 > cx0 && cy1 > cy0) {
                        // an instance can only win in sub-tiles holding a word behind it (a word
                        // at its depth or farther): the walk covers the bounding box of those
                        // sub-tiles within its rectangle
                        const uint32_t fd = (uint32_t)(frag >> 32);
                        uint32_t lx0 = kTile, lx1 = 0, ly0 = kTile, ly1 = 0;
                        for (uint32_t sy = (cy0 - oy) / kSub; sy <= (cy1 - 1u - oy) / kSub; sy++)
                            for (uint32_t sx = (cx0 - ox) / kSub; sx <= (cx1 - 1u - ox) / kSub; sx++)
                                if (fd <= s_sub[sy * (kTile / kSub) + sx]) {
                                    lx0 = min(lx0, sx * kSub);
                                    lx1 = max(lx1, sx * kSub + kSub);
                                    ly0 = min(ly0, sy * kSub);
                                    ly1 = max(ly1, sy * kSub + kSub);
                                }
                        if (lx0 < lx1) {
                            cx0 = max(cx0, ox + lx0);
                            cx1 = min(cx1, ox + lx1);
                            cy0 = max(cy0, oy + ly0);
                            cy1 = min(cy1, oy + ly1);
                            // the walk visits 2x2 pixel quads: qw per row of quads, qh rows
                            // (rw, rh <= 16); ceil(1024 / qw) turns the walk's f / qw (f < 64)
                            // into a multiply and shift
                            const uint32_t rw = cx1 - cx0, rh = cy1 - cy0, qw = (rw + 1u) >> 1;
                            rect = (cx0 - ox) | ((cy0 - oy) << 4) | ((rw - 1u) << 8) | ((rh - 1u) << 12) |
                                   (((1024u + qw - 1u) / qw) << 16);
                            cnt = qw * ((rh + 1u) >> 1);
                        }
                    }
                    const ConicExp2 ce = conic_exp2(p.conic_x, p.conic_y, p.conic_z);
                    s_frag[tid] = frag;
                    s_msx[tid] = p.mean_x - 0.5f;
                    s_msy[tid] = p.mean_y - 0.5f;
                    s_o[tid] = p.opacity;
                    // log2(kMinAlpha / opacity) = -log2(255 opacity) (power_threshold is
                    // log(255 opacity)) less a margin of 0.01 that dwarfs the error of this
                    // estimate, so no fragment passing the exact alpha test is ever below it
                    s_thr[tid] = -p.power_threshold * kLog2e - 0.01f;
                    s_ca[tid] = ce.a;
                    s_cb[tid] = ce.b;
                    s_cc[tid] = ce.c;
                }
            }
            s_rect[tid] = rect;
            // pixel quads of the batch's live instances, flattened over the workgroup's lanes
            uint32_t total;
            const uint32_t incl = block_inclusive_scan(cnt, s_wave, total);
            // Each lane walks a contiguous range of per = ceil(total / 256) of the batch's 2x2
            // pixel quads (an instance's rectangle advances without division; the quad's pixels
            // beyond an odd width or height are skipped). Fragments that some sample word of
            // their pixel still lies behind go to the wave's LDS queue; alpha and the samples are
            // evaluated on full queues of 64, so no lane idles on a rejected fragment. Each
            // instance names itself to the lanes whose range starts in its quads
            // [incl - cnt, incl) (t per < incl <= total <= 256 per keeps t below 256).
            const uint32_t per = (total + kBatchBlock - 1u) / kBatchBlock;
            if (cnt)
                for (uint32_t t = (incl - cnt + per - 1u) / per; t * per < incl; t++) s_start[t] = (uint16_t)tid;
            s_end[tid] = incl;
            __syncthreads();
            uint32_t k = tid * per;
            const uint32_t kend = min(k + per, total);
            uint32_t j = 0, jend = 0, rect_j = 0, qw = 1, q = 0, r = 0, fdepth = 0;
            float msx = 0.0f, msy = 0.0f, thr = 0.0f;
            ConicExp2 ce = { 0.0f, 0.0f, 0.0f };
            if (k < kend) {
                j = s_start[tid];
                jend = s_end[j];
                const uint32_t f = k - (j ? s_end[j - 1u] : 0u);
                rect_j = s_rect[j];
                qw = (((rect_j >> 8) & 0xfu) + 2u) >> 1;
                q = (f * (rect_j >> 16)) >> 10;  // f / qw for f < 64, qw <= 8
                r = f - q * qw;
                fdepth = (uint32_t)(s_frag[j] >> 32);
                msx = s_msx[j];
                msy = s_msy[j];
                thr = s_thr[j];
                ce = { s_ca[j], s_cb[j], s_cc[j] };
            }
            uint16_t* queue = s_queue[tid / 64u];
            uint32_t count = 0;  // wave-uniform
            for (uint32_t it = 0; it < per; it++, k++) {
                // lanes past their range compute on stale (in-bounds) state and queue nothing
                const bool valid = k < kend;
                if (valid && k == jend) {
                    while (s_end[j] <= k) j++;
                    jend = s_end[j];
                    rect_j = s_rect[j];
                    qw = (((rect_j >> 8) & 0xfu) + 2u) >> 1;
                    q = 0;
                    r = 0;
                    fdepth = (uint32_t)(s_frag[j] >> 32);
                    msx = s_msx[j];
                    msy = s_msy[j];
                    thr = s_thr[j];
                    ce = { s_ca[j], s_cb[j], s_cc[j] };
                }
                // a lane whose range ended may sit one quad row below its rectangle: the row is
                // taken modulo kRows so its reads stay inside s_pmax (valid lanes have ly < kRows)
                const uint32_t lx = (rect_j & 0xfu) + 2u * r, ly = (((rect_j >> 4) & 0xfu) + 2u * q) & (kRows - 1u);
                const uint32_t l = ly * kTile + lx;
                const bool two_x = 2u * r + 1u < ((rect_j >> 8) & 0xfu) + 1u;
                const bool two_y = 2u * q + 1u < ((rect_j >> 12) & 0xfu) + 1u;
                if (valid && ++r == qw) {
                    r = 0;
                    q++;
                }
                // some sample word of the pixel was behind the fragment at the batch start,
                // and the fragment lies inside the instance's alpha >= 1/255 ellipse (the
                // exponents are the ones the alpha evaluation computes; pixel coordinates below
                // 2^24 are exact floats, so p + 1.0f is (float)(p + 1))
                const float fx = (float)(ox + lx), fy = (float)(oy + ly);
                const float dx0 = fx - msx, dx1 = (fx + 1.0f) - msx;
                const float dy0 = fy - msy, dy1 = (fy + 1.0f) - msy;
                // Candidate masks: every test is one comparison whose lane mask (ballot) is taken
                // on its own and the masks are combined with scalar ANDs (a ballot of a combined
                // condition would be materialised per lane and compared again). Every lane of the
                // wave is active here.
                const uint64_t mv = __builtin_amdgcn_ballot_w64(valid);
                const uint64_t mx = __builtin_amdgcn_ballot_w64(two_x), my = __builtin_amdgcn_ballot_w64(two_y);
                const uint64_t m00 = mv & __builtin_amdgcn_ballot_w64(fdepth <= s_pmax[l]) &
                                     __builtin_amdgcn_ballot_w64(!(fragment_exponent(ce, dx0, dy0) < thr));
                const uint64_t m01 = mv & mx & __builtin_amdgcn_ballot_w64(fdepth <= s_pmax[l + 1u]) &
                                     __builtin_amdgcn_ballot_w64(!(fragment_exponent(ce, dx1, dy0) < thr));
                const uint64_t m10 = mv & my & __builtin_amdgcn_ballot_w64(fdepth <= s_pmax[l + kTile]) &
                                     __builtin_amdgcn_ballot_w64(!(fragment_exponent(ce, dx0, dy1) < thr));
                const uint64_t m11 = mv & mx & my & __builtin_amdgcn_ballot_w64(fdepth <= s_pmax[l + kTile + 1u]) &
                                     __builtin_amdgcn_ballot_w64(!(fragment_exponent(ce, dx1, dy1) < thr));
                if (!(m00 | m01 | m10 | m11)) continue;
                // queue slots: the candidates of lower lanes first (mbcnt counts the set bits of
                // a mask below this lane), the quad's four pixels one after another
                const auto below = [](uint64_t m) {
                    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                };
                const uint32_t entry = (j << 8) | l;
                const uint32_t n00 = (uint32_t)__popcll(m00), n01 = (uint32_t)__popcll(m01);
                const uint32_t n10 = (uint32_t)__popcll(m10), n11 = (uint32_t)__popcll(m11);
                if (__builtin_amdgcn_inverse_ballot_w64(m00)) queue[count + below(m00)] = (uint16_t)entry;
                if (__builtin_amdgcn_inverse_ballot_w64(m01)) queue[count + n00 + below(m01)] = (uint16_t)(entry + 1u);
                if (__builtin_amdgcn_inverse_ballot_w64(m10))
                    queue[count + n00 + n01 + below(m10)] = (uint16_t)(entry + kTile);
                if (__builtin_amdgcn_inverse_ballot_w64(m11))
                    queue[count + n00 + n01 + n10 + below(m11)] = (uint16_t)(entry + kTile + 1u);
                count += n00 + n01 + n10 + n11;
                // count < 64 before this step and at most 256 entries were added: full queues
                // of 64 are evaluated, then fewer than 64 entries move to the front
                if (count >= 64u) {
                    __builtin_amdgcn_wave_barrier();
                    uint32_t head = 0;
                    do {
                        tile_fragment<kLayers, kPix>(queue[head + lane], layer, s_frag, s_msx, s_msy, s_o, s_ca, s_cb,
                                                     s_cc, ox, oy, G, fs_group);
                        head += 64u;
                    } while (count - head >= 64u);
                    count -= head;
                    __builtin_amdgcn_wave_barrier();
                    const uint32_t rest = lane < count ? queue[head + lane] : 0u;
                    __builtin_amdgcn_wave_barrier();
                    if (lane < count) queue[lane] = rest;
                    __builtin_amdgcn_wave_barrier();
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (lane < count)
                tile_fragment<kLayers, kPix>(queue[lane], layer, s_frag, s_msx, s_msy, s_o, s_ca, s_cb, s_cc, ox, oy, G,
                                       fs_group);
        }
        __syncthreads();
        for (uint32_t i = tid; i < G * kPix; i += kBatchBlock) {
            const uint32_t p = i % kPix;
            if (inside(p))
                fbv[(size_t)(s0 + i / kPix) * P + (size_t)(oy + p / kTile) * W + ox + p % kTile] = layer[i / kPix][p];
        }
        __syncthreads();
    }
}

__global__ void single_view_table_kernel(ViewDesc d, uint32_t views_bytes, SingleViewPrefixes prefixes,
                                         const float* __restrict__ w2c, const float* __restrict__ cam_position,
                                         char* __restrict__ table)
{
    if (threadIdx.x != 0) return;
    d.r1 = make_float4(w2c[0], w2c[1], w2c[2], w2c[3]);
    d.r2 = make_float4(w2c[4], w2c[5], w2c[6], w2c[7]);
    d.r3 = make_float4(w2c[8], w2c[9], w2c[10], w2c[11]);
    d.cam_x = cam_position[0];
    d.cam_y = cam_position[1];
    d.cam_z = cam_position[2];
    *reinterpret_cast<ViewDesc*>(table) = d;
    uint32_t* prefix = reinterpret_cast<uint32_t*>(table + views_bytes);
    for (uint32_t i = 0; i < 6u; i++) prefix[i] = prefixes.w[i];
}

hipError_t launch_single_view_table(const ViewDesc& desc, size_t views_bytes, const SingleViewPrefixes& prefixes,
                                    const float* w2c, const float* cam_position, void* table, hipStream_t stream)
{
    hipLaunchKernelGGL(single_view_table_kernel, dim3(1), dim3(64), 0, stream, desc, (uint32_t)views_bytes, prefixes,
                       w2c, cam_position, static_cast<char*>(table));
    return hipGetLastError();
}

hipError_t launch_batch_tiles(const SceneArgs& sc, const ViewDesc* views, uint32_t n_views, const uint32_t* tile_prefix,
                              const uint32_t* raster_prefix, const TileBuffers& tb, uint32_t layers,
                              BatchCounters* counters, uint64_t* fb, hipEvent_t raster_event, hipStream_t stream)
{
    if (n_views == 0 || tb.n_tiles == 0) return hipSuccess;
    const uint32_t n_chunks = (sc.n + kBatchBlock - 1u) / kBatchBlock;
    const uint64_t n_items64 = (uint64_t)n_chunks * n_views;
    if (n_items64 >= 0xffffffffull) return hipErrorInvalidValue;
    int device = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return e;
    // each binning pass: as many workgroups as fit the device at once (chunks are dealt
    // round-robin, so a workgroup that waits for a slot would lengthen the tail)
    uint64_t grids[2];
    for (int f = 0; f < 2; f++) {
        const uint32_t n_bin_chunks = (sc.n + (f ? kBinChunk<true> : kBinChunk<false>) - 1u) /
                                      (f ? kBinChunk<true> : kBinChunk<false>);
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, f ? reinterpret_cast<const void*>(batch_bin_kernel<true>)
                       : reinterpret_cast<const void*>(batch_bin_kernel<false>),
            kBatchBlock, 0);
        if (e != hipSuccess) return e;
        grids[f] = (uint64_t)cus * (per_cu < 1 ? 1 : per_cu);
        if (grids[f] > n_bin_chunks) grids[f] = n_bin_chunks;
        if (grids[f] < 1) grids[f] = 1;
    }
    // the scene's own cache of the view-independent terms, or one computed here per render
    const float4* cov3d = sc.cov3d ? sc.cov3d : tb.cov3d;
    if (sc.n > 0) {
        hipLaunchKernelGGL(chunk_centroid_kernel, dim3(n_chunks), dim3(kBatchBlock), 0, stream, sc, tb.centroids);
        hipLaunchKernelGGL(view_key_range_kernel, dim3(n_views), dim3(kRangeThreads), 0, stream, views, tb.centroids,
                           n_chunks, tb.view_keys);
        if (!sc.cov3d) hipLaunchKernelGGL(cov3d_kernel, dim3(n_chunks), dim3(kBatchBlock), 0, stream, sc, tb.cov3d);
        hipLaunchKernelGGL((batch_bin_kernel<false>), dim3((unsigned)grids[0]), dim3(kBatchBlock), 0, stream, sc, cov3d,
                           views, n_views, tb.tile_count, tb.tile_cursor, tb.offsets,
                           tb.entries, tb.view_keys, PairRecords{ tb.pair_rect, tb.compact_records }, counters);
    }
    hipLaunchKernelGGL((exclusive_scan_kernel<CountInput>), dim3(1), dim3(kScanThreads), 0, stream,
                       CountInput{ tb.tile_count }, tb.offsets, nullptr, tb.n_tiles, tb.n_tiles, &counters->n_instances,
                       &counters->overflow, tb.capacity, nullptr);
    if (sc.n > 0) {
        hipLaunchKernelGGL((batch_bin_kernel<true>), dim3((unsigned)grids[1]), dim3(kBatchBlock), 0, stream, sc, cov3d,
                           views, n_views, tb.tile_count, tb.tile_cursor, tb.offsets,
                           tb.entries, tb.view_keys, PairRecords{ tb.pair_rect, tb.compact_records }, counters);
    }
    const dim3 order_grid((tb.n_tiles + kOrderThreads - 1u) / kOrderThreads);
    hipLaunchKernelGGL(bin_order_kernel<false>, order_grid, dim3(kOrderThreads), 0, stream, tb.offsets, tb.n_tiles,
                       tb.tile_count, counters);
    hipLaunchKernelGGL(bin_order_kernel<true>, order_grid, dim3(kOrderThreads), 0, stream, tb.offsets, tb.n_tiles,
                       tb.tile_count, counters);
    hipLaunchKernelGGL(tile_sort_kernel, dim3(tb.n_tiles), dim3(kBatchBlock), 0, stream, n_views, tb.tile_count,
                       tile_prefix, tb.view_keys, tb.offsets, tb.entries, tb.sorted_keys, tb.sorted_instances, counters);
    if (raster_event) {
        e = hipEventRecord(raster_event, stream);
        if (e != hipSuccess) return e;
    }
    // one workgroup per 16x16 tile, 4^b per bin of (16 << b) pixels (the caller keeps the total
    // below 2^31; it equals raster_prefix[n_views]), rounded up to a multiple of 32; two per tile
    // for the half-tile kernel, rounded up to a multiple of 64
    const dim3 raster_grid((tb.n_raster + 31u) & ~31u);
    const dim3 half_grid((2u * tb.n_raster + 63u) & ~63u);
    // Views of one sample take the one-layer kernel (least LDS per workgroup); the others take
    // 8 samples per pass over whole tiles or 16 per pass over half tiles: by default 8 up to 8
    // samples and 16 above, `layers` 8 or 16 for every view of several samples. Every launch
    // spans all views' tiles and returns at once on the others' (measured faster than launches
    // over class-ordered tiles, whose sequential tails cost more than the early returns).
    const uint32_t kAll = 0xffffffffu;
    const uint32_t mid_hi = layers == 8u ? kAll : layers == 16u ? 1u : 8u;
    const uint32_t many_lo = layers == 16u ? 2u : layers == 8u ? kAll : 9u;
    if (tb.spp_max >= 1u && tb.spp_min <= 1u)
        hipLaunchKernelGGL((batch_tile_raster_kernel<1, kTile>), raster_grid, dim3(kBatchBlock), 0, stream, sc, cov3d,
                           views, n_views, raster_prefix, tb.offsets, tb.sorted_instances, tb.sorted_keys, counters,
                           fb, 1u, 1u);
    if (tb.spp_max >= 2u && tb.spp_min <= mid_hi)
        hipLaunchKernelGGL((batch_tile_raster_kernel<8, kTile>), raster_grid, dim3(kBatchBlock), 0, stream, sc, cov3d,
                           views, n_views, raster_prefix, tb.offsets, tb.sorted_instances, tb.sorted_keys, counters,
                           fb, 2u, mid_hi);
    if (tb.spp_max >= many_lo)
        hipLaunchKernelGGL((batch_tile_raster_kernel<16, kTile / 2u>), half_grid, dim3(kBatchBlock), 0, stream, sc,
                           cov3d, views, n_views, raster_prefix, tb.offsets, tb.sorted_instances, tb.sorted_keys,
                           counters, fb, many_lo, kAll);
    return hipGetLastError();
}

}  // namespace sgs
