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
                                                                       uint32_t entry_frags)
{
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
        const ViewRun vr = view_run(vd, fb);
        const uint32_t g = chunk * kBatchBlock + tid;
        Projected p;
        bool active =
            g < sc.n && project_gaussian(sc.means, sc.scales, sc.rotations, sc.opacities, g, vd.r1, vd.r2, vd.r3,
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
                                   hipStream_t stream)
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
                       t.sequential_threshold, t.big_threshold, t.entry_fragments);
    return hipGetLastError();
}

}  // namespace sgs
