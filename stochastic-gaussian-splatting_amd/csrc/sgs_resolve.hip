// Sample and resolve kernels: shade the single winner of each framebuffer (sub)pixel with
// its view-dependent SH colour, box-filter the subpixels and merge with the history
// (RESET / STATIC weighted average / REPROJECT with Catmull-Rom history and YCoCg soft
// rejection). Restates resolve.cu:27-349 and eval_sh (gs_utils.h:197-243).
#include "sgs_internal.h"

namespace sgs {

constexpr uint32_t kBlockSample = 256;
constexpr uint32_t kBlockResolve = 256;

// spatio-temporal reuse constants (resolve.cu:16-23)
constexpr float kHistoryWeight = 0.7f;
constexpr float kGammaStatic = 4.0f;
constexpr float kGammaMotion = 1.0f;
constexpr float kVelocityThreshold = 4.0f;
constexpr float kMinExtentLuma = 0.01f;
constexpr float kMinExtentChroma = 0.005f;
constexpr float kBorderMargin = 2.0f;

#define SH_C0 0.28209479177387814f
#define SH_C1 0.48860251190291987f
#define SH_C2A 1.0925484305920792f
#define SH_C2B 0.94617469575755997f
#define SH_C2C 0.31539156525251999f
#define SH_C2D 0.54627421529603959f
#define SH_C3A 0.59004358992664352f
#define SH_C3B 1.7701307697799304f
#define SH_C3C 2.8906114426405538f
#define SH_C3D 0.45704579946446572f
#define SH_C3E 2.2852289973223288f
#define SH_C3F 1.865881662950577f
#define SH_C3G 1.1195289977703462f
#define SH_C3H 1.4453057213202769f

// real SH up to degree 3 (B = 1, 4, 9, 16 bases), colour = max(0.5 + sum, 0)
template <uint32_t B>
__device__ __forceinline__ float3 eval_sh(const float* __restrict__ c, float x, float y, float z)
{
    float out[3];
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        float col = 0.5f + SH_C0 * c[ch];
        if constexpr (B > 1) {
            col = col - SH_C1 * y * c[3 + ch] + SH_C1 * z * c[6 + ch] - SH_C1 * x * c[9 + ch];
            if constexpr (B > 4) {
                const float xx = x * x, yy = y * y, zz = z * z;
                const float xy = x * y, xz = x * z, yz = y * z;
                col = col + SH_C2A * xy * c[12 + ch] - SH_C2A * yz * c[15 + ch] + (SH_C2B * zz - SH_C2C) * c[18 + ch]
                      - SH_C2A * xz * c[21 + ch] + SH_C2D * (xx - yy) * c[24 + ch];
                if constexpr (B > 9) {
                    col = col + y * (SH_C3A * yy - SH_C3B * xx) * c[27 + ch] + SH_C3C * xy * z * c[30 + ch]
                          + y * (SH_C3D - SH_C3E * zz) * c[33 + ch] + z * (SH_C3F * zz - SH_C3G) * c[36 + ch]
                          + x * (SH_C3D - SH_C3E * zz) * c[39 + ch] + SH_C3H * z * (xx - yy) * c[42 + ch]
                          + x * (SH_C3B * yy - SH_C3A * xx) * c[45 + ch];
                }
            }
        }
        out[ch] = fmaxf(col, 0.0f);
    }
    return make_float3(out[0], out[1], out[2]);
}

// colour of Gaussian g seen from cam_position: normalize(mean - cam) (resolve.cu:72-75)
template <uint32_t B>
__device__ __forceinline__ float3 shade(const SceneArgs& sc, float cpx, float cpy, float cpz, uint32_t g)
{
    const float vx = sc.means[3u * g + 0u] - cpx;
    const float vy = sc.means[3u * g + 1u] - cpy;
    const float vz = sc.means[3u * g + 2u] - cpz;
    const float inv = 1.0f / sqrtf(vx * vx + vy * vy + vz * vz);
    float c[B * 3];
    const float* src = sc.sh + (size_t)g * (B * 3u);
    if constexpr ((B * 3) % 4 == 0) {
        // (N,B,3) rows of 12 or 48 floats: 16-byte aligned when the array is
#pragma unroll
        for (uint32_t i = 0; i < B * 3 / 4; i++) {
            const float4 v = reinterpret_cast<const float4*>(src)[i];
            c[4 * i + 0] = v.x;
            c[4 * i + 1] = v.y;
            c[4 * i + 2] = v.z;
            c[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < B * 3; i++) c[i] = src[i];
    }
    return eval_sh<B>(c, vx * inv, vy * inv, vz * inv);
}

__device__ __forceinline__ float reuse_alpha(const ConicExp2 ce, float o, float dx, float dy)
{
    const float a = fragment_alpha(ce, o, dx, dy);
    return a < kMinAlpha ? 0.0f : a;
}

template <uint32_t B, bool kReuse>
__global__ __launch_bounds__(kBlockSample) void sample_kernel(SceneArgs sc, const float4* __restrict__ w2c,
                                                              const float* __restrict__ cam_position, FbCamera cam,
                                                              const uint64_t* __restrict__ fb,
                                                              float4* __restrict__ samples,
                                                              float4* __restrict__ reuse_weights,
                                                              uint2* __restrict__ k_prim, float4* __restrict__ k_occ,
                                                              uint32_t n_pixels)
{
    const uint32_t pix = blockIdx.x * kBlockSample + threadIdx.x;
    if (pix >= n_pixels) return;
    const uint64_t data = fb[pix];
    const uint32_t dk = (uint32_t)(data >> 32);
    const bool has = dk != 0xffffffffu;
    const uint32_t g = has ? (uint32_t)data : 0u;
    float4 smp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (has) {
        const float3 col = shade<B>(sc, cam_position[0], cam_position[1], cam_position[2], g);
        smp = make_float4(col.x, col.y, col.z, u2f(dk));
    }
    samples[pix] = smp;
    if constexpr (!kReuse) return;

    // project the winner again to evaluate it at the four neighbours (resolve.cu:81-117)
    Projected p;
    bool projected = false;
    if (has) projected = project_scene(sc, g, w2c[0], w2c[1], w2c[2], cam, p);
    float alpha = 1.0f;
    float na[4] = { 1.0f, 1.0f, 1.0f, 1.0f };
    float sdepth = __builtin_huge_valf();
    if (projected) {
        const float dx = (float)(pix % cam.width_i) - (p.mean_x - 0.5f);
        const float dy = (float)(pix / cam.width_i) - (p.mean_y - 0.5f);
        const ConicExp2 ce = conic_exp2(p.conic_x, p.conic_y, p.conic_z);
        alpha = fmaxf(reuse_alpha(ce, p.opacity, dx, dy), kMinAlpha);
        na[0] = reuse_alpha(ce, p.opacity, dx - 1.0f, dy);
        na[1] = reuse_alpha(ce, p.opacity, dx + 1.0f, dy);
        na[2] = reuse_alpha(ce, p.opacity, dx, dy - 1.0f);
        na[3] = reuse_alpha(ce, p.opacity, dx, dy + 1.0f);
        sdepth = smp.w;
    }
    float w[4];
#pragma unroll
    for (int c = 0; c < 4; c++) w[c] = (has && !projected) ? 0.0f : na[c] / alpha;

    // change in occlusion from the k-buffer (resolve.cu:119-141)
    bool unique = projected;
    uint2 prim[kKBuffer];
    float4 occ[kKBuffer];
#pragma unroll
    for (uint32_t k = 0; k < kKBuffer; k++) {
        const size_t e = (size_t)k * n_pixels + pix;
        prim[k] = k_prim[e];
        occ[k] = k_occ[e];
        if (prim[k].x == g) unique = false;
        if (u2f(prim[k].y) < sdepth) {
            w[0] = w[0] * occ[k].x;
            w[1] = w[1] * occ[k].y;
            w[2] = w[2] * occ[k].z;
            w[3] = w[3] * occ[k].w;
        }
    }
    if (unique) {
#pragma unroll
        for (uint32_t k = kKBuffer - 1; k > 0; k--) {
            const size_t e = (size_t)k * n_pixels + pix;
            k_prim[e] = prim[k - 1];
            k_occ[e] = occ[k - 1];
        }
        k_prim[pix] = make_uint2(g, dk);
        const float den = 1.0f - fminf(alpha, kMaxAlpha);
        k_occ[pix] = make_float4((1.0f - fminf(na[0], kMaxAlpha)) / den, (1.0f - fminf(na[1], kMaxAlpha)) / den,
                                 (1.0f - fminf(na[2], kMaxAlpha)) / den, (1.0f - fminf(na[3], kMaxAlpha)) / den);
    }
    reuse_weights[pix] = make_float4(w[0], w[1], w[2], w[3]);
}

template <uint32_t B>
static void sample_dispatch_b(bool reuse, unsigned grid, hipStream_t stream, const SceneArgs& sc, const float4* w2c,
                              const float* cp, const FbCamera& cam, const uint64_t* fb, float4* samples, float4* rw,
                              uint2* kp, float4* ko, uint32_t n)
{
    if (reuse)
        hipLaunchKernelGGL((sample_kernel<B, true>), dim3(grid), dim3(kBlockSample), 0, stream, sc, w2c, cp, cam, fb,
                           samples, rw, kp, ko, n);
    else
        hipLaunchKernelGGL((sample_kernel<B, false>), dim3(grid), dim3(kBlockSample), 0, stream, sc, w2c, cp, cam, fb,
                           samples, rw, kp, ko, n);
}

hipError_t launch_sample(const SceneArgs& sc, const float* w2c, const float* cam_position, const FbCamera& cam,
                         const uint64_t* fb, float4* samples, float4* reuse_weights, uint2* k_prim, float4* k_occ,
                         bool spatial_reuse, hipStream_t stream)
{
    const uint32_t n = cam.width_i * cam.height_i;
    if (n == 0) return hipSuccess;
    const unsigned grid = (n + kBlockSample - 1u) / kBlockSample;
    const float4* w = reinterpret_cast<const float4*>(w2c);
    switch (sc.sh_bases) {
        case 1: sample_dispatch_b<1>(spatial_reuse, grid, stream, sc, w, cam_position, cam, fb, samples, reuse_weights, k_prim, k_occ, n); break;
        case 4: sample_dispatch_b<4>(spatial_reuse, grid, stream, sc, w, cam_position, cam, fb, samples, reuse_weights, k_prim, k_occ, n); break;
        case 9: sample_dispatch_b<9>(spatial_reuse, grid, stream, sc, w, cam_position, cam, fb, samples, reuse_weights, k_prim, k_occ, n); break;
        default: sample_dispatch_b<16>(spatial_reuse, grid, stream, sc, w, cam_position, cam, fb, samples, reuse_weights, k_prim, k_occ, n); break;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// resolve (resolve.cu:146-349)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }

__device__ __forceinline__ float4 catmull_rom_weights(float t)
{
    const float t2 = t * t;
    return make_float4(t * (-0.5f + t * (1.0f - 0.5f * t)), 1.0f + t2 * (-2.5f + 1.5f * t),
                       t * (0.5f + t * (2.0f - 1.5f * t)), t2 * (-0.5f + 0.5f * t));
}

__device__ __forceinline__ float3 rgb_to_ycocg(float r, float g, float b)
{
    return make_float3(0.25f * r + 0.5f * g + 0.25f * b, 0.5f * r - 0.5f * b, -0.25f * r + 0.5f * g - 0.25f * b);
}

// running sums of the box filter over the subpixels of one output pixel
struct SpatialSums {
    float c0, c1, c2, d, a, w;  // colour, depth, alpha, weight mass (pre-normalisation)
    float dout;                 // depth sum of covered subpixels
    uint32_t valid;             // covered subpixels
};

__device__ __forceinline__ void add_subpixel(SpatialSums& s, const float4 sm, float c0, float c1, float c2, float dsum,
                                             float asum, float wsum)
{
    if (sm.w > 0.0f) {
        s.dout = s.dout + sm.w;
        s.valid++;
    }
    const float r = 1.0f / wsum;
    s.c0 = s.c0 + c0 * r;
    s.c1 = s.c1 + c1 * r;
    s.c2 = s.c2 + c2 * r;
    s.d = s.d + dsum * r;
    s.a = s.a + asum * r;
    s.w = s.w + wsum;
}

struct TemporalOut {
    float c0, c1, c2, d, a, w;
};

// STATIC: weighted cumulative moving average (resolve.cu:250-258)
__device__ __forceinline__ TemporalOut temporal_static(const SpatialSums& s, const HistoryPtrs& h, uint32_t pix)
{
    TemporalOut o;
    const float hm = h.weight_mass[pix];
    o.w = hm + s.w;
    const float hf = hm / o.w;
    o.c0 = lerpf(s.c0, h.image[3u * pix + 0u], hf);
    o.c1 = lerpf(s.c1, h.image[3u * pix + 1u], hf);
    o.c2 = lerpf(s.c2, h.image[3u * pix + 2u], hf);
    o.d = lerpf(s.d, h.depth[pix], hf);
    o.a = lerpf(s.a, h.alpha[pix], hf);
    return o;
}

// one output pixel from the sample buffer: box filter (+ spatial reuse) and temporal mode
template <bool kReuse>
__device__ __forceinline__ TemporalOut resolve_from_samples(const FbCamera& cam, uint32_t width, uint32_t height,
                                                            float focal_x, float focal_y, float center_x,
                                                            float center_y, const float4* __restrict__ samples,
                                                            const float4* __restrict__ rw, const HistoryPtrs& hist,
                                                            const float* __restrict__ m, int32_t mode, uint32_t pix)
{
    const uint32_t ss = (uint32_t)cam.ss;
    const uint32_t bx = (pix % width) * ss, by = (pix / width) * ss;
    const uint32_t FW = cam.width_i, FH = cam.height_i;
    const float nf = 1.0f / (float)(ss * ss);
    SpatialSums s = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0u };
    for (uint32_t oy = 0; oy < ss; oy++) {
        for (uint32_t ox = 0; ox < ss; ox++) {
            const uint32_t fx = bx + ox, fy = by + oy;
            const uint32_t fi = fy * FW + fx;
            const float4 sm = samples[fi];
            float c0 = sm.x, c1 = sm.y, c2 = sm.z;
            float dsum = sm.w;
            float asum = sm.w > 0.0f ? 1.0f : 0.0f;
            float wsum = 1.0f;
            if constexpr (kReuse) {
                // neighbours stored their weights towards this subpixel (resolve.cu:225-230)
                auto add = [&](uint32_t ni, float wk) {
                    const float4 ns = samples[ni];
                    c0 = c0 + wk * ns.x;
                    c1 = c1 + wk * ns.y;
                    c2 = c2 + wk * ns.z;
                    dsum = dsum + wk * ns.w;
                    asum = asum + (ns.w > 0.0f ? wk : 0.0f);
                    wsum = wsum + wk;
                };
                if (fx > 0u) add(fi - 1u, rw[fi - 1u].y);
                if (fx < FW - 1u) add(fi + 1u, rw[fi + 1u].x);
                if (fy > 0u) add(fi - FW, rw[fi - FW].w);
                if (fy < FH - 1u) add(fi + FW, rw[fi + FW].z);
            }
            add_subpixel(s, sm, c0, c1, c2, dsum, asum, wsum);
        }
    }
    s.c0 = s.c0 * nf;
    s.c1 = s.c1 * nf;
    s.c2 = s.c2 * nf;
    s.d = s.d * nf;
    s.a = s.a * nf;
    s.w = s.w * nf;

    TemporalOut o = { s.c0, s.c1, s.c2, s.d, s.a, s.w };
    if (mode == SGS_TEMPORAL_STATIC) {
        o = temporal_static(s, hist, pix);
    } else if (mode == SGS_TEMPORAL_REPROJECT) {
        const uint32_t x = pix % width, y = pix / width;
        const float z = s.valid > 0u ? s.dout / (float)s.valid : cam.far_plane;
        const float posx = (float)x + 0.5f, posy = (float)y + 0.5f;
        const float vpx = (posx - center_x) / focal_x * z;
        const float vpy = (posy - center_y) / focal_y * z;
        const float vpz = z;
        const float pz = m[8] * vpx + m[9] * vpy + m[10] * vpz + m[11];
        if (pz > 0.0f) {
            const float ppx = focal_x * (m[0] * vpx + m[1] * vpy + m[2] * vpz + m[3]) / pz + center_x;
            const float ppy = focal_y * (m[4] * vpx + m[5] * vpy + m[6] * vpz + m[7]) / pz + center_y;
            const bool inside = ppx >= kBorderMargin && ppx <= (float)width - kBorderMargin && ppy >= kBorderMargin &&
                                ppy <= (float)height - kBorderMargin;
            if (inside) {
                // Catmull-Rom history fetch, all 16 taps inside by the border margin
                const float tx = ppx - 0.5f, ty = ppy - 0.5f;
                const float tbx = floorf(tx), tby = floorf(ty);
                const float4 wx4 = catmull_rom_weights(tx - tbx), wy4 = catmull_rom_weights(ty - tby);
                const float wx[4] = { wx4.x, wx4.y, wx4.z, wx4.w };
                const float wy[4] = { wy4.x, wy4.y, wy4.z, wy4.w };
                const uint32_t hbx = (uint32_t)tbx - 1u, hby = (uint32_t)tby - 1u;
                float ch0 = 0.0f, ch1 = 0.0f, ch2 = 0.0f, dh = 0.0f, ah = 0.0f;
                for (uint32_t j = 0; j < 4; j++) {
                    for (uint32_t i = 0; i < 4; i++) {
                        const size_t hi = (size_t)(hby + j) * width + hbx + i;
                        const float wgt = wx[i] * wy[j];
                        ch0 = ch0 + wgt * hist.image[3 * hi + 0];
                        ch1 = ch1 + wgt * hist.image[3 * hi + 1];
                        ch2 = ch2 + wgt * hist.image[3 * hi + 2];
                        dh = dh + wgt * hist.depth[hi];
                        ah = ah + wgt * hist.alpha[hi];
                    }
                }
                // YCoCg moments of the current samples in a 3x3 window
                float mean[3] = { 0.0f, 0.0f, 0.0f }, msq[3] = { 0.0f, 0.0f, 0.0f };
                for (int oy = -1; oy <= 1; oy++) {
                    for (int ox = -1; ox <= 1; ox++) {
                        int nxi = (int)x + ox, nyi = (int)y + oy;
                        nxi = nxi < 0 ? 0 : (nxi > (int)width - 1 ? (int)width - 1 : nxi);
                        nyi = nyi < 0 ? 0 : (nyi > (int)height - 1 ? (int)height - 1 : nyi);
                        const uint32_t nx = (uint32_t)nxi, ny = (uint32_t)nyi;
                        float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
                        for (uint32_t sy = 0; sy < ss; sy++)
                            for (uint32_t sx = 0; sx < ss; sx++) {
                                const float4 sm = samples[(ny * ss + sy) * FW + nx * ss + sx];
                                n0 = n0 + sm.x;
                                n1 = n1 + sm.y;
                                n2 = n2 + sm.z;
                            }
                        const float3 ycc = rgb_to_ycocg(n0 * nf, n1 * nf, n2 * nf);
                        mean[0] = mean[0] + ycc.x;
                        mean[1] = mean[1] + ycc.y;
                        mean[2] = mean[2] + ycc.z;
                        msq[0] = msq[0] + ycc.x * ycc.x;
                        msq[1] = msq[1] + ycc.y * ycc.y;
                        msq[2] = msq[2] + ycc.z * ycc.z;
                    }
                }
                float sd[3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    mean[c] = mean[c] / 9.0f;
                    msq[c] = msq[c] / 9.0f;
                    sd[c] = sqrtf(fmaxf(msq[c] - mean[c] * mean[c], 0.0f));
                }
                // soft history rejection with velocity-adaptive half-extents
                const float dvx = posx - ppx, dvy = posy - ppy;
                const float velocity = sqrtf(dvx * dvx + dvy * dvy);
                const float gamma = lerpf(kGammaStatic, kGammaMotion, fminf(velocity / kVelocityThreshold, 1.0f));
                const float minext[3] = { kMinExtentLuma, kMinExtentChroma, kMinExtentChroma };
                const float3 ych = rgb_to_ycocg(ch0, ch1, ch2);
                const float yv[3] = { ych.x, ych.y, ych.z };
                float disc[3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float ext = fmaxf(gamma * sd[c], minext[c]);
                    disc[c] = fabsf(yv[c] - mean[c]) / ext;
                }
                const float dist = fmaxf(disc[0], fmaxf(disc[1], disc[2]));
                const float discount = dist > 1.0f ? kHistoryWeight / dist : kHistoryWeight;
                o.c0 = lerpf(s.c0, ch0, discount);
                o.c1 = lerpf(s.c1, ch1, discount);
                o.c2 = lerpf(s.c2, ch2, discount);
                o.d = lerpf(s.d, dh, discount);
                o.a = lerpf(s.a, ah, discount);
                o.w = s.w / (1.0f - discount);
            }
        }
    }
    return o;
}

__device__ __forceinline__ void store_output(const OutputPtrs& out, uint32_t pix, const TemporalOut& o)
{
    out.image[3u * pix + 0u] = o.c0;
    out.image[3u * pix + 1u] = o.c1;
    out.image[3u * pix + 2u] = o.c2;
    out.depth[pix] = o.d;
    out.alpha[pix] = o.a;
    out.weight_mass[pix] = o.w;
}

template <bool kReuse>
__global__ __launch_bounds__(kBlockResolve) void resolve_kernel(FbCamera cam, uint32_t width, uint32_t height,
                                                                float focal_x, float focal_y, float center_x,
                                                                float center_y, const float4* __restrict__ samples,
                                                                const float4* __restrict__ rw, HistoryPtrs hist,
                                                                const float* __restrict__ m, OutputPtrs out,
                                                                int32_t mode)
{
    const uint32_t pix = blockIdx.x * kBlockResolve + threadIdx.x;
    if (pix >= width * height) return;
    store_output(out, pix, resolve_from_samples<kReuse>(cam, width, height, focal_x, focal_y, center_x, center_y,
                                                        samples, rw, hist, m, mode, pix));
}

hipError_t launch_resolve(const FbCamera& cam, uint32_t width, uint32_t height, float focal_x, float focal_y,
                          float center_x, float center_y, const float4* samples, const float4* reuse_weights,
                          const HistoryPtrs& hist, const float* m, OutputPtrs out, bool spatial_reuse,
                          int32_t temporal_mode, hipStream_t stream)
{
    const uint32_t n = width * height;
    if (n == 0) return hipSuccess;
    const unsigned grid = (n + kBlockResolve - 1u) / kBlockResolve;
    if (spatial_reuse)
        hipLaunchKernelGGL((resolve_kernel<true>), dim3(grid), dim3(kBlockResolve), 0, stream, cam, width, height,
                           focal_x, focal_y, center_x, center_y, samples, reuse_weights, hist, m, out, temporal_mode);
    else
        hipLaunchKernelGGL((resolve_kernel<false>), dim3(grid), dim3(kBlockResolve), 0, stream, cam, width, height,
                           focal_x, focal_y, center_x, center_y, samples, reuse_weights, hist, m, out, temporal_mode);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// fused shade + resolve for the common case (no spatial reuse, RESET or STATIC): each
// output pixel shades its ss x ss winners directly, so the (P,4) sample buffer is never
// written. Same operations in the same order as sample_kernel + resolve_kernel.
// ------------------------------------------------------------------------------------------
// box filter of one output pixel shading its ss x ss framebuffer winners directly
// (identical operations to sample_kernel followed by resolve_from_samples without reuse)
template <uint32_t B>
__device__ __forceinline__ SpatialSums shade_box(const SceneArgs& sc, float cpx, float cpy, float cpz,
                                                 const FbCamera& cam, uint32_t width, const uint64_t* __restrict__ fb,
                                                 uint32_t pix)
{
    const uint32_t ss = (uint32_t)cam.ss;
    const uint32_t bx = (pix % width) * ss, by = (pix / width) * ss;
    const float nf = 1.0f / (float)(ss * ss);
    SpatialSums s = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0u };
    for (uint32_t oy = 0; oy < ss; oy++) {
        for (uint32_t ox = 0; ox < ss; ox++) {
            const uint64_t data = fb[(size_t)(by + oy) * cam.width_i + bx + ox];
            const uint32_t dk = (uint32_t)(data >> 32);
            float4 sm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (dk != 0xffffffffu) {
                const float3 col = shade<B>(sc, cpx, cpy, cpz, (uint32_t)data);
                sm = make_float4(col.x, col.y, col.z, u2f(dk));
            }
            add_subpixel(s, sm, sm.x, sm.y, sm.z, sm.w, sm.w > 0.0f ? 1.0f : 0.0f, 1.0f);
        }
    }
    s.c0 = s.c0 * nf;
    s.c1 = s.c1 * nf;
    s.c2 = s.c2 * nf;
    s.d = s.d * nf;
    s.a = s.a * nf;
    s.w = s.w * nf;
    return s;
}

template <uint32_t B>
__global__ __launch_bounds__(kBlockResolve) void shade_resolve_kernel(SceneArgs sc, const float* __restrict__ cam_position,
                                                                      FbCamera cam, uint32_t width, uint32_t height,
                                                                      const uint64_t* __restrict__ fb, HistoryPtrs hist,
                                                                      OutputPtrs out, int32_t mode)
{
    const uint32_t pix = blockIdx.x * kBlockResolve + threadIdx.x;
    if (pix >= width * height) return;
    const SpatialSums s = shade_box<B>(sc, cam_position[0], cam_position[1], cam_position[2], cam, width, fb, pix);
    TemporalOut o = { s.c0, s.c1, s.c2, s.d, s.a, s.w };
    if (mode == SGS_TEMPORAL_STATIC) o = temporal_static(s, hist, pix);
    store_output(out, pix, o);
}

hipError_t launch_shade_resolve(const SceneArgs& sc, const float* cam_position, const FbCamera& cam, uint32_t width,
                                uint32_t height, const uint64_t* fb, const HistoryPtrs& hist, OutputPtrs out,
                                int32_t temporal_mode, hipStream_t stream)
{
    const uint32_t n = width * height;
    if (n == 0) return hipSuccess;
    const unsigned grid = (n + kBlockResolve - 1u) / kBlockResolve;
    switch (sc.sh_bases) {
        case 1: hipLaunchKernelGGL((shade_resolve_kernel<1>), dim3(grid), dim3(kBlockResolve), 0, stream, sc, cam_position, cam, width, height, fb, hist, out, temporal_mode); break;
        case 4: hipLaunchKernelGGL((shade_resolve_kernel<4>), dim3(grid), dim3(kBlockResolve), 0, stream, sc, cam_position, cam, width, height, fb, hist, out, temporal_mode); break;
        case 9: hipLaunchKernelGGL((shade_resolve_kernel<9>), dim3(grid), dim3(kBlockResolve), 0, stream, sc, cam_position, cam, width, height, fb, hist, out, temporal_mode); break;
        default: hipLaunchKernelGGL((shade_resolve_kernel<16>), dim3(grid), dim3(kBlockResolve), 0, stream, sc, cam_position, cam, width, height, fb, hist, out, temporal_mode); break;
    }
    return hipGetLastError();
}

}  // namespace sgs

// ------------------------------------------------------------------------------------------
// batched resolve: every output pixel of every view folds its view's spp samples in order,
// sample 0 with the view's temporal mode and the rest STATIC, the operations of spp
// successive single-sample calls (the history of sample s is the output of sample s - 1).
// The colour of a winner depends only on the view and the Gaussian, and the pixels of a tile
// share most of their winners across pixels and samples, so each workgroup (one 16x16 tile of
// image pixels) keeps its distinct winners in an LDS hash table. Every framebuffer word is read
// once: the pixels take kChunk words each into registers and insert their winners, the
// workgroup's lanes shade the winners new in the table, then each pixel folds its kChunk words
// with the colours looked up in the table. A winner the table cannot take (kShadeProbes
// occupied slots) is shaded by its pixel, so the colours are those of shade() in every case.
// ------------------------------------------------------------------------------------------
namespace sgs {

constexpr uint32_t kShadeLog2 = 11;
constexpr uint32_t kShadeSlots = 1u << kShadeLog2;
constexpr uint32_t kShadeProbes = 64;
constexpr uint32_t kNoWinner = 0xffffffffu;  // empty slot (Gaussian indices are below 2^32 - 1)
constexpr uint32_t kChunk = 4;  // framebuffer words per pixel in registers between the phases

__device__ __forceinline__ uint32_t shade_slot(uint32_t g)
{
    return (g * 0x9E3779B1u) >> (32u - kShadeLog2);  // multiplicative (Fibonacci) hashing
}

template <uint32_t B>
__global__ __launch_bounds__(kBlockResolve) void batch_resolve_kernel(SceneArgs sc, const ViewDesc* __restrict__ views,
                                                                      uint32_t n_views,
                                                                      const uint32_t* __restrict__ block_prefix,
                                                                      const uint64_t* __restrict__ fb,
                                                                      const float4* __restrict__ samples,
                                                                      uint32_t level)
{
    __shared__ uint32_t s_key[kShadeSlots];
    __shared__ float s_red[kShadeSlots], s_green[kShadeSlots], s_blue[kShadeSlots];
    __shared__ uint16_t s_new[kShadeSlots];  // slots in the order their winners were inserted
    __shared__ uint32_t s_n_new;
    // view of this block: last v with block_prefix[v] <= blockIdx.x
    uint32_t lo = 0, hi = n_views - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (block_prefix[mid] <= blockIdx.x) lo = mid;
        else hi = mid - 1u;
    }
    const ViewDesc& vd = views[lo];
    // views whose history is another view's output of the batch resolve in a later launch;
    // views with spatial reuse are resolved sample by sample
    if (vd.level != level || vd.reuse) return;
    // 16x16 tile t of the view's image pixels, row-major over ceil(W/16) tiles per row
    const uint32_t t = blockIdx.x - block_prefix[lo], tw = (vd.width + kResolveTile - 1u) / kResolveTile;
    const uint32_t px = (t % tw) * kResolveTile + threadIdx.x % kResolveTile;
    const uint32_t py = (t / tw) * kResolveTile + threadIdx.x / kResolveTile;
    const bool active = px < vd.width && py < vd.height;
    const uint32_t pix = active ? py * vd.width + px : 0u;
    const size_t P = (size_t)vd.cam.width_i * vd.cam.height_i;
    const uint64_t* fbv = fb + vd.fb_offset;
    const uint32_t s0 = vd.mode == SGS_TEMPORAL_REPROJECT ? 1u : 0u;
    const uint32_t ss = (uint32_t)vd.cam.ss;
    const uint32_t bx = px * ss, by = py * ss;
    for (uint32_t i = threadIdx.x; i < kShadeSlots; i += kBlockResolve) s_key[i] = kNoWinner;
    if (threadIdx.x == 0) s_n_new = 0u;
    TemporalOut o = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
    if (s0 == 1u && active)
        o = resolve_from_samples<false>(vd.cam, vd.width, vd.height, vd.focal_x, vd.focal_y, vd.center_x, vd.center_y,
                                        samples + vd.samples_offset, nullptr, vd.hist, vd.v2p, vd.mode, pix);
    __syncthreads();
    // the pixel's framebuffer words in (sample, subpixel row, subpixel column) order, kChunk at a
    // time (the count is the same for every pixel of the view, so the barriers are uniform)
    const uint32_t n_words = (vd.spp - s0) * ss * ss;
    const float nf = 1.0f / (float)(ss * ss);
    uint32_t ls = s0, ly = 0, lx = 0;  // load cursor
    uint32_t fs = s0, fy = 0, fx = 0;  // fold cursor
    SpatialSums sp = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0u };
    uint32_t last = kNoWinner, cg = kNoWinner, n_done = 0;
    float3 cc = make_float3(0.0f, 0.0f, 0.0f);
    for (uint32_t c0 = 0; c0 < n_words; c0 += kChunk) {
        uint64_t w[kChunk];
#pragma unroll
        for (uint32_t u = 0; u < kChunk; u++) {
            w[u] = ~0ull;
            if (c0 + u < n_words) {
                if (active) w[u] = fbv[(size_t)ls * P + (size_t)(by + ly) * vd.cam.width_i + bx + lx];
                if (++lx == ss) {
                    lx = 0;
                    if (++ly == ss) {
                        ly = 0;
                        ls++;
                    }
                }
            }
        }
        // 1. the chunk's winners into the table (open addressing with linear probing; slots are
        // never freed, so a winner that found no slot never gets one later)
#pragma unroll
        for (uint32_t u = 0; u < kChunk; u++) {
            const uint32_t g = (uint32_t)w[u];
            if ((uint32_t)(w[u] >> 32) == 0xffffffffu || g == last) continue;
            last = g;
            uint32_t h = shade_slot(g);
            for (uint32_t k = 0; k < kShadeProbes; k++) {
                uint32_t old = s_key[h];
                if (old == kNoWinner) old = atomicCAS(&s_key[h], kNoWinner, g);
                if (old == kNoWinner) {
                    s_new[atomicAdd(&s_n_new, 1u)] = (uint16_t)h;
                    break;
                }
                if (old == g) break;
                h = (h + 1u) & (kShadeSlots - 1u);
            }
        }
        __syncthreads();
        // 2. the winners new in this chunk shaded once, the workgroup's lanes over the list
        const uint32_t n_new = s_n_new;
        for (uint32_t i = n_done + threadIdx.x; i < n_new; i += kBlockResolve) {
            const uint32_t h = s_new[i];
            const float3 c = shade<B>(sc, vd.cam_x, vd.cam_y, vd.cam_z, s_key[h]);
            s_red[h] = c.x;
            s_green[h] = c.y;
            s_blue[h] = c.z;
        }
        n_done = n_new;
        __syncthreads();
        if (!active) continue;
        // 3. the chunk's words folded in order; a sample's box filter completes at its last word
#pragma unroll
        for (uint32_t u = 0; u < kChunk; u++) {
            if (c0 + u >= n_words) break;
            const uint32_t dk = (uint32_t)(w[u] >> 32);
            float4 sm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (dk != 0xffffffffu) {
                const uint32_t g = (uint32_t)w[u];
                if (g != cg) {
                    uint32_t h = shade_slot(g), k = 0;
                    for (; k < kShadeProbes; k++) {
                        const uint32_t key = s_key[h];
                        if (key == g || key == kNoWinner) break;
                        h = (h + 1u) & (kShadeSlots - 1u);
                    }
                    if (k < kShadeProbes && s_key[h] == g) cc = make_float3(s_red[h], s_green[h], s_blue[h]);
                    else cc = shade<B>(sc, vd.cam_x, vd.cam_y, vd.cam_z, g);
                    cg = g;
                }
                sm = make_float4(cc.x, cc.y, cc.z, u2f(dk));
            }
            add_subpixel(sp, sm, sm.x, sm.y, sm.z, sm.w, sm.w > 0.0f ? 1.0f : 0.0f, 1.0f);
            if (++fx < ss) continue;
            fx = 0;
            if (++fy < ss) continue;
            fy = 0;
            sp.c0 = sp.c0 * nf;
            sp.c1 = sp.c1 * nf;
            sp.c2 = sp.c2 * nf;
            sp.d = sp.d * nf;
            sp.a = sp.a * nf;
            sp.w = sp.w * nf;
            if (fs == 0) {
                o = { sp.c0, sp.c1, sp.c2, sp.d, sp.a, sp.w };
                if (vd.mode == SGS_TEMPORAL_STATIC) o = temporal_static(sp, vd.hist, pix);
            } else {
                // STATIC against the previous sample's output (resolve.cu:250-258)
                const float hm = o.w;
                const float wsum = hm + sp.w;
                const float hf = hm / wsum;
                o.c0 = lerpf(sp.c0, o.c0, hf);
                o.c1 = lerpf(sp.c1, o.c1, hf);
                o.c2 = lerpf(sp.c2, o.c2, hf);
                o.d = lerpf(sp.d, o.d, hf);
                o.a = lerpf(sp.a, o.a, hf);
                o.w = wsum;
            }
            fs++;
            sp = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0u };
        }
    }
    if (active) store_output(vd.out, pix, o);
}

hipError_t launch_batch_resolve(const SceneArgs& sc, const ViewDesc* views, uint32_t n_views,
                                const uint32_t* block_prefix, uint32_t n_blocks, const uint64_t* fb,
                                const float4* samples, uint32_t level, hipStream_t stream)
{
    if (n_blocks == 0 || n_views == 0) return hipSuccess;
    switch (sc.sh_bases) {
        case 1: hipLaunchKernelGGL((batch_resolve_kernel<1>), dim3(n_blocks), dim3(kBlockResolve), 0, stream, sc, views, n_views, block_prefix, fb, samples, level); break;
        case 4: hipLaunchKernelGGL((batch_resolve_kernel<4>), dim3(n_blocks), dim3(kBlockResolve), 0, stream, sc, views, n_views, block_prefix, fb, samples, level); break;
        case 9: hipLaunchKernelGGL((batch_resolve_kernel<9>), dim3(n_blocks), dim3(kBlockResolve), 0, stream, sc, views, n_views, block_prefix, fb, samples, level); break;
        default: hipLaunchKernelGGL((batch_resolve_kernel<16>), dim3(n_blocks), dim3(kBlockResolve), 0, stream, sc, views, n_views, block_prefix, fb, samples, level); break;
    }
    return hipGetLastError();
}

}  // namespace sgs
