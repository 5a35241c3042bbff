// Device-side building blocks of the stochastic splatting renderer for gfx950 (wave64).
//
// Every floating-point expression here is evaluated in a fixed order; the library is built
// with -ffp-contract=off (no implicit FMA) and IEEE division/sqrt, and uses its own
// exp/exp2/log/sin/cos polynomials, so results are reproducible bit for bit on any IEEE
// machine that evaluates the same sequence. Reference routines restated here are cited as
// file:line of StochasticGSCudaBackend/StochasticGSCudaBackend/.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgs {

// ------------------------------------------------------------------------------------------
// constants (gs_utils.h:13-18, math_utils.h:10, resolve.h:9)
// ------------------------------------------------------------------------------------------
constexpr float kMinAlpha = 1.0f / 255.0f;
constexpr float kLog2e = 1.44269504088896340736f;
constexpr float kLn2Hi = 0.693359375f;
constexpr float kLn2Lo = -2.12194440054713770e-4f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr uint32_t kKBuffer = 4;
constexpr float kMaxAlpha = 0.99f;
constexpr uint64_t kEmpty = ~0ull;

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

// ------------------------------------------------------------------------------------------
// math: range reduction + Taylor polynomials by Horner's rule with fma
// ------------------------------------------------------------------------------------------
// 2^x; 0 below -125 (never subnormal) and for NaN
__device__ __forceinline__ float exp2_f(float x)
{
    if (!(x >= -125.0f)) return 0.0f;
    if (x >= 128.0f) return __builtin_huge_valf();
    const float n = rintf(x);
    const float r = x - n;
    float p = 1.52527338040598380e-5f;
    p = fmaf(p, r, 1.54035303933816060e-4f);
    p = fmaf(p, r, 1.33335581464284410e-3f);
    p = fmaf(p, r, 9.61812910762847700e-3f);
    p = fmaf(p, r, 5.55041086648215760e-2f);
    p = fmaf(p, r, 2.40226506959100700e-1f);
    p = fmaf(p, r, 6.93147180559945300e-1f);
    p = fmaf(p, r, 1.0f);
    const int ni = (int)n;
    const int nc = ni > 127 ? 127 : ni;
    float result = p * u2f((uint32_t)(nc + 127) << 23);
    if (ni > 127) result = result * 2.0f;
    return result;
}

// e^x (Cody-Waite reduction); 0 below -86.5
__device__ __forceinline__ float exp_f(float x)
{
    if (x != x) return x;
    if (x > 88.5f) return __builtin_huge_valf();
    if (x < -86.5f) return 0.0f;
    const float n = rintf(x * kLog2e);
    float r = fmaf(-n, kLn2Hi, x);
    r = fmaf(-n, kLn2Lo, r);
    float p = 1.98412698412698410e-4f;
    p = fmaf(p, r, 1.38888888888888890e-3f);
    p = fmaf(p, r, 8.33333333333333300e-3f);
    p = fmaf(p, r, 4.16666666666666640e-2f);
    p = fmaf(p, r, 1.66666666666666660e-1f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    const int ni = (int)n;
    const int nc = ni > 127 ? 127 : ni;
    float result = p * u2f((uint32_t)(nc + 127) << 23);
    if (ni > 127) result = result * 2.0f;
    return result;
}

// natural logarithm: x = m*2^e, log(m) = 2 atanh((m-1)/(m+1))
__device__ __forceinline__ float log_f(float x)
{
    if (x != x || x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_huge_valf();
    if (x == __builtin_huge_valf()) return x;
    int e = 0;
    uint32_t b = f2u(x);
    if (b < 0x00800000u) {
        x = x * 8388608.0f;
        b = f2u(x);
        e = -23;
    }
    e += (int)(b >> 23) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421354f) {
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (m + 1.0f);
    const float z = s * s;
    float p = 1.81818181818181820e-1f;
    p = fmaf(p, z, 2.22222222222222220e-1f);
    p = fmaf(p, z, 2.85714285714285710e-1f);
    p = fmaf(p, z, 4.00000000000000000e-1f);
    p = fmaf(p, z, 6.66666666666666630e-1f);
    p = fmaf(p, z, 2.0f);
    const float lm = s * p;
    const float ef = (float)e;
    return fmaf(ef, kLn2Hi, fmaf(ef, kLn2Lo, lm));
}

// sin(2 pi u), cos(2 pi u) with an exact quadrant reduction on u
__device__ __forceinline__ void sincos_2pi(float u, float& s, float& c)
{
    const float q = rintf(u * 4.0f);
    const float r = u - q * 0.25f;
    const float a = r * kTwoPi;
    const float a2 = a * a;
    float sp = 2.75573192239858900e-6f;
    sp = fmaf(sp, a2, -1.98412698412698410e-4f);
    sp = fmaf(sp, a2, 8.33333333333333300e-3f);
    sp = fmaf(sp, a2, -1.66666666666666660e-1f);
    const float sa = fmaf(a * a2, sp, a);
    float cp = -2.75573192239858900e-7f;
    cp = fmaf(cp, a2, 2.48015873015873000e-5f);
    cp = fmaf(cp, a2, -1.38888888888888890e-3f);
    cp = fmaf(cp, a2, 4.16666666666666640e-2f);
    cp = fmaf(cp, a2, -0.5f);
    const float ca = fmaf(a2, cp, 1.0f);
    switch (((int)q) & 3) {
        case 0: s = sa; c = ca; break;
        case 1: s = ca; c = -sa; break;
        case 2: s = -sa; c = -ca; break;
        default: s = -ca; c = sa; break;
    }
}

// ------------------------------------------------------------------------------------------
// counter-based RNG (math_utils.h:45-79, 113-136)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint2 make_seed(uint32_t idx, uint32_t frame)
{
    uint32_t h1 = hash32(idx ^ (frame * 0x9E3779B9u));
    uint32_t h2 = hash32((idx * 0xBB67AE85u) ^ frame);
    h1 ^= h2 * 0x3C6EF372u;
    h2 ^= h1 * 0xA54FF53Au;
    return make_uint2(h1, h2);
}

__device__ __forceinline__ uint2 pcg2d_uint(uint2 v)
{
    uint32_t x = v.x * 1664525u + 1013904223u;
    uint32_t y = v.y * 1664525u + 1013904223u;
    x += y * 1664525u;
    y += x * 1664525u;
    x ^= x >> 16;
    y ^= y >> 16;
    x += y * 1664525u;
    y += x * 1664525u;
    x ^= x >> 16;
    y ^= y >> 16;
    return make_uint2(x, y);
}

__device__ __forceinline__ float2 pcg2d(uint2& state)
{
    state = pcg2d_uint(state);
    return make_float2((float)state.x * 0x1p-32f, (float)state.y * 0x1p-32f);
}

__device__ __forceinline__ uint32_t pcg_hash(uint32_t seed)
{
    const uint32_t state = seed * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// uniform [0,1) for fragment (instance, x, y) of a frame
__device__ __forceinline__ float rand_uniform(uint32_t frame, uint32_t instance, uint32_t x, uint32_t y)
{
    const uint32_t seed = instance ^ (frame * 1973u) ^ (x * 9277u) ^ (y * 2699u);
    return u2f(0x3f800000u | (pcg_hash(seed) >> 9u)) - 1.0f;
}

// 128x128 blue-noise slice with per-primitive R2 offsets, 0.32 fixed-point rotation
__device__ __forceinline__ float blue_noise(const uint8_t* __restrict__ slice, uint32_t rotation, uint32_t instance,
                                            uint32_t x, uint32_t y)
{
    const uint32_t ox = (instance * 0xC13FA9A9u) >> 25u;
    const uint32_t oy = (instance * 0x91E10DA5u) >> 25u;
    const uint32_t texel = ((y + oy) & 127u) * 128u + ((x + ox) & 127u);
    const uint32_t value = ((uint32_t)slice[texel] << 24u) + (1u << 23u) + rotation;
    return (float)(value >> 8u) * 0x1p-24f;
}

// ------------------------------------------------------------------------------------------
// GPS sampling math (math_utils.h:14-43, 81-111)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float dilog(float x)
{
    float y = -0.952623943f;
    y = fmaf(y, x, 2.995873777f);
    y = fmaf(y, x, -3.762115247f);
    y = fmaf(y, x, 2.334605281f);
    y = fmaf(y, x, -0.773941317f);
    y = fmaf(y, x, 5.97506396e-02f);
    y = fmaf(y, x, -0.256617940f);
    y = fmaf(y, x, 2.000061866f);
    y = y * x;
    const float s = 1.0f - x;
    if (s > 0.0f) y = y + s * log_f(fmaxf(s, 1.0e-37f));
    return y;
}

__device__ __forceinline__ float inv_dilog(float x)
{
    const float t = fminf(x / 1.6449340668482264f, 1.0f);
    float y = -1.27463503e+01f;
    y = fmaf(y, t, 5.88993459e+01f);
    y = fmaf(y, t, -1.16025780e+02f);
    y = fmaf(y, t, 1.26945827e+02f);
    y = fmaf(y, t, -8.43108826e+01f);
    y = fmaf(y, t, 3.48799862e+01f);
    y = fmaf(y, t, -8.89606235e+00f);
    y = fmaf(y, t, 1.38640936e+00f);
    y = fmaf(y, t, -7.80640876e-01f);
    y = fmaf(y, t, 1.64841888e+00f);
    y = fmaf(y, t, -2.82836687e-05f);
    return y;
}

__device__ __forceinline__ float2 box_muller(float u1, float u2)
{
    const float R = sqrtf(fmaxf(-2.0f * log_f(fmaxf(fminf(u1, 1.0f), 1.0e-37f)), 0.0f));
    float s, c;
    sincos_2pi(u2, s, c);
    return make_float2(R * c, R * s);
}

// radius-corrected Box-Muller: offsets of an opacity-weighted Gaussian (math_utils.h:89-96)
__device__ __forceinline__ float2 corrected_box_muller(float u1, float u2, float alpha_rcp, float alpha_dilog)
{
    const float a = alpha_rcp * inv_dilog((1.0f - u1) * alpha_dilog);
    const float R = sqrtf(fmaxf(-2.0f * log_f(fmaxf(fminf(a, 1.0f), 1.0e-37f)), 0.0f));
    float s, c;
    sincos_2pi(u2, s, c);
    return make_float2(R * c, R * s);
}

__device__ __forceinline__ uint32_t f2u_rn_sat(float v)
{
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967040.0f) return 0xffffffffu;
    return (uint32_t)rintf(v);
}

// normal approximation of a Poisson draw with Cornish-Fisher correction (math_utils.h:98-111)
__device__ __forceinline__ uint32_t poisson(uint2& state, float lambda)
{
    const float2 u = pcg2d(state);
    const float w = box_muller(u.x, u.y).x;
    const float w2 = w * w;
    const float w3 = w2 * w;
    const float w4 = w2 * w2;
    const float s = sqrtf(lambda);
    const float s_rcp = 1.0f / s;
    const float l_rcp = 1.0f / lambda;
    float kf = lambda + s * w + (w2 - 1.0f) / 6.0f;
    kf = kf + s_rcp * (-(1.0f / 36.0f) * w - (1.0f / 72.0f) * w3);
    kf = kf + l_rcp * (-(8.0f / 405.0f) + (7.0f / 810.0f) * w2 + (1.0f / 270.0f) * w4);
    return f2u_rn_sat(fmaxf(kf, 0.0f));
}

// ------------------------------------------------------------------------------------------
// camera in framebuffer (sub)pixels
// ------------------------------------------------------------------------------------------
struct FbCamera {
    float width, height;        // framebuffer size as float
    float focal_x, focal_y, center_x, center_y;
    float near_plane, far_plane;
    float ss;                   // supersampling factor as float
    uint32_t width_i, height_i; // framebuffer size
    int32_t proper_antialiasing;
    // view-space clamp of the Jacobian's evaluation point, 15% outside the framebuffer
    // (gs_utils.h:84-90): (-0.15 * width - center_x) / focal_x etc., computed once per view
    float clip_left, clip_right, clip_top, clip_bottom;
};

struct Projected {
    float depth, opacity, determinant, power_threshold;
    float mean_x, mean_y;
    float conic_x, conic_y, conic_z;
    float cov_x, cov_y, cov_z;
    float bx0, bx1, by0, by1;
};

// view-space depth of Gaussian g (the projection's depth, bit for bit)
__device__ __forceinline__ float view_depth(const float* __restrict__ means, uint32_t g, const float4 r3)
{
    const float mx = means[3u * g + 0u], my = means[3u * g + 1u], mz = means[3u * g + 2u];
    return r3.x * mx + r3.y * my + r3.z * mz + r3.w;
}

// View-independent part of the projection (gs_utils.h:38-80): sigmoid opacity and the 3D
// covariance R S S^T R^T of the normalised quaternion. False if the Gaussian is culled in
// every view (opacity below 1/255 or a degenerate quaternion). pt: log(255 opacity), the
// power threshold of views without proper antialiasing, when known (NaN: computed per view).
struct Cov3d {
    float c11, c12, c13, c22, c23, c33, opacity;
    float pt;
};

__device__ __forceinline__ bool gaussian_cov3d(const float* __restrict__ scales, const float* __restrict__ rotations,
                                               const float* __restrict__ opacities, uint32_t g, Cov3d& c)
{
    c.opacity = 1.0f / (1.0f + exp_f(-opacities[g]));
    c.pt = __builtin_nanf("");
    if (c.opacity < kMinAlpha) return false;

    const float vx = exp_f(2.0f * scales[3u * g + 0u]);
    const float vy = exp_f(2.0f * scales[3u * g + 1u]);
    const float vz = exp_f(2.0f * scales[3u * g + 2u]);
    const float4 q = reinterpret_cast<const float4*>(rotations)[g];
    const float qr = q.x, qi = q.y, qj = q.z, qk = q.w;
    const float ii = qi * qi, jj = qj * qj, kk = qk * qk;
    const float ij = qi * qj, ik = qi * qk, jk = qj * qk;
    const float ri = qr * qi, rj = qr * qj, rk = qr * qk;
    const float norm_sq = qr * qr + ii + jj + kk;
    if (norm_sq < 1e-8f) return false;
    const float nr = 1.0f / norm_sq;
    const float R11 = 1.0f - 2.0f * (jj + kk) * nr, R12 = 2.0f * (ij - rk) * nr, R13 = 2.0f * (ik + rj) * nr;
    const float R21 = 2.0f * (ij + rk) * nr, R22 = 1.0f - 2.0f * (ii + kk) * nr, R23 = 2.0f * (jk - ri) * nr;
    const float R31 = 2.0f * (ik - rj) * nr, R32 = 2.0f * (jk + ri) * nr, R33 = 1.0f - 2.0f * (ii + jj) * nr;
    const float S11 = R11 * vx, S12 = R12 * vy, S13 = R13 * vz;
    const float S21 = R21 * vx, S22 = R22 * vy, S23 = R23 * vz;
    const float S31 = R31 * vx, S32 = R32 * vy, S33 = R33 * vz;
    c.c11 = S11 * R11 + S12 * R12 + S13 * R13;
    c.c12 = S11 * R21 + S12 * R22 + S13 * R23;
    c.c13 = S11 * R31 + S12 * R32 + S13 * R33;
    c.c22 = S21 * R21 + S22 * R22 + S23 * R23;
    c.c23 = S21 * R31 + S22 * R32 + S23 * R33;
    c.c33 = S31 * R31 + S32 * R32 + S33 * R33;
    return true;
}

// View-dependent part of the EWA projection (gs_utils.h:80-193) for a Gaussian at view
// depth `depth` (already inside [near, far]); returns false if culled.
__device__ __forceinline__ bool project_cov3d(float mx, float my, float mz, float depth, const Cov3d& cv,
                                              const float4 r1, const float4 r2, const float4 r3, const FbCamera& cam,
                                              Projected& o)
{
    float opacity = cv.opacity;
    const float c11 = cv.c11, c12 = cv.c12, c13 = cv.c13, c22 = cv.c22, c23 = cv.c23, c33 = cv.c33;
    const float x = (r1.x * mx + r1.y * my + r1.z * mz + r1.w) / depth;
    const float y = (r2.x * mx + r2.y * my + r2.z * mz + r2.w) / depth;

    const float xc = fmaxf(cam.clip_left, fminf(x, cam.clip_right));
    const float yc = fmaxf(cam.clip_top, fminf(y, cam.clip_bottom));
    const float j11 = cam.focal_x / depth;
    const float j13 = -j11 * xc;
    const float j22 = cam.focal_y / depth;
    const float j23 = -j22 * yc;
    const float a1 = j11 * r1.x + j13 * r3.x, a2 = j11 * r1.y + j13 * r3.y, a3 = j11 * r1.z + j13 * r3.z;
    const float b1 = j22 * r2.x + j23 * r3.x, b2 = j22 * r2.y + j23 * r3.y, b3 = j22 * r2.z + j23 * r3.z;
    const float p1 = a1 * c11 + a2 * c12 + a3 * c13;
    const float p2 = a1 * c12 + a2 * c22 + a3 * c23;
    const float p3 = a1 * c13 + a2 * c23 + a3 * c33;
    const float q1 = b1 * c11 + b2 * c12 + b3 * c13;
    const float q2 = b1 * c12 + b2 * c22 + b3 * c23;
    const float q3 = b1 * c13 + b2 * c23 + b3 * c33;
    float cxx = p1 * a1 + p2 * a2 + p3 * a3;
    const float cxy = p1 * b1 + p2 * b2 + p3 * b3;
    float cyy = q1 * b1 + q2 * b2 + q3 * b3;
    const float det_raw = cxx * cyy - cxy * cxy;
    const float kernel = (cam.proper_antialiasing ? 0.1f : 0.3f) * cam.ss * cam.ss;
    cxx = cxx + kernel;
    cyy = cyy + kernel;
    const float det = cxx * cyy - cxy * cxy;
    if (det < 1e-6f) return false;
    o.conic_x = cyy / det;
    o.conic_y = -cxy / det;
    o.conic_z = cxx / det;
    if (cam.proper_antialiasing) {
        opacity = opacity * sqrtf(fmaxf(det_raw / det, 0.0f));
        if (opacity < kMinAlpha) return false;
    }
    o.mean_x = x * cam.focal_x + cam.center_x;
    o.mean_y = y * cam.focal_y + cam.center_y;
    // without antialiasing the opacity, hence the threshold, is the same in every view
    const float pt = !cam.proper_antialiasing && cv.pt == cv.pt ? cv.pt : log_f(opacity * 255.0f);
    const float cutoff = 2.0f * pt;
    const float ex = sqrtf(cxx * cutoff);
    const float ey = sqrtf(cyy * cutoff);
    o.bx0 = o.mean_x - ex;
    o.bx1 = o.mean_x + ex;
    o.by0 = o.mean_y - ey;
    o.by1 = o.mean_y + ey;
    if (!(o.bx0 < cam.width) || !(o.bx1 >= 0.0f) || !(o.by0 < cam.height) || !(o.by1 >= 0.0f)) return false;
    o.depth = depth;
    o.opacity = opacity;
    o.determinant = det;
    o.power_threshold = pt;
    o.cov_x = cxx;
    o.cov_y = cxy;
    o.cov_z = cyy;
    return true;
}

// EWA projection with culling (gs_utils.h:38-193); returns false if culled.
// w2c rows r1, r2, r3 (row-major, 4 floats each).
__device__ __forceinline__ bool project_gaussian(const float* __restrict__ means, const float* __restrict__ scales,
                                                 const float* __restrict__ rotations,
                                                 const float* __restrict__ opacities, uint32_t g,
                                                 const float4 r1, const float4 r2, const float4 r3,
                                                 const FbCamera& cam, Projected& o)
{
    const float mx = means[3u * g + 0u], my = means[3u * g + 1u], mz = means[3u * g + 2u];
    const float depth = view_depth(means, g, r3);
    if (!(depth >= cam.near_plane && depth <= cam.far_plane)) return false;
    Cov3d cv;
    if (!gaussian_cov3d(scales, rotations, opacities, g, cv)) return false;
    return project_cov3d(mx, my, mz, depth, cv, r1, r2, r3, cam, o);
}

// Cached view-independent part: two float4 per Gaussian, (c11, c12, c13, c22) and
// (c23, c33, opacity, log(255 opacity)); opacity < 0 marks a Gaussian culled in every view.
__device__ __forceinline__ void store_cov3d(float4* __restrict__ cache, uint32_t g, bool valid, const Cov3d& c)
{
    cache[2u * g] = make_float4(c.c11, c.c12, c.c13, c.c22);
    cache[2u * g + 1u] = make_float4(c.c23, c.c33, valid ? c.opacity : -1.0f, valid ? log_f(c.opacity * 255.0f) : 0.0f);
}

// project_gaussian with the view-independent part read from the cache (bit-identical)
__device__ __forceinline__ bool project_cached(const float* __restrict__ means, const float4* __restrict__ cache,
                                               uint32_t g, const float4 r1, const float4 r2, const float4 r3,
                                               const FbCamera& cam, Projected& o)
{
    const float mx = means[3u * g + 0u], my = means[3u * g + 1u], mz = means[3u * g + 2u];
    const float depth = view_depth(means, g, r3);
    if (!(depth >= cam.near_plane && depth <= cam.far_plane)) return false;
    const float4 a = cache[2u * g], b = cache[2u * g + 1u];
    if (b.z < 0.0f) return false;
    const Cov3d cv = { a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w };
    return project_cov3d(mx, my, mz, depth, cv, r1, r2, r3, cam, o);
}

// ------------------------------------------------------------------------------------------
// fragments (scatter_utils.cuh:33-46)
// ------------------------------------------------------------------------------------------
struct ConicExp2 {
    float a, b, c;
};

__device__ __forceinline__ ConicExp2 conic_exp2(float cx, float cy, float cz)
{
    return { -0.5f * kLog2e * cx, -kLog2e * cy, -0.5f * kLog2e * cz };
}

// exponent of a fragment at offset (dx, dy) from the mean, in log2 units
__device__ __forceinline__ float fragment_exponent(const ConicExp2 ce, float dx, float dy)
{
    const float t = fmaf(ce.a, dx, ce.b * dy);
    return fmaf(dx, t, ce.c * dy * dy);
}

__device__ __forceinline__ float fragment_alpha(const ConicExp2 ce, float opacity, float dx, float dy)
{
    return opacity * exp2_f(fminf(fragment_exponent(ce, dx, dy), 0.0f));
}

__device__ __forceinline__ uint64_t fragment_key(float depth, uint32_t g)
{
    return ((uint64_t)f2u(depth) << 32) | (uint64_t)g;
}

struct Aabb {
    uint32_t x0, x1, y0, y1;
};

__device__ __forceinline__ Aabb pixel_aabb(const Projected& p, const FbCamera& cam)
{
    Aabb b;
    b.x0 = (uint32_t)floorf(fmaxf(p.bx0, 0.0f));
    b.x1 = (uint32_t)ceilf(fminf(p.bx1, cam.width));
    b.y0 = (uint32_t)floorf(fmaxf(p.by0, 0.0f));
    b.y1 = (uint32_t)ceilf(fminf(p.by1, cam.height));
    return b;
}

// ------------------------------------------------------------------------------------------
// GPS point k of a Gaussian: per-point counter seed, radius-corrected Box-Muller offset through
// the Cholesky factor L of cov2d, power test at the pixel centre (scatter_utils.cuh:115-130).
__device__ __forceinline__ bool gps_point_at(uint32_t k, uint32_t seed_x, uint32_t seed_y, float mx, float my, float L0,
                                          float L1, float L2, float cx, float cy, float cz, float pt, float o_rcp,
                                          float o_dilog, const FbCamera& cam, uint32_t& pixel)
{
    uint2 st = make_uint2(seed_x ^ ((k + 1u) * 0x9E3779B9u), seed_y);
    const float2 u = pcg2d(st);
    const float2 xy = corrected_box_muller(u.x, u.y, o_rcp, o_dilog);
    const float sx = mx + L0 * xy.x;
    const float sy = my + (L1 * xy.x + L2 * xy.y);
    if (!(sx < cam.width && sx >= 0.0f && sy < cam.height && sy >= 0.0f)) return false;
    const float dcx = truncf(sx) + 0.5f - mx;
    const float dcy = truncf(sy) + 0.5f - my;
    const float power = 0.5f * (cx * dcx * dcx + cz * dcy * dcy) + cy * dcx * dcy;
    if (power > pt) return false;
    pixel = (uint32_t)sy * cam.width_i + (uint32_t)sx;
    return true;
}

// wave-uniform broadcast of a lane's value through an SGPR (v_readlane_b32)
__device__ __forceinline__ uint32_t bcast(uint32_t v, int lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ float bcast(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ uint64_t bcast(uint64_t v, int lane)
{
    const uint32_t lo = bcast((uint32_t)v, lane), hi = bcast((uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

}  // namespace sgs
