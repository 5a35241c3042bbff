/*
 * Scalar CPU oracle of the sorting-free stochastic splatting renderer.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels, never part of the
 * product path. Compile with -O2 -ffp-contract=off (no fast math): every FP32
 * operation below is evaluated exactly as written, in source order.
 *
 * Citations are to the reference tree StochasticGSCudaBackend/StochasticGSCudaBackend/.
 * The reference is built with -use_fast_math (setup.py:84), so its exact bits are not
 * reachable; this oracle fixes one IEEE-reproducible evaluation order and the HIP kernels
 * follow it. How each function is pinned (tests/ file in brackets):
 *   orc_exp2f/expf/logf/sincos_2pi  FP64 libm within a few ulp, special values [test_oracle_math]
 *   RNG (hash32 .. blue_noise)       independent Python-integer restatement + uniformity [test_oracle_rng]
 *   dilog / inv_dilog                FP64 Li2 series, Li2(1)=pi^2/6, inverse round trip [test_oracle_math]
 *   poisson                          mean/variance against lambda [test_oracle_math]
 *   corrected_box_muller             radial CDF of the opacity-weighted Gaussian [test_oracle_math]
 *   orc_project                      FP64 closed-form EWA (Zwicker et al. 2001), cull cases [test_oracle_project]
 *   orc_eval_sh                      FP64 real SH basis from associated Legendre polynomials [test_oracle_shade]
 *   catmull_rom / ycocg              partition of unity, linear reproduction, inverse transform [test_oracle_shade]
 *   orc_accumulate (ST)              pixel-major brute force without AABBs, thread-count invariance,
 *                                    convergence to orc_blend_fp64 [test_oracle_accumulate]
 *   orc_accumulate (GPS)             point counts against lambda, winners inside the image,
 *                                    point density against the Gaussian [test_oracle_accumulate]
 *   orc_resolve                      box filter = subpixel mean, STATIC = running mean (closed form),
 *                                    REPROJECT with identity motion on a constant field [test_oracle_resolve];
 *                                    REPROJECT under motion: whole-pixel translations read one known texel,
 *                                    checkerboard moments fix gamma(velocity), linear-ramp histories are
 *                                    fetched at the FP64 reprojected position [test_oracle_reuse_motion]
 *   orc_sample spatial reuse         weights = FP64 alpha ratios of the winner, k-buffer occlusion product,
 *                                    insertion and eviction, neighbour components in the resolve
 *                                    [test_oracle_reuse_motion]
 *   everything                       ASan + UBSan build over every entry point [test_oracle_sanitize]
 */
#include "sgs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------ */
/* bit helpers                                                                                */
/* ------------------------------------------------------------------------------------------ */
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ------------------------------------------------------------------------------------------ */
/* math library: range reduction + Taylor polynomials evaluated by Horner's rule with fma.    */
/* exp/exp2: Cody & Waite (1980) reduction x = n*ln2 + r; log: m*2^e with                     */
/* log(m) = 2*atanh((m-1)/(m+1)); sin/cos(2*pi*u): exact quadrant reduction on u.             */
/* ------------------------------------------------------------------------------------------ */
#define ORC_LOG2E 1.44269504088896340736f
#define ORC_LN2_HI 0.693359375f
#define ORC_LN2_LO -2.12194440054713770e-4f
#define ORC_TWO_PI 6.28318530717958647692f

/* 2^x; 0 below -125 (never subnormal) and for NaN. */
float orc_exp2f(float x)
{
    if (!(x >= -125.0f)) return 0.0f;
    if (x >= 128.0f) return INFINITY;
    const float n = rintf(x);
    const float r = x - n; /* exact, |r| <= 0.5 */
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

/* e^x; 0 below -86.5 (never subnormal). */
float orc_expf(float x)
{
    if (x != x) return x;
    if (x > 88.5f) return INFINITY;
    if (x < -86.5f) return 0.0f;
    const float n = rintf(x * ORC_LOG2E);
    float r = fmaf(-n, ORC_LN2_HI, x);
    r = fmaf(-n, ORC_LN2_LO, r);
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

/* natural logarithm */
float orc_logf(float x)
{
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    uint32_t b = f2u(x);
    if (b < 0x00800000u) { x = x * 8388608.0f; b = f2u(x); e = -23; }
    e += (int)(b >> 23) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421354f) { m = m * 0.5f; e += 1; }
    const float f = m - 1.0f;           /* exact */
    const float s = f / (m + 1.0f);
    const float z = s * s;
    float p = 1.81818181818181820e-1f; /* 2/11 */
    p = fmaf(p, z, 2.22222222222222220e-1f);
    p = fmaf(p, z, 2.85714285714285710e-1f);
    p = fmaf(p, z, 4.00000000000000000e-1f);
    p = fmaf(p, z, 6.66666666666666630e-1f);
    p = fmaf(p, z, 2.0f);
    const float lm = s * p;
    const float ef = (float)e;
    return fmaf(ef, ORC_LN2_HI, fmaf(ef, ORC_LN2_LO, lm));
}

/* sin(2*pi*u), cos(2*pi*u) */
void orc_sincos_2pi(float u, float* s, float* c)
{
    const float q = rintf(u * 4.0f);
    const float r = u - q * 0.25f;      /* exact, |r| <= 1/8 */
    const float a = r * ORC_TWO_PI;
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
        case 0: *s = sa; *c = ca; break;
        case 1: *s = ca; *c = -sa; break;
        case 2: *s = -sa; *c = -ca; break;
        default: *s = -ca; *c = sa; break;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* integer RNG: math_utils.h:45-79 (hash32, make_seed, pcg2d), :113-122 (pcg_hash, rand),     */
/* :124-136 (blue noise)                                                                      */
/* ------------------------------------------------------------------------------------------ */
uint32_t orc_hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

void orc_make_seed(uint32_t idx, uint32_t frame, uint32_t out[2])
{
    uint32_t h1 = orc_hash32(idx ^ (frame * 0x9E3779B9u));
    uint32_t h2 = orc_hash32((idx * 0xBB67AE85u) ^ frame);
    h1 ^= h2 * 0x3C6EF372u;
    h2 ^= h1 * 0xA54FF53Au;
    out[0] = h1;
    out[1] = h2;
}

void orc_pcg2d_uint(const uint32_t in[2], uint32_t out[2])
{
    uint32_t x = in[0] * 1664525u + 1013904223u;
    uint32_t y = in[1] * 1664525u + 1013904223u;
    x += y * 1664525u;
    y += x * 1664525u;
    x ^= x >> 16;
    y ^= y >> 16;
    x += y * 1664525u;
    y += x * 1664525u;
    x ^= x >> 16;
    y ^= y >> 16;
    out[0] = x;
    out[1] = y;
}

void orc_pcg2d(uint32_t state[2], float out[2])
{
    uint32_t next[2];
    orc_pcg2d_uint(state, next);
    state[0] = next[0];
    state[1] = next[1];
    out[0] = (float)next[0] * 0x1p-32f;
    out[1] = (float)next[1] * 0x1p-32f;
}

uint32_t orc_pcg_hash(uint32_t seed)
{
    const uint32_t state = seed * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

float orc_rand(uint32_t frame, uint32_t instance, uint32_t x, uint32_t y)
{
    const uint32_t seed = instance ^ (frame * 1973u) ^ (x * 9277u) ^ (y * 2699u);
    return u2f(0x3f800000u | (orc_pcg_hash(seed) >> 9u)) - 1.0f;
}

float orc_blue_noise(const uint8_t* slice, uint32_t rotation, uint32_t instance, uint32_t x, uint32_t y)
{
    const uint32_t offset_x = (instance * 0xC13FA9A9u) >> 25u;
    const uint32_t offset_y = (instance * 0x91E10DA5u) >> 25u;
    const uint32_t texel = ((y + offset_y) % 128u) * 128u + (x + offset_x) % 128u;
    const uint32_t value = ((uint32_t)slice[texel] << 24u) + (1u << 23u) + rotation;
    return (float)(value >> 8u) * 0x1p-24f;
}

/* Per-point counter-based seed (accumulate_hybrid_gps.cu:163-165) used for every GPS point,
 * scattered or chunked, so the result does not depend on the schedule. Point p uses
 * index p + 1: index 0 would reuse the stream of the Poisson count draw. */
void orc_gps_point_seed(uint32_t gaussian_idx, uint32_t frame, uint32_t point_idx, uint32_t out[2])
{
    orc_make_seed(gaussian_idx, frame, out);
    out[0] ^= (point_idx + 1u) * 0x9E3779B9u;
}

/* ------------------------------------------------------------------------------------------ */
/* GPS sampling math: math_utils.h:14-43 (dilog, inv_dilog), :81-96 (Box-Muller variants),   */
/* :98-111 (Poisson). Polynomial coefficients are the reference's approximation.              */
/* ------------------------------------------------------------------------------------------ */
float orc_dilog(float x)
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
    if (s > 0.0f) y = y + s * orc_logf(fmaxf(s, 1.0e-37f));
    return y;
}

float orc_inv_dilog(float x)
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

void orc_box_muller(float u1, float u2, float out[2])
{
    const float R = sqrtf(fmaxf(-2.0f * orc_logf(fmaxf(fminf(u1, 1.0f), 1.0e-37f)), 0.0f));
    float s, c;
    orc_sincos_2pi(u2, &s, &c);
    out[0] = R * c;
    out[1] = R * s;
}

void orc_corrected_box_muller(float u1, float u2, float alpha_rcp, float alpha_dilog, float out[2])
{
    const float a = alpha_rcp * orc_inv_dilog((1.0f - u1) * alpha_dilog);
    const float R = sqrtf(fmaxf(-2.0f * orc_logf(fmaxf(fminf(a, 1.0f), 1.0e-37f)), 0.0f));
    float s, c;
    orc_sincos_2pi(u2, &s, &c);
    out[0] = R * c;
    out[1] = R * s;
}

/* round-half-even float -> u32 of a non-negative value, saturating like __float2uint_rn */
static inline uint32_t f2u_rn_sat(float v)
{
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967040.0f) return 0xffffffffu;
    return (uint32_t)rintf(v);
}

uint32_t orc_poisson(uint32_t state[2], float lambda)
{
    float u[2], bm[2];
    orc_pcg2d(state, u);
    orc_box_muller(u[0], u[1], bm);
    const float w = bm[0];
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

/* ------------------------------------------------------------------------------------------ */
/* projection: gs_utils.h:38-193 (constants :13-18)                                           */
/* ------------------------------------------------------------------------------------------ */
#define ORC_MIN_ALPHA (1.0f / 255.0f)

void orc_project(const orc_scene* sc, uint32_t g, const float* w2c,
                 float width, float height, float focal_x, float focal_y, float center_x, float center_y,
                 float near_plane, float far_plane, int32_t proper_antialiasing, float ss,
                 orc_proj* o)
{
    int active = 1;
    const float mx = sc->means[3 * g + 0], my = sc->means[3 * g + 1], mz = sc->means[3 * g + 2];
    const float* r1 = w2c;
    const float* r2 = w2c + 4;
    const float* r3 = w2c + 8;
    const float depth = r3[0] * mx + r3[1] * my + r3[2] * mz + r3[3];
    o->in_depth_range = (depth >= near_plane && depth <= far_plane);
    if (!o->in_depth_range) active = 0;
    float opacity = 1.0f / (1.0f + orc_expf(-sc->opacities[g]));
    if (opacity < ORC_MIN_ALPHA) active = 0;

    const float vx = orc_expf(2.0f * sc->scales[3 * g + 0]);
    const float vy = orc_expf(2.0f * sc->scales[3 * g + 1]);
    const float vz = orc_expf(2.0f * sc->scales[3 * g + 2]);
    const float qr = sc->rotations[4 * g + 0], qi = sc->rotations[4 * g + 1];
    const float qj = sc->rotations[4 * g + 2], qk = sc->rotations[4 * g + 3];
    const float ii = qi * qi, jj = qj * qj, kk = qk * qk;
    const float ij = qi * qj, ik = qi * qk, jk = qj * qk;
    const float ri = qr * qi, rj = qr * qj, rk = qr * qk;
    const float norm_sq = qr * qr + ii + jj + kk;
    if (norm_sq < 1e-8f) active = 0;
    const float nr = 1.0f / norm_sq;
    const float R11 = 1.0f - 2.0f * (jj + kk) * nr, R12 = 2.0f * (ij - rk) * nr, R13 = 2.0f * (ik + rj) * nr;
    const float R21 = 2.0f * (ij + rk) * nr, R22 = 1.0f - 2.0f * (ii + kk) * nr, R23 = 2.0f * (jk - ri) * nr;
    const float R31 = 2.0f * (ik - rj) * nr, R32 = 2.0f * (jk + ri) * nr, R33 = 1.0f - 2.0f * (ii + jj) * nr;
    const float S11 = R11 * vx, S12 = R12 * vy, S13 = R13 * vz;
    const float S21 = R21 * vx, S22 = R22 * vy, S23 = R23 * vz;
    const float S31 = R31 * vx, S32 = R32 * vy, S33 = R33 * vz;
    const float c11 = S11 * R11 + S12 * R12 + S13 * R13;
    const float c12 = S11 * R21 + S12 * R22 + S13 * R23;
    const float c13 = S11 * R31 + S12 * R32 + S13 * R33;
    const float c22 = S21 * R21 + S22 * R22 + S23 * R23;
    const float c23 = S21 * R31 + S22 * R32 + S23 * R33;
    const float c33 = S31 * R31 + S32 * R32 + S33 * R33;

    const float x = (r1[0] * mx + r1[1] * my + r1[2] * mz + r1[3]) / depth;
    const float y = (r2[0] * mx + r2[1] * my + r2[2] * mz + r2[3]) / depth;

    const float clip_left = (-0.15f * width - center_x) / focal_x;
    const float clip_right = (1.15f * width - center_x) / focal_x;
    const float clip_top = (-0.15f * height - center_y) / focal_y;
    const float clip_bottom = (1.15f * height - center_y) / focal_y;
    const float xc = fmaxf(clip_left, fminf(x, clip_right));
    const float yc = fmaxf(clip_top, fminf(y, clip_bottom));
    const float j11 = focal_x / depth;
    const float j13 = -j11 * xc;
    const float j22 = focal_y / depth;
    const float j23 = -j22 * yc;
    const float a1 = j11 * r1[0] + j13 * r3[0], a2 = j11 * r1[1] + j13 * r3[1], a3 = j11 * r1[2] + j13 * r3[2];
    const float b1 = j22 * r2[0] + j23 * r3[0], b2 = j22 * r2[1] + j23 * r3[1], b3 = j22 * r2[2] + j23 * r3[2];
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
    const float kernel = (proper_antialiasing ? 0.1f : 0.3f) * ss * ss;
    cxx = cxx + kernel;
    cyy = cyy + kernel;
    const float det = cxx * cyy - cxy * cxy;
    if (det < 1e-6f) active = 0;
    o->conic[0] = cyy / det;
    o->conic[1] = -cxy / det;
    o->conic[2] = cxx / det;
    if (proper_antialiasing) {
        opacity = opacity * sqrtf(fmaxf(det_raw / det, 0.0f));
        if (opacity < ORC_MIN_ALPHA) active = 0;
    }
    o->mean2d[0] = x * focal_x + center_x;
    o->mean2d[1] = y * focal_y + center_y;
    const float pt = orc_logf(opacity * 255.0f);
    const float cutoff = 2.0f * pt;
    const float ex = sqrtf(cxx * cutoff);
    const float ey = sqrtf(cyy * cutoff);
    o->bounds[0] = o->mean2d[0] - ex;
    o->bounds[1] = o->mean2d[0] + ex;
    o->bounds[2] = o->mean2d[1] - ey;
    o->bounds[3] = o->mean2d[1] + ey;
    /* off screen, or undefined bounds */
    if (!(o->bounds[0] < width) || !(o->bounds[1] >= 0.0f) || !(o->bounds[2] < height) || !(o->bounds[3] >= 0.0f))
        active = 0;
    o->active = active;
    o->depth = depth;
    o->opacity = opacity;
    o->determinant = det;
    o->power_threshold = pt;
    o->cov2d[0] = cxx;
    o->cov2d[1] = cxy;
    o->cov2d[2] = cyy;
}

/* ------------------------------------------------------------------------------------------ */
/* accumulation                                                                               */
/* ------------------------------------------------------------------------------------------ */
typedef struct fbcam {
    float W, H, fx, fy, cx, cy, ss;
    uint32_t Wi, Hi, ssi;
} fbcam;

/* render.cu:52 (ss = max(1, raw)) and the subpixel scaling of accumulate_scatter.cu:180-189 */
static fbcam make_fbcam(const orc_camera* cam)
{
    fbcam f;
    f.ssi = cam->supersampling_factor < 1 ? 1u : (uint32_t)cam->supersampling_factor;
    f.ss = (float)f.ssi;
    f.Wi = cam->width * f.ssi;
    f.Hi = cam->height * f.ssi;
    f.W = (float)f.Wi;
    f.H = (float)f.Hi;
    f.fx = cam->focal_x * f.ss;
    f.fy = cam->focal_y * f.ss;
    f.cx = cam->center_x * f.ss;
    f.cy = cam->center_y * f.ss;
    return f;
}

static void project_fb(const orc_scene* s, const orc_camera* cam, const fbcam* f, uint32_t g, orc_proj* p)
{
    orc_project(s, g, cam->w2c, f->W, f->H, f->fx, f->fy, f->cx, f->cy, cam->near_plane, cam->far_plane,
                cam->proper_antialiasing, f->ss, p);
}

static inline void fb_min(uint64_t* p, uint64_t v, int atomic)
{
    if (!atomic) {
        if (v < *p) *p = v;
        return;
    }
    uint64_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

/* scale_conic_for_exp2 (gs_utils.h:21-24): alpha = o * 2^(dx*(a*dx + b*dy) + c*dy*dy) */
static inline void conic_exp2(const float conic[3], float ce[3])
{
    ce[0] = -0.5f * ORC_LOG2E * conic[0];
    ce[1] = -ORC_LOG2E * conic[1];
    ce[2] = -0.5f * ORC_LOG2E * conic[2];
}

/* one fragment's alpha (scatter_utils.cuh:37-40), in the fixed fma order shared with the kernels */
static inline float st_alpha(const float ce[3], float opacity, float dx, float dy)
{
    const float t = fmaf(ce[0], dx, ce[1] * dy);
    const float e = fmaf(dx, t, ce[2] * dy * dy);
    return opacity * orc_exp2f(fminf(e, 0.0f));
}

static inline float st_threshold(uint32_t frame, uint32_t g, uint32_t px, uint32_t py,
                                 const uint8_t* bn, uint32_t bn_rot)
{
    return bn ? orc_blue_noise(bn, bn_rot, g, px, py) : orc_rand(frame, g, px, py);
}

/* pixel AABB of a projected Gaussian (accumulate_scatter.cu:62-70) */
static inline void st_aabb(const orc_proj* p, const fbcam* f, uint32_t* x0, uint32_t* x1, uint32_t* y0, uint32_t* y1)
{
    *x0 = (uint32_t)floorf(fmaxf(p->bounds[0], 0.0f));
    *x1 = (uint32_t)ceilf(fminf(p->bounds[1], f->W));
    *y0 = (uint32_t)floorf(fmaxf(p->bounds[2], 0.0f));
    *y1 = (uint32_t)ceilf(fminf(p->bounds[3], f->H));
}

static inline uint64_t fragment_key(float depth, uint32_t g)
{
    return ((uint64_t)f2u(depth) << 32) | (uint64_t)g;
}

/* stochastic transparency of one Gaussian: accumulate_st_cu + scatter_fragments
 * (accumulate_scatter.cu:45-77, scatter_utils.cuh:14-95); fragment order is irrelevant */
static void st_gaussian(uint32_t g, const orc_proj* p, const fbcam* f, uint32_t frame,
                        const uint8_t* bn, uint32_t bn_rot, const uint32_t crop[4], uint64_t* fb,
                        int atomic, uint64_t* tested, uint64_t* accepted)
{
    uint32_t x0, x1, y0, y1;
    st_aabb(p, f, &x0, &x1, &y0, &y1);
    const float msx = p->mean2d[0] - 0.5f;
    const float msy = p->mean2d[1] - 0.5f;
    float ce[3];
    conic_exp2(p->conic, ce);
    const uint64_t frag = fragment_key(p->depth, g);
    const uint32_t xs = x0 > crop[0] ? x0 : crop[0], xe = x1 < crop[2] ? x1 : crop[2];
    const uint32_t ys = y0 > crop[1] ? y0 : crop[1], ye = y1 < crop[3] ? y1 : crop[3];
    const uint32_t cw = crop[2] - crop[0];
    for (uint32_t py = ys; py < ye; py++) {
        for (uint32_t px = xs; px < xe; px++) {
            (*tested)++;
            const float dx = (float)px - msx;
            const float dy = (float)py - msy;
            const float alpha = st_alpha(ce, p->opacity, dx, dy);
            if (alpha < ORC_MIN_ALPHA) continue;
            if (st_threshold(frame, g, px, py, bn, bn_rot) >= alpha) continue;
            (*accepted)++;
            fb_min(&fb[(size_t)(py - crop[1]) * cw + (px - crop[0])], frag, atomic);
        }
    }
}

static inline uint32_t gps_max_points(const orc_camera* cam, const fbcam* f)
{
    if (cam->max_gps_points) return cam->max_gps_points;
    return f2u_rn_sat(f->W * f->H / (2.0f * f->ss * f->ss)); /* accumulate_scatter.cu:128 */
}

/* GPS point count (accumulate_scatter.cu:120-129) */
static uint32_t gps_count(uint32_t g, const orc_proj* p, uint32_t frame, uint32_t max_points,
                          float* importance_out, uint32_t* true_out, float* dilog_out)
{
    const float od = orc_dilog(p->opacity);
    const float importance = ORC_TWO_PI * sqrtf(p->determinant) * od;
    uint32_t seed[2];
    orc_make_seed(g, frame, seed);
    const uint32_t ct = orc_poisson(seed, importance);
    if (importance_out) *importance_out = importance;
    if (true_out) *true_out = ct;
    if (dilog_out) *dilog_out = od;
    return ct < max_points ? ct : max_points;
}

/* Gaussian point splatting of one Gaussian: accumulate_gps_cu + scatter_points
 * (accumulate_scatter.cu:103-146, scatter_utils.cuh:98-183), per-point counter seeds */
static void gps_gaussian(uint32_t g, const orc_proj* p, const fbcam* f, uint32_t frame, uint32_t max_points,
                         const uint32_t crop[4], uint64_t* fb, int atomic, uint64_t* tested, uint64_t* accepted)
{
    /* a point's pixel centre lies inside the power_threshold ellipse, so its pixel is within
     * half a pixel of the bounds: Gaussians whose bounds miss the crop by > 1 px add nothing */
    if (p->bounds[1] + 1.0f < (float)crop[0] || p->bounds[0] - 1.0f >= (float)crop[2] ||
        p->bounds[3] + 1.0f < (float)crop[1] || p->bounds[2] - 1.0f >= (float)crop[3])
        return;
    float od;
    const uint32_t count = gps_count(g, p, frame, max_points, NULL, NULL, &od);
    if (count == 0) return;
    const float c0 = sqrtf(p->cov2d[0]);
    const float c1 = p->cov2d[1] / c0;
    const float c2 = sqrtf(p->cov2d[2] - c1 * c1);
    const float o_rcp = 1.0f / fmaxf(p->opacity, 1.0e-37f);
    const uint64_t frag = fragment_key(p->depth, g);
    const uint32_t cw = crop[2] - crop[0];
    uint32_t base[2];
    orc_make_seed(g, frame, base);
    for (uint32_t k = 0; k < count; k++) {
        (*tested)++;
        uint32_t st[2] = { base[0] ^ ((k + 1u) * 0x9E3779B9u), base[1] };
        float u[2], xy[2];
        orc_pcg2d(st, u);
        orc_corrected_box_muller(u[0], u[1], o_rcp, od, xy);
        const float sx = p->mean2d[0] + c0 * xy[0];
        const float sy = p->mean2d[1] + (c1 * xy[0] + c2 * xy[1]);
        if (!(sx < f->W && sx >= 0.0f && sy < f->H && sy >= 0.0f)) continue;
        const float dcx = truncf(sx) + 0.5f - p->mean2d[0];
        const float dcy = truncf(sy) + 0.5f - p->mean2d[1];
        const float power = 0.5f * (p->conic[0] * dcx * dcx + p->conic[2] * dcy * dcy) + p->conic[1] * dcx * dcy;
        if (power > p->power_threshold) continue;
        const uint32_t px = (uint32_t)sx, py = (uint32_t)sy;
        if (px < crop[0] || px >= crop[2] || py < crop[1] || py >= crop[3]) continue;
        (*accepted)++;
        fb_min(&fb[(size_t)(py - crop[1]) * cw + (px - crop[0])], frag, atomic);
    }
}

void orc_accumulate(const orc_scene* s, const orc_camera* cam, uint32_t frame_idx,
                    const uint8_t* bn, uint32_t bn_rot, const uint32_t crop[4], uint64_t* fb,
                    orc_stats* stats, int n_threads)
{
    const fbcam f = make_fbcam(cam);
    const uint32_t max_points = gps_max_points(cam, &f);
    uint64_t n_z = 0, n_p = 0, n_t = 0, n_a = 0;
    const int atomic = n_threads > 1;
    const int64_t n = (int64_t)s->n_gaussians;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) num_threads(n_threads > 1 ? n_threads : 1) \
    reduction(+ : n_z, n_p, n_t, n_a)
#endif
    for (int64_t gi = 0; gi < n; gi++) {
        const uint32_t g = (uint32_t)gi;
        orc_proj p;
        project_fb(s, cam, &f, g, &p);
        n_z += (uint64_t)p.in_depth_range;
        if (!p.active) continue;
        n_p++;
        if (cam->use_gps) gps_gaussian(g, &p, &f, frame_idx, max_points, crop, fb, atomic, &n_t, &n_a);
        else st_gaussian(g, &p, &f, frame_idx, bn, bn_rot, crop, fb, atomic, &n_t, &n_a);
    }
    if (stats) {
        stats->n_in_depth_range = n_z;
        stats->n_projected = n_p;
        stats->n_tested = n_t;
        stats->n_accepted = n_a;
    }
}

void orc_accumulate_st_bruteforce(const orc_scene* s, const orc_camera* cam, uint32_t frame_idx,
                                  const uint8_t* bn, uint32_t bn_rot, const uint32_t crop[4], uint64_t* fb)
{
    const fbcam f = make_fbcam(cam);
    const uint32_t n = s->n_gaussians;
    orc_proj* proj = (orc_proj*)malloc(sizeof(orc_proj) * (n ? n : 1));
    for (uint32_t g = 0; g < n; g++) project_fb(s, cam, &f, g, &proj[g]);
    const uint32_t cw = crop[2] - crop[0];
    for (uint32_t py = crop[1]; py < crop[3]; py++) {
        for (uint32_t px = crop[0]; px < crop[2]; px++) {
            uint64_t best = ~0ull;
            for (uint32_t g = 0; g < n; g++) {
                const orc_proj* p = &proj[g];
                if (!p->active) continue;
                float ce[3];
                conic_exp2(p->conic, ce);
                const float dx = (float)px - (p->mean2d[0] - 0.5f);
                const float dy = (float)py - (p->mean2d[1] - 0.5f);
                const float alpha = st_alpha(ce, p->opacity, dx, dy);
                if (alpha < ORC_MIN_ALPHA) continue;
                if (st_threshold(frame_idx, g, px, py, bn, bn_rot) >= alpha) continue;
                const uint64_t frag = fragment_key(p->depth, g);
                if (frag < best) best = frag;
            }
            fb[(size_t)(py - crop[1]) * cw + (px - crop[0])] = best;
        }
    }
    free(proj);
}

void orc_gps_counts(const orc_scene* s, const orc_camera* cam, uint32_t frame_idx,
                    uint32_t* counts_capped, uint32_t* counts_true, float* importance)
{
    const fbcam f = make_fbcam(cam);
    const uint32_t max_points = gps_max_points(cam, &f);
    for (uint32_t g = 0; g < s->n_gaussians; g++) {
        orc_proj p;
        project_fb(s, cam, &f, g, &p);
        if (!p.active) {
            counts_capped[g] = 0;
            if (counts_true) counts_true[g] = 0;
            if (importance) importance[g] = 0.0f;
            continue;
        }
        float imp;
        uint32_t ct;
        counts_capped[g] = gps_count(g, &p, frame_idx, max_points, &imp, &ct, NULL);
        if (counts_true) counts_true[g] = ct;
        if (importance) importance[g] = imp;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* sample (resolve.cu:27-144) and eval_sh (gs_utils.h:197-243)                                */
/* ------------------------------------------------------------------------------------------ */
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

void orc_eval_sh(const float* c, const float d[3], uint32_t n_bases, float out[3])
{
    const float x = d[0], y = d[1], z = d[2];
    for (int ch = 0; ch < 3; ch++) {
        float col = 0.5f + SH_C0 * c[ch];
        if (n_bases > 1) {
            col = col - SH_C1 * y * c[3 + ch] + SH_C1 * z * c[6 + ch] - SH_C1 * x * c[9 + ch];
            if (n_bases > 4) {
                const float xx = x * x, yy = y * y, zz = z * z;
                const float xy = x * y, xz = x * z, yz = y * z;
                col = col + SH_C2A * xy * c[12 + ch] - SH_C2A * yz * c[15 + ch] + (SH_C2B * zz - SH_C2C) * c[18 + ch]
                      - SH_C2A * xz * c[21 + ch] + SH_C2D * (xx - yy) * c[24 + ch];
                if (n_bases > 9) {
                    col = col + y * (SH_C3A * yy - SH_C3B * xx) * c[27 + ch] + SH_C3C * xy * z * c[30 + ch]
                          + y * (SH_C3D - SH_C3E * zz) * c[33 + ch] + z * (SH_C3F * zz - SH_C3G) * c[36 + ch]
                          + x * (SH_C3D - SH_C3E * zz) * c[39 + ch] + SH_C3H * z * (xx - yy) * c[42 + ch]
                          + x * (SH_C3B * yy - SH_C3A * xx) * c[45 + ch];
                }
            }
        }
        out[ch] = fmaxf(col, 0.0f);
    }
}

/* colour of the Gaussian seen from the camera: normalize(mean - cam) (resolve.cu:72-75) */
static void shade(const orc_scene* s, const orc_camera* cam, uint32_t g, float out[3])
{
    const float vx = s->means[3 * g + 0] - cam->cam_position[0];
    const float vy = s->means[3 * g + 1] - cam->cam_position[1];
    const float vz = s->means[3 * g + 2] - cam->cam_position[2];
    const float inv = 1.0f / sqrtf(vx * vx + vy * vy + vz * vz);
    const float d[3] = { vx * inv, vy * inv, vz * inv };
    orc_eval_sh(s->sh + (size_t)g * s->n_sh_bases * 3u, d, s->n_sh_bases, out);
}

#define K_BUFFER 4
#define MAX_ALPHA 0.99f

static inline float reuse_alpha(const float ce[3], float opacity, float dx, float dy)
{
    const float a = st_alpha(ce, opacity, dx, dy);
    return a < ORC_MIN_ALPHA ? 0.0f : a;
}

void orc_sample(const orc_scene* s, const orc_camera* cam, const uint64_t* fb,
                float* samples, float* reuse_weights, int32_t* k_prim, float* k_occ)
{
    const fbcam f = make_fbcam(cam);
    const uint32_t P = f.Wi * f.Hi;
    for (uint32_t pix = 0; pix < P; pix++) {
        const uint64_t data = fb[pix];
        const uint32_t dk = (uint32_t)(data >> 32);
        const int has = dk != 0xffffffffu;
        const uint32_t g = has ? (uint32_t)(data & 0xffffffffu) : 0u;
        float smp[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
        if (has) {
            shade(s, cam, g, smp);
            smp[3] = u2f(dk);
        }
        memcpy(&samples[4 * (size_t)pix], smp, sizeof smp);
        if (!cam->spatial_reuse) continue;

        orc_proj p;
        int projected = 0;
        if (has) {
            project_fb(s, cam, &f, g, &p);
            projected = p.active;
        }
        float alpha = 1.0f;
        float na[4] = { 1.0f, 1.0f, 1.0f, 1.0f };
        float sdepth = INFINITY;
        if (projected) {
            const float dx = (float)(pix % f.Wi) - (p.mean2d[0] - 0.5f);
            const float dy = (float)(pix / f.Wi) - (p.mean2d[1] - 0.5f);
            float ce[3];
            conic_exp2(p.conic, ce);
            alpha = fmaxf(reuse_alpha(ce, p.opacity, dx, dy), ORC_MIN_ALPHA);
            na[0] = reuse_alpha(ce, p.opacity, dx - 1.0f, dy);
            na[1] = reuse_alpha(ce, p.opacity, dx + 1.0f, dy);
            na[2] = reuse_alpha(ce, p.opacity, dx, dy - 1.0f);
            na[3] = reuse_alpha(ce, p.opacity, dx, dy + 1.0f);
            sdepth = smp[3];
        }
        float w[4];
        for (int c = 0; c < 4; c++) w[c] = (has && !projected) ? 0.0f : na[c] / alpha;

        int unique = projected;
        uint32_t prim[K_BUFFER][2];
        float occ[K_BUFFER][4];
        for (uint32_t k = 0; k < K_BUFFER; k++) {
            const size_t e = (size_t)k * P + pix;
            prim[k][0] = (uint32_t)k_prim[2 * e];
            prim[k][1] = (uint32_t)k_prim[2 * e + 1];
            memcpy(occ[k], &k_occ[4 * e], sizeof occ[k]);
            if (prim[k][0] == g) unique = 0;
            if (u2f(prim[k][1]) < sdepth)
                for (int c = 0; c < 4; c++) w[c] = w[c] * occ[k][c];
        }
        if (unique) {
            for (uint32_t k = K_BUFFER - 1; k > 0; k--) {
                const size_t e = (size_t)k * P + pix;
                k_prim[2 * e] = (int32_t)prim[k - 1][0];
                k_prim[2 * e + 1] = (int32_t)prim[k - 1][1];
                memcpy(&k_occ[4 * e], occ[k - 1], sizeof occ[k - 1]);
            }
            k_prim[2 * (size_t)pix] = (int32_t)g;
            k_prim[2 * (size_t)pix + 1] = (int32_t)dk;
            const float den = 1.0f - fminf(alpha, MAX_ALPHA);
            for (int c = 0; c < 4; c++) k_occ[4 * (size_t)pix + c] = (1.0f - fminf(na[c], MAX_ALPHA)) / den;
        }
        memcpy(&reuse_weights[4 * (size_t)pix], w, sizeof w);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* resolve (resolve.cu:146-349)                                                               */
/* ------------------------------------------------------------------------------------------ */
void orc_catmull_rom_weights(float t, float out[4])
{
    const float t2 = t * t;
    out[0] = t * (-0.5f + t * (1.0f - 0.5f * t));
    out[1] = 1.0f + t2 * (-2.5f + 1.5f * t);
    out[2] = t * (0.5f + t * (2.0f - 1.5f * t));
    out[3] = t2 * (-0.5f + 0.5f * t);
}

void orc_rgb_to_ycocg(const float c[3], float out[3])
{
    out[0] = 0.25f * c[0] + 0.5f * c[1] + 0.25f * c[2];
    out[1] = 0.5f * c[0] - 0.5f * c[2];
    out[2] = -0.25f * c[0] + 0.5f * c[1] - 0.25f * c[2];
}

static inline float lerpf(float a, float b, float t) { return a + t * (b - a); }

int orc_resolve(const orc_camera* cam, const float* samples, const float* rw,
                const float* h_image, const float* h_depth, const float* h_alpha, const float* h_weight,
                const float* m, int32_t temporal_mode,
                float* image, float* depth_map, float* alpha_map, float* weight_map)
{
    if (temporal_mode < 0 || temporal_mode > 2) return -1;
    const fbcam f = make_fbcam(cam);
    const uint32_t W = cam->width, H = cam->height, ss = f.ssi;
    const uint32_t FW = f.Wi, FH = f.Hi;
    const float nf = 1.0f / (float)(ss * ss);
    for (uint32_t pix = 0; pix < W * H; pix++) {
        const uint32_t bx = (pix % W) * ss, by = (pix / W) * ss;
        float cs[3] = { 0.0f, 0.0f, 0.0f };
        float ds = 0.0f, as = 0.0f, ws = 0.0f, dout = 0.0f;
        uint32_t vc = 0;
        for (uint32_t oy = 0; oy < ss; oy++) {
            for (uint32_t ox = 0; ox < ss; ox++) {
                const uint32_t fx = bx + ox, fy = by + oy;
                const uint32_t fi = fy * FW + fx;
                const float* sm = &samples[4 * (size_t)fi];
                if (sm[3] > 0.0f) {
                    dout = dout + sm[3];
                    vc++;
                }
                float c0 = sm[0], c1 = sm[1], c2 = sm[2];
                float dsum = sm[3];
                float asum = sm[3] > 0.0f ? 1.0f : 0.0f;
                float wsum = 1.0f;
                if (cam->spatial_reuse) {
                    uint32_t nidx[4];
                    float nw[4];
                    int nn = 0;
                    if (fx > 0) { nidx[nn] = fi - 1; nw[nn++] = rw[4 * (size_t)(fi - 1) + 1]; }
                    if (fx < FW - 1) { nidx[nn] = fi + 1; nw[nn++] = rw[4 * (size_t)(fi + 1) + 0]; }
                    if (fy > 0) { nidx[nn] = fi - FW; nw[nn++] = rw[4 * (size_t)(fi - FW) + 3]; }
                    if (fy < FH - 1) { nidx[nn] = fi + FW; nw[nn++] = rw[4 * (size_t)(fi + FW) + 2]; }
                    for (int k = 0; k < nn; k++) {
                        const float* ns = &samples[4 * (size_t)nidx[k]];
                        const float wk = nw[k];
                        c0 = c0 + wk * ns[0];
                        c1 = c1 + wk * ns[1];
                        c2 = c2 + wk * ns[2];
                        dsum = dsum + wk * ns[3];
                        asum = asum + (ns[3] > 0.0f ? wk : 0.0f);
                        wsum = wsum + wk;
                    }
                }
                const float r = 1.0f / wsum;
                cs[0] = cs[0] + c0 * r;
                cs[1] = cs[1] + c1 * r;
                cs[2] = cs[2] + c2 * r;
                ds = ds + dsum * r;
                as = as + asum * r;
                ws = ws + wsum;
            }
        }
        cs[0] = cs[0] * nf;
        cs[1] = cs[1] * nf;
        cs[2] = cs[2] * nf;
        ds = ds * nf;
        as = as * nf;
        ws = ws * nf;

        float col[3] = { cs[0], cs[1], cs[2] };
        float dep = ds, alp = as, wm = ws;
        if (temporal_mode == 1) {
            const float hm = h_weight[pix];
            wm = hm + ws;
            const float hf = hm / wm;
            for (int c = 0; c < 3; c++) col[c] = lerpf(cs[c], h_image[3 * (size_t)pix + c], hf);
            dep = lerpf(ds, h_depth[pix], hf);
            alp = lerpf(as, h_alpha[pix], hf);
        } else if (temporal_mode == 2) {
            const uint32_t x = pix % W, y = pix / W;
            const float z = vc > 0 ? dout / (float)vc : cam->far_plane;
            const float posx = (float)x + 0.5f, posy = (float)y + 0.5f;
            const float vpx = (posx - cam->center_x) / cam->focal_x * z;
            const float vpy = (posy - cam->center_y) / cam->focal_y * z;
            const float vpz = z;
            const float pz = m[8] * vpx + m[9] * vpy + m[10] * vpz + m[11];
            if (pz > 0.0f) {
                const float ppx = cam->focal_x * (m[0] * vpx + m[1] * vpy + m[2] * vpz + m[3]) / pz + cam->center_x;
                const float ppy = cam->focal_y * (m[4] * vpx + m[5] * vpy + m[6] * vpz + m[7]) / pz + cam->center_y;
                const int inside = ppx >= 2.0f && ppx <= (float)W - 2.0f && ppy >= 2.0f && ppy <= (float)H - 2.0f;
                if (inside) {
                    const float tx = ppx - 0.5f, ty = ppy - 0.5f;
                    const float tbx = floorf(tx), tby = floorf(ty);
                    float wx[4], wy[4];
                    orc_catmull_rom_weights(tx - tbx, wx);
                    orc_catmull_rom_weights(ty - tby, wy);
                    const uint32_t hbx = (uint32_t)tbx - 1u, hby = (uint32_t)tby - 1u;
                    float ch[3] = { 0.0f, 0.0f, 0.0f };
                    float dh = 0.0f, ah = 0.0f;
                    for (uint32_t j = 0; j < 4; j++) {
                        for (uint32_t i = 0; i < 4; i++) {
                            const size_t hi = (size_t)(hby + j) * W + hbx + i;
                            const float w = wx[i] * wy[j];
                            ch[0] = ch[0] + w * h_image[3 * hi + 0];
                            ch[1] = ch[1] + w * h_image[3 * hi + 1];
                            ch[2] = ch[2] + w * h_image[3 * hi + 2];
                            dh = dh + w * h_depth[hi];
                            ah = ah + w * h_alpha[hi];
                        }
                    }
                    float mean[3] = { 0.0f, 0.0f, 0.0f }, msq[3] = { 0.0f, 0.0f, 0.0f };
                    for (int oy = -1; oy <= 1; oy++) {
                        for (int ox = -1; ox <= 1; ox++) {
                            int nxi = (int)x + ox, nyi = (int)y + oy;
                            nxi = nxi < 0 ? 0 : (nxi > (int)W - 1 ? (int)W - 1 : nxi);
                            nyi = nyi < 0 ? 0 : (nyi > (int)H - 1 ? (int)H - 1 : nyi);
                            const uint32_t nx = (uint32_t)nxi, ny = (uint32_t)nyi;
                            float nc[3] = { 0.0f, 0.0f, 0.0f };
                            for (uint32_t sy = 0; sy < ss; sy++)
                                for (uint32_t sx = 0; sx < ss; sx++) {
                                    const float* sm = &samples[4 * ((size_t)(ny * ss + sy) * FW + nx * ss + sx)];
                                    nc[0] = nc[0] + sm[0];
                                    nc[1] = nc[1] + sm[1];
                                    nc[2] = nc[2] + sm[2];
                                }
                            const float ncs[3] = { nc[0] * nf, nc[1] * nf, nc[2] * nf };
                            float ycc[3];
                            orc_rgb_to_ycocg(ncs, ycc);
                            for (int c = 0; c < 3; c++) {
                                mean[c] = mean[c] + ycc[c];
                                msq[c] = msq[c] + ycc[c] * ycc[c];
                            }
                        }
                    }
                    float sd[3];
                    for (int c = 0; c < 3; c++) {
                        mean[c] = mean[c] / 9.0f;
                        msq[c] = msq[c] / 9.0f;
                        sd[c] = sqrtf(fmaxf(msq[c] - mean[c] * mean[c], 0.0f));
                    }
                    const float dvx = posx - ppx, dvy = posy - ppy;
                    const float velocity = sqrtf(dvx * dvx + dvy * dvy);
                    const float gamma = lerpf(4.0f, 1.0f, fminf(velocity / 4.0f, 1.0f));
                    const float minext[3] = { 0.01f, 0.005f, 0.005f };
                    float ych[3];
                    orc_rgb_to_ycocg(ch, ych);
                    float dist = 0.0f;
                    float disc[3];
                    for (int c = 0; c < 3; c++) {
                        const float ext = fmaxf(gamma * sd[c], minext[c]);
                        disc[c] = fabsf(ych[c] - mean[c]) / ext;
                    }
                    dist = fmaxf(disc[0], fmaxf(disc[1], disc[2]));
                    const float discount = dist > 1.0f ? 0.7f / dist : 0.7f;
                    for (int c = 0; c < 3; c++) col[c] = lerpf(cs[c], ch[c], discount);
                    dep = lerpf(ds, dh, discount);
                    alp = lerpf(as, ah, discount);
                    wm = ws / (1.0f - discount);
                }
            }
        }
        image[3 * (size_t)pix + 0] = col[0];
        image[3 * (size_t)pix + 1] = col[1];
        image[3 * (size_t)pix + 2] = col[2];
        depth_map[pix] = dep;
        alpha_map[pix] = alp;
        weight_map[pix] = wm;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* full render (render.cu:16-158)                                                             */
/* ------------------------------------------------------------------------------------------ */
int orc_render(const orc_scene* s, const orc_camera* cam, uint32_t frame_idx,
               const uint8_t* blue_noise, uint32_t blue_noise_numel,
               int32_t* k_prim, float* k_occ,
               const float* h_image, const float* h_depth, const float* h_alpha, const float* h_weight,
               const float* v2pv, int32_t temporal_mode,
               float* image, float* depth, float* alpha, float* weight_mass,
               uint64_t* fb_out, orc_stats* stats, int n_threads)
{
    if (temporal_mode < 0 || temporal_mode > 2) return -1;
    const fbcam f = make_fbcam(cam);
    const size_t P = (size_t)f.Wi * f.Hi;
    uint64_t* fb = fb_out ? fb_out : (uint64_t*)malloc(P * 8u + 8u);
    float* samples = (float*)malloc(P * 16u + 16u);
    float* rw = cam->spatial_reuse ? (float*)malloc(P * 16u + 16u) : NULL;
    memset(fb, 0xff, P * 8u);
    const uint8_t* slice = NULL;
    uint32_t rot = 0;
    if (blue_noise_numel > 0 && !cam->use_gps) {
        slice = blue_noise + (size_t)(frame_idx % 64u) * 128u * 128u; /* render.cu:77-82 */
        rot = (frame_idx / 64u) * 0x9E3779B9u;
    }
    const uint32_t crop[4] = { 0, 0, f.Wi, f.Hi };
    orc_accumulate(s, cam, frame_idx, slice, rot, crop, fb, stats, n_threads);
    orc_sample(s, cam, fb, samples, rw, k_prim, k_occ);
    const int rc = orc_resolve(cam, samples, rw, h_image, h_depth, h_alpha, h_weight, v2pv, temporal_mode,
                               image, depth, alpha, weight_mass);
    free(samples);
    free(rw);
    if (!fb_out) free(fb);
    return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* FP64 sorted front-to-back blend: the expectation of one ST sample                          */
/* ------------------------------------------------------------------------------------------ */
typedef struct blend_frag {
    uint64_t key;
    double alpha;
    double color[3];
    double depth;
} blend_frag;

static int cmp_frag(const void* a, const void* b)
{
    const uint64_t ka = ((const blend_frag*)a)->key, kb = ((const blend_frag*)b)->key;
    return ka < kb ? -1 : (ka > kb ? 1 : 0);
}

void orc_blend_fp64(const orc_scene* s, const orc_camera* cam, double* image, double* depth, double* alpha)
{
    const fbcam f = make_fbcam(cam);
    const uint32_t n = s->n_gaussians;
    orc_proj* proj = (orc_proj*)malloc(sizeof(orc_proj) * (n ? n : 1));
    float (*col)[3] = (float (*)[3])malloc(sizeof(float) * 3 * (n ? n : 1));
    uint32_t (*box)[4] = (uint32_t (*)[4])malloc(sizeof(uint32_t) * 4 * (n ? n : 1));
    for (uint32_t g = 0; g < n; g++) {
        project_fb(s, cam, &f, g, &proj[g]);
        if (proj[g].active) {
            shade(s, cam, g, col[g]);
            st_aabb(&proj[g], &f, &box[g][0], &box[g][1], &box[g][2], &box[g][3]);
        }
    }
    blend_frag* list = (blend_frag*)malloc(sizeof(blend_frag) * (n ? n : 1));
    const uint32_t W = cam->width, H = cam->height, ss = f.ssi;
    memset(image, 0, sizeof(double) * 3 * W * H);
    memset(depth, 0, sizeof(double) * W * H);
    memset(alpha, 0, sizeof(double) * W * H);
    const double nf = 1.0 / (double)(ss * ss);
    for (uint32_t py = 0; py < f.Hi; py++) {
        for (uint32_t px = 0; px < f.Wi; px++) {
            uint32_t cnt = 0;
            for (uint32_t g = 0; g < n; g++) {
                const orc_proj* p = &proj[g];
                if (!p->active) continue;
                if (px < box[g][0] || px >= box[g][1] || py < box[g][2] || py >= box[g][3]) continue;
                float ce[3];
                conic_exp2(p->conic, ce);
                const float a = st_alpha(ce, p->opacity, (float)px - (p->mean2d[0] - 0.5f),
                                         (float)py - (p->mean2d[1] - 0.5f));
                if (a < ORC_MIN_ALPHA) continue;
                list[cnt].key = fragment_key(p->depth, g);
                list[cnt].alpha = a > 1.0f ? 1.0 : (double)a;
                list[cnt].color[0] = col[g][0];
                list[cnt].color[1] = col[g][1];
                list[cnt].color[2] = col[g][2];
                list[cnt].depth = p->depth;
                cnt++;
            }
            qsort(list, cnt, sizeof(blend_frag), cmp_frag);
            double T = 1.0, C[3] = { 0.0, 0.0, 0.0 }, D = 0.0;
            for (uint32_t k = 0; k < cnt; k++) {
                const double w = list[k].alpha * T;
                C[0] += w * list[k].color[0];
                C[1] += w * list[k].color[1];
                C[2] += w * list[k].color[2];
                D += w * list[k].depth;
                T *= 1.0 - list[k].alpha;
            }
            const size_t o = (size_t)(py / ss) * W + px / ss;
            image[3 * o + 0] += C[0] * nf;
            image[3 * o + 1] += C[1] * nf;
            image[3 * o + 2] += C[2] * nf;
            depth[o] += D * nf;
            alpha[o] += (1.0 - T) * nf;
        }
    }
    free(list);
    free(box);
    free(col);
    free(proj);
}
