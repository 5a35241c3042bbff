/*
 * Driver of the CPU oracle for the AddressSanitizer / UndefinedBehaviorSanitizer build
 * (`make -C oracle sanitize`). TEST INFRASTRUCTURE ONLY, like the oracle itself.
 *
 * Runs every entry point of sgs_oracle.h over small seeded scenes at odd image sizes
 * (framebuffer edges, supersampling 1-3, crops touching the borders, GPS with and without
 * a point cap, spatial reuse over several frames, STATIC and REPROJECT histories, blue
 * noise, one and several OpenMP threads), so out-of-bounds accesses and undefined
 * arithmetic anywhere in the oracle abort the run.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sgs_oracle.h"

static uint32_t lcg_state = 12345u;
static float frand(void)
{
    lcg_state = lcg_state * 1664525u + 1013904223u;
    return (float)(lcg_state >> 8) * (1.0f / 16777216.0f);
}
static float nrand(void)
{
    const float u1 = frand() + 1e-7f, u2 = frand();
    return sqrtf(-2.0f * logf(u1)) * cosf(6.2831853f * u2);
}

static void make_scene(uint32_t n, uint32_t bases, float log_scale, float* m, float* s, float* r, float* o, float* sh)
{
    for (uint32_t i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            m[3 * i + k] = 2.0f * frand() - 1.0f;
            s[3 * i + k] = log_scale + 0.7f * nrand();
        }
        float q[4], l = 0.0f;
        for (int k = 0; k < 4; k++) { q[k] = nrand(); l += q[k] * q[k]; }
        l = sqrtf(l);
        for (int k = 0; k < 4; k++) r[4 * i + k] = (i % 97 == 5) ? 0.0f : q[k] / l;  /* some zero quaternions */
        o[i] = 2.0f * nrand();
        for (uint32_t b = 0; b < 3 * bases; b++) sh[(size_t)i * 3 * bases + b] = b < 3 ? 0.5f + 0.2f * nrand() : 0.05f * nrand();
    }
}

static void look_at(float angle, float dist, orc_camera* c, uint32_t w, uint32_t h)
{
    const float ex = dist * sinf(angle), ey = 0.2f, ez = -dist * cosf(angle);
    float z[3] = { -ex, -ey, -ez };
    float zl = sqrtf(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    for (int k = 0; k < 3; k++) z[k] /= zl;
    /* x = (0,1,0) x z, y = z x x */
    float x[3] = { z[2], 0.0f, -z[0] };
    float xl = sqrtf(x[0] * x[0] + x[2] * x[2]);
    x[0] /= xl; x[2] /= xl;
    const float y[3] = { z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0] };
    const float* R[3] = { x, y, z };
    const float e[3] = { ex, ey, ez };
    memset(c, 0, sizeof *c);
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) c->w2c[4 * i + j] = R[i][j];
        c->w2c[4 * i + 3] = -(R[i][0] * e[0] + R[i][1] * e[1] + R[i][2] * e[2]);
    }
    c->w2c[15] = 1.0f;
    c->cam_position[0] = ex; c->cam_position[1] = ey; c->cam_position[2] = ez;
    c->width = w; c->height = h;
    c->focal_x = c->focal_y = 0.5f * (float)h / tanf(0.5235988f);
    c->center_x = 0.5f * (float)w; c->center_y = 0.5f * (float)h;
    c->near_plane = 0.2f; c->far_plane = 1000.0f;
}

static int run_case(uint32_t n, uint32_t bases, uint32_t w, uint32_t h, int ss, int gps, int reuse, int aa, int threads)
{
    float* m = malloc(sizeof(float) * 3 * n);
    float* s = malloc(sizeof(float) * 3 * n);
    float* r = malloc(sizeof(float) * 4 * n);
    float* o = malloc(sizeof(float) * n);
    float* sh = malloc(sizeof(float) * 3 * bases * n);
    make_scene(n, bases, -3.4f, m, s, r, o, sh);
    const orc_scene sc = { m, s, r, o, sh, n, bases };
    orc_camera cam, cam2;
    look_at(0.3f, 3.0f, &cam, w, h);
    look_at(0.33f, 3.05f, &cam2, w, h);
    cam.supersampling_factor = cam2.supersampling_factor = ss;
    cam.use_gps = cam2.use_gps = gps;
    cam.spatial_reuse = cam2.spatial_reuse = reuse;
    cam.proper_antialiasing = cam2.proper_antialiasing = aa;
    cam.max_gps_points = cam2.max_gps_points = gps == 2 ? 16u : 0u;
    const uint32_t W = w, H = h, FW = w * (uint32_t)ss, FH = h * (uint32_t)ss, P = FW * FH;
    const uint32_t K = reuse ? 4u : 0u;
    int32_t* kp = malloc(sizeof(int32_t) * 2 * (K ? K : 1) * P);
    float* ko = malloc(sizeof(float) * 4 * (K ? K : 1) * P);
    for (size_t i = 0; i < 2u * (K ? K : 1) * P; i++) kp[i] = -1;
    memset(ko, 0, sizeof(float) * 4 * (K ? K : 1) * P);
    float* img[2]; float* dep[2]; float* alp[2]; float* wm[2];
    for (int k = 0; k < 2; k++) {
        img[k] = calloc((size_t)3 * W * H, sizeof(float));
        dep[k] = calloc((size_t)W * H, sizeof(float));
        alp[k] = calloc((size_t)W * H, sizeof(float));
        wm[k] = calloc((size_t)W * H, sizeof(float));
    }
    uint64_t* fb = malloc(sizeof(uint64_t) * P);
    uint8_t* bn = malloc(64u * 128u * 128u);
    for (uint32_t i = 0; i < 64u * 128u * 128u; i++) bn[i] = (uint8_t)(frand() * 256.0f);
    float v2p[16];
    /* previous view <- current view: w2c_prev * inverse(w2c_cur) for rigid transforms */
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 3; k++) acc += cam.w2c[4 * i + k] * cam2.w2c[4 * j + k];
            v2p[4 * i + j] = acc;
        }
        float t = cam.w2c[4 * i + 3];
        for (int j = 0; j < 3; j++) t -= v2p[4 * i + j] * cam2.w2c[4 * j + 3];
        v2p[4 * i + 3] = t;
    }
    v2p[12] = v2p[13] = v2p[14] = 0.0f; v2p[15] = 1.0f;
    orc_stats st;
    int rc = 0;
    int cur = 0;
    for (int f = 0; f < 4 && !rc; f++) {
        const int mode = f == 0 ? 0 : (f == 3 ? 2 : 1);
        const orc_camera* c = f == 3 ? &cam2 : &cam;
        const int prev = 1 - cur;
        rc |= orc_render(&sc, c, (uint32_t)(f * 70), f == 1 && !gps ? bn : NULL, f == 1 && !gps ? 64u * 128u * 128u : 0u,
                         K ? kp : NULL, K ? ko : NULL, img[prev], dep[prev], alp[prev], wm[prev],
                         mode == 2 ? v2p : NULL, mode, img[cur], dep[cur], alp[cur], wm[cur], fb, &st, threads);
        cur = prev;
    }
    /* crops touching every border, brute force, GPS counts, FP64 blend */
    const uint32_t crops[3][4] = { { 0, 0, FW, FH }, { 0, FH / 2, FW / 3 + 1, FH }, { FW - 3, 0, FW, 2 } };
    for (int k = 0; k < 3; k++) {
        const uint32_t cw = crops[k][2] - crops[k][0], ch = crops[k][3] - crops[k][1];
        uint64_t* cf = malloc(sizeof(uint64_t) * cw * ch);
        for (uint32_t i = 0; i < cw * ch; i++) cf[i] = ~0ull;
        orc_accumulate(&sc, &cam, 9u, NULL, 0u, crops[k], cf, &st, threads);
        if (!gps) orc_accumulate_st_bruteforce(&sc, &cam, 9u, bn, 7u, crops[k], cf);
        free(cf);
    }
    uint32_t* cc = malloc(sizeof(uint32_t) * n);
    uint32_t* ct = malloc(sizeof(uint32_t) * n);
    float* imp = malloc(sizeof(float) * n);
    orc_gps_counts(&sc, &cam, 3u, cc, ct, imp);
    if (!gps && w * h <= 400u) {
        double* bi = malloc(sizeof(double) * 3 * W * H);
        double* bd = malloc(sizeof(double) * W * H);
        double* ba = malloc(sizeof(double) * W * H);
        orc_blend_fp64(&sc, &cam, bi, bd, ba);
        free(bi); free(bd); free(ba);
    }
    free(cc); free(ct); free(imp);
    for (int k = 0; k < 2; k++) { free(img[k]); free(dep[k]); free(alp[k]); free(wm[k]); }
    free(fb); free(bn); free(kp); free(ko); free(m); free(s); free(r); free(o); free(sh);
    return rc;
}

int main(void)
{
    /* scalar primitives over their whole input ranges */
    for (int i = -2000; i <= 2000; i++) {
        const float x = (float)i * 0.0731f;
        volatile float a = orc_exp2f(x) + orc_expf(x) + orc_logf(fabsf(x) + 1e-30f);
        float sn, cs, out[4];
        orc_sincos_2pi((float)(i + 2000) / 4001.0f, &sn, &cs);
        orc_catmull_rom_weights((float)(i + 2000) / 4001.0f, out);
        a += orc_dilog((float)(i + 2000) / 4001.0f) + orc_inv_dilog((float)(i + 2000) / 4001.0f * 1.6449f);
        uint32_t seed[2] = { (uint32_t)i, 77u };
        a += (float)orc_poisson(seed, fabsf(x) * 100.0f);
        (void)a;
    }
    struct { uint32_t n, bases, w, h; int ss, gps, reuse, aa, threads; } cases[] = {
        { 300, 16, 17, 13, 1, 0, 0, 0, 1 }, { 300, 9, 19, 11, 2, 0, 1, 0, 1 }, { 250, 4, 23, 9, 3, 0, 0, 1, 2 },
        { 300, 1, 16, 16, 1, 1, 0, 0, 1 }, { 200, 16, 21, 15, 2, 2, 1, 0, 2 }, { 400, 16, 31, 7, 1, 0, 1, 1, 3 },
        { 1, 16, 5, 3, 1, 0, 1, 0, 1 }, { 0, 16, 8, 8, 1, 0, 1, 0, 1 },
    };
    for (size_t k = 0; k < sizeof cases / sizeof cases[0]; k++) {
        if (run_case(cases[k].n, cases[k].bases, cases[k].w, cases[k].h, cases[k].ss, cases[k].gps, cases[k].reuse,
                     cases[k].aa, cases[k].threads)) {
            fprintf(stderr, "case %zu: render failed\n", k);
            return 1;
        }
    }
    printf("oracle sanitizer run: %zu cases clean\n", sizeof cases / sizeof cases[0]);
    return 0;
}
