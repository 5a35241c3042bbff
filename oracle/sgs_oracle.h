/*
 * Scalar CPU oracle of the sorting-free stochastic splatting renderer.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP kernels in
 * stochastic-gaussian-splatting_amd/csrc. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product path never calls it.
 *
 * It restates, one function per reference routine, the algorithm of
 *   StochasticGSCudaBackend/StochasticGSCudaBackend/{gs_utils.h, math_utils.h,
 *   scatter_utils.cuh, accumulate_scatter.cu, accumulate_hybrid*.cu, resolve.cu, render.cu}
 * (file:line citations at each definition in sgs_oracle.c). Every floating-point
 * operation is written out in a fixed order and the file is compiled with
 * -ffp-contract=off and without fast math, so the HIP kernels, which follow the
 * same sequence, reproduce its framebuffers bit for bit.
 *
 * Parity status of each function is listed in the header comment of sgs_oracle.c.
 */
#ifndef SGS_ORACLE_H
#define SGS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene {
    const float* means;           /* (N,3) world positions */
    const float* scales;          /* (N,3) log scales */
    const float* rotations;       /* (N,4) quaternion (r,i,j,k), unnormalized */
    const float* opacities;       /* (N,) logits */
    const float* sh;              /* (N,B,3) SH coefficients */
    uint32_t n_gaussians;
    uint32_t n_sh_bases;          /* B in {1,4,9,16} */
} orc_scene;

/* Camera in image pixels (not subpixels), like render_utils.h:21-32. */
typedef struct orc_camera {
    float w2c[16];                /* row-major world->view, +z forward */
    float cam_position[3];
    uint32_t width, height;
    float focal_x, focal_y, center_x, center_y;
    float near_plane, far_plane;
    int32_t proper_antialiasing;
    int32_t supersampling_factor; /* values < 1 mean 1 (render.cu:52) */
    int32_t use_gps;
    int32_t spatial_reuse;
    uint32_t max_gps_points;      /* 0: reference heuristic W*H/(2 ss^2) */
} orc_camera;

typedef struct orc_proj {
    int32_t active;
    int32_t in_depth_range;
    float depth, opacity, determinant, power_threshold;
    float mean2d[2];
    float conic[3];
    float cov2d[3];
    float bounds[4];              /* x_min, x_max, y_min, y_max (float) */
} orc_proj;

typedef struct orc_stats {
    uint64_t n_in_depth_range;    /* Gaussians inside [near, far] */
    uint64_t n_projected;         /* Gaussians that survive every cull */
    uint64_t n_tested;            /* ST: fragments evaluated; GPS: points drawn */
    uint64_t n_accepted;          /* atomicMin operations issued */
} orc_stats;

/* ---- math primitives (bit-exact twins of the device library) ---- */
float orc_exp2f(float x);
float orc_expf(float x);
float orc_logf(float x);
void  orc_sincos_2pi(float u, float* s, float* c);

/* ---- integer RNG (math_utils.h:45-79,113-136) ---- */
uint32_t orc_hash32(uint32_t x);
void     orc_make_seed(uint32_t idx, uint32_t frame, uint32_t out[2]);
void     orc_pcg2d_uint(const uint32_t in[2], uint32_t out[2]);
void     orc_pcg2d(uint32_t state[2], float out[2]);
uint32_t orc_pcg_hash(uint32_t seed);
float    orc_rand(uint32_t frame, uint32_t instance, uint32_t x, uint32_t y);
float    orc_blue_noise(const uint8_t* slice, uint32_t rotation, uint32_t instance, uint32_t x, uint32_t y);
void     orc_gps_point_seed(uint32_t gaussian_idx, uint32_t frame, uint32_t point_idx, uint32_t out[2]);

/* ---- GPS sampling math (math_utils.h:14-43,81-111) ---- */
float    orc_dilog(float x);
float    orc_inv_dilog(float x);
void     orc_box_muller(float u1, float u2, float out[2]);
void     orc_corrected_box_muller(float u1, float u2, float alpha_rcp, float alpha_dilog, float out[2]);
uint32_t orc_poisson(uint32_t state[2], float lambda);

/* ---- projection and shading ---- */
void orc_project(const orc_scene* s, uint32_t idx, const float* w2c,
                 float width, float height, float focal_x, float focal_y, float center_x, float center_y,
                 float near_plane, float far_plane, int32_t proper_antialiasing, float supersampling_factor,
                 orc_proj* out);
void orc_eval_sh(const float* coefficients, const float direction[3], uint32_t n_bases, float out[3]);
void orc_catmull_rom_weights(float t, float out[4]);
void orc_rgb_to_ycocg(const float rgb[3], float out[3]);

/* ---- accumulation: closest accepted fragment per framebuffer (sub)pixel ----
 * The framebuffer covers the crop [x0,x1) x [y0,y1) of the (W*ss) x (H*ss)
 * framebuffer; fb[(py-y0)*(x1-x0) + (px-x0)]. It must be filled with ~0 first.
 * n_threads > 1 runs OpenMP over Gaussians with a CAS-based u64 min. */
void orc_accumulate(const orc_scene* s, const orc_camera* cam, uint32_t frame_idx,
                    const uint8_t* blue_noise_slice, uint32_t blue_noise_rotation,
                    const uint32_t crop[4], uint64_t* fb, orc_stats* stats, int n_threads);

/* Pixel-major brute force over every Gaussian without bounding boxes (ST only). */
void orc_accumulate_st_bruteforce(const orc_scene* s, const orc_camera* cam, uint32_t frame_idx,
                                  const uint8_t* blue_noise_slice, uint32_t blue_noise_rotation,
                                  const uint32_t crop[4], uint64_t* fb);

/* GPS point count of every Gaussian (0 for culled ones), capped like the renderer. */
void orc_gps_counts(const orc_scene* s, const orc_camera* cam, uint32_t frame_idx,
                    uint32_t* counts_capped, uint32_t* counts_true, float* importance);

/* ---- sample + resolve (resolve.cu) on full framebuffers ---- */
void orc_sample(const orc_scene* s, const orc_camera* cam, const uint64_t* fb,
                float* samples, float* reuse_weights, int32_t* k_prim, float* k_occ);
int  orc_resolve(const orc_camera* cam, const float* samples, const float* reuse_weights,
                 const float* h_image, const float* h_depth, const float* h_alpha, const float* h_weight,
                 const float* view_to_previous_view, int32_t temporal_mode,
                 float* image, float* depth, float* alpha, float* weight_mass);

/* Full single-sample render (render.cu:16-158): clear, accumulate, sample, resolve. */
int orc_render(const orc_scene* s, const orc_camera* cam, uint32_t frame_idx,
               const uint8_t* blue_noise, uint32_t blue_noise_numel,
               int32_t* k_prim, float* k_occ,
               const float* h_image, const float* h_depth, const float* h_alpha, const float* h_weight,
               const float* view_to_previous_view, int32_t temporal_mode,
               float* image, float* depth, float* alpha, float* weight_mass,
               uint64_t* fb_out, orc_stats* stats, int n_threads);

/* FP64 front-to-back blend of every fragment the ST estimator can accept, sorted by
 * (depth, index), with per-fragment alpha and colour taken from the FP32 pipeline:
 * the expectation of one stochastic-transparency sample. Outputs premultiplied
 * image (H,W,3), depth (H,W), alpha (H,W) after the ss x ss box filter. */
void orc_blend_fp64(const orc_scene* s, const orc_camera* cam,
                    double* image, double* depth, double* alpha);

#ifdef __cplusplus
}
#endif
#endif
