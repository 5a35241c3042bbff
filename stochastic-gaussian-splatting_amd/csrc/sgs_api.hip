// C ABI entry points of libsgs.so (declared in include/sgs.h): argument validation,
// workspace carve-up and stage dispatch. Replaces the host dispatcher and pybind11
// binding of render.cu:16-165.
#include <cmath>
#include <cstring>
#include <string>

#include "sgs_internal.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* where)
{
    return fail(SGS_ERROR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

}  // namespace

namespace sgs {

FbCamera make_fb_camera(uint32_t width, uint32_t height, float focal_x, float focal_y, float center_x, float center_y,
                        float near_plane, float far_plane, int32_t proper_antialiasing, int32_t supersampling_factor)
{
    // render.cu:52 and the subpixel scaling of accumulate_scatter.cu:180-189
    FbCamera c;
    const uint32_t ss = supersampling_factor < 1 ? 1u : (uint32_t)supersampling_factor;
    c.ss = (float)ss;
    c.width_i = width * ss;
    c.height_i = height * ss;
    c.width = (float)c.width_i;
    c.height = (float)c.height_i;
    c.focal_x = focal_x * c.ss;
    c.focal_y = focal_y * c.ss;
    c.center_x = center_x * c.ss;
    c.center_y = center_y * c.ss;
    c.near_plane = near_plane;
    c.far_plane = far_plane;
    c.proper_antialiasing = proper_antialiasing ? 1 : 0;
    return c;
}

uint32_t gps_max_points(const FbCamera& cam, uint32_t override_points)
{
    if (override_points) return override_points;
    // accumulate_scatter.cu:128, round-half-even like __float2uint_rn
    const float v = cam.width * cam.height / (2.0f * cam.ss * cam.ss);
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967040.0f) return 0xffffffffu;
    return (uint32_t)std::nearbyint(v);
}

static Tuning resolve_tuning(const sgs_tuning* t, uint32_t n_gaussians)
{
    Tuning r;
    r.sequential_threshold = (t && t->sequential_threshold) ? t->sequential_threshold : 8u;
    r.binning_pixel_threshold = (t && t->binning_pixel_threshold) ? t->binning_pixel_threshold : 512u;
    r.binning_point_threshold = (t && t->binning_point_threshold) ? t->binning_point_threshold : 512u;
    uint32_t mb = (t && t->max_binned_gaussians) ? t->max_binned_gaussians : (1u << 20);
    if (mb > n_gaussians) mb = n_gaussians;
    r.max_binned = mb < 1u ? 1u : mb;
    r.tile_size = (t && t->tile_size) ? t->tile_size : 32u;
    r.max_tile_instances = (t && t->max_tile_instances) ? t->max_tile_instances : (1u << 24);
    r.max_gps_points = t ? t->max_gps_points : 0u;
    return r;
}

}  // namespace sgs

using namespace sgs;

namespace {

struct RenderLayout {
    size_t fb, samples, reuse, hybrid, total;
};

size_t hybrid_bytes(const sgs_settings* s, const FbCamera& cam, const Tuning& t)
{
    if (!s->hybrid_scheduling) return 0;
    return s->use_gps ? hybrid_gps_scratch_bytes(t) : hybrid_st_scratch_bytes(cam.width_i, cam.height_i, t);
}

RenderLayout render_layout(const sgs_settings* s, const FbCamera& cam, const Tuning& t)
{
    const size_t P = (size_t)cam.width_i * cam.height_i;
    Arena a;
    RenderLayout L;
    L.fb = a.take(8u * P);
    L.samples = a.take(16u * P);
    L.reuse = s->spatial_reuse ? a.take(16u * P) : 0;
    const size_t hb = hybrid_bytes(s, cam, t);
    L.hybrid = hb ? a.take(hb) : 0;
    L.total = a.offset;
    return L;
}

int check_settings(const sgs_settings* s)
{
    if (!s) return fail(SGS_ERROR_INVALID_ARGUMENT, "settings is NULL");
    if (!s->w2c || !s->cam_position) return fail(SGS_ERROR_INVALID_ARGUMENT, "settings.w2c / cam_position is NULL");
    if (s->width == 0 || s->height == 0) return fail(SGS_ERROR_INVALID_ARGUMENT, "image width and height must be > 0");
    const uint64_t ss = s->supersampling_factor < 1 ? 1u : (uint64_t)s->supersampling_factor;
    const uint64_t fw = (uint64_t)s->width * ss, fh = (uint64_t)s->height * ss;
    if (fw > 65535u || fh > 65535u || fw * fh >= (1ull << 32))
        return fail(SGS_ERROR_INVALID_ARGUMENT, "framebuffer too large: each side must be <= 65535 and W*H < 2^32");
    if (!(s->focal_x != 0.0f) || !(s->focal_y != 0.0f)) return fail(SGS_ERROR_INVALID_ARGUMENT, "focal length is 0");
    return SGS_SUCCESS;
}

int check_gaussians(const sgs_gaussians* g)
{
    if (!g) return fail(SGS_ERROR_INVALID_ARGUMENT, "gaussians is NULL");
    if (g->n_gaussians > 0 && (!g->means || !g->scales || !g->rotations || !g->opacities || !g->sh_coefficients))
        return fail(SGS_ERROR_INVALID_ARGUMENT, "a Gaussian parameter pointer is NULL");
    if (reinterpret_cast<uintptr_t>(g->rotations) % 16u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "rotations must be 16-byte aligned");
    const uint32_t b = g->total_sh_bases;
    if (b != 1 && b != 4 && b != 9 && b != 16)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "total_sh_bases must be 1, 4, 9 or 16 (SH degree 0-3)");
    if ((b == 4 || b == 16) && reinterpret_cast<uintptr_t>(g->sh_coefficients) % 16u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "sh_coefficients must be 16-byte aligned");
    return SGS_SUCCESS;
}

int check_tuning(const sgs_tuning* t)
{
    if (t && t->tile_size && t->tile_size != 16u && t->tile_size != 32u && t->tile_size != 64u)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "tuning.tile_size must be 16, 32 or 64");
    return SGS_SUCCESS;
}

SceneArgs scene_args(const sgs_gaussians* g)
{
    return { g->means, g->scales, g->rotations, g->opacities, g->sh_coefficients, g->n_gaussians, g->total_sh_bases };
}

// clear + accumulate into fb
int accumulate_into(const sgs_gaussians* g, const sgs_settings* s, const FbCamera& cam, const uint8_t* blue_noise,
                    uint64_t blue_noise_numel, uint64_t* fb, uint32_t frame_idx, const Tuning& t, uint64_t* counters,
                    void* hybrid_scratch, hipStream_t stream)
{
    const size_t P = (size_t)cam.width_i * cam.height_i;
    hipError_t e = launch_fill_u64(fb, P, kEmpty, stream);
    if (e != hipSuccess) return hip_fail(e, "framebuffer clear");
    if (counters) {
        e = launch_fill_u64(counters, 4, 0ull, stream);
        if (e != hipSuccess) return hip_fail(e, "counter clear");
    }
    // one blue-noise slice per frame, rotated by the golden ratio per pass (render.cu:77-82)
    const uint8_t* slice = nullptr;
    uint32_t rotation = 0;
    if (blue_noise && blue_noise_numel > 0 && !s->use_gps) {
        slice = blue_noise + (size_t)(frame_idx % 64u) * 128u * 128u;
        rotation = (frame_idx / 64u) * 0x9E3779B9u;
    }
    const SceneArgs sc = scene_args(g);
    if (s->use_gps) {
        const uint32_t max_points = gps_max_points(cam, t.max_gps_points);
        if (s->hybrid_scheduling)
            e = run_hybrid_gps(sc, s->w2c, cam, fb, frame_idx, max_points, t, hybrid_scratch, counters, stream);
        else
            e = launch_gps_scatter(sc, s->w2c, cam, fb, frame_idx, max_points, t.sequential_threshold, BinArgs{},
                                   counters, stream);
    } else {
        if (s->hybrid_scheduling)
            e = run_hybrid_st(sc, s->w2c, cam, fb, frame_idx, slice, rotation, t, hybrid_scratch, counters, stream);
        else
            e = launch_st_scatter(sc, s->w2c, cam, fb, frame_idx, slice, rotation, t.sequential_threshold, BinArgs{},
                                  counters, stream);
    }
    if (e != hipSuccess) return hip_fail(e, "accumulation");
    return SGS_SUCCESS;
}

int check_blue_noise(const uint8_t* bn, uint64_t numel)
{
    if (numel != 0 && numel != 64u * 128u * 128u)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "blue_noise must hold 0 or 64*128*128 elements");
    if (numel != 0 && !bn) return fail(SGS_ERROR_INVALID_ARGUMENT, "blue_noise is NULL");
    return SGS_SUCCESS;
}

}  // namespace

extern "C" {

int sgs_k_buffer_size(void) { return SGS_K_BUFFER_SIZE; }

const char* sgs_last_error(void) { return g_last_error.c_str(); }

size_t sgs_render_workspace_size(const sgs_settings* settings, uint32_t n_gaussians, const sgs_tuning* tuning)
{
    if (!settings) return 0;
    const FbCamera cam = make_fb_camera(settings->width, settings->height, settings->focal_x, settings->focal_y,
                                        settings->center_x, settings->center_y, settings->near_plane,
                                        settings->far_plane, settings->proper_antialiasing,
                                        settings->supersampling_factor);
    return render_layout(settings, cam, resolve_tuning(tuning, n_gaussians)).total;
}

int sgs_render(const sgs_gaussians* gaussians, const sgs_settings* s, const uint8_t* blue_noise,
               uint64_t blue_noise_numel, int32_t* k_buffer_primitives, float* k_buffer_occlusion,
               const sgs_history* history, const float* view_to_previous_view, int32_t temporal_mode,
               const sgs_history* out, uint32_t frame_idx, const sgs_tuning* tuning, void* workspace,
               size_t workspace_bytes, void* stream_ptr)
{
    int rc;
    if ((rc = check_gaussians(gaussians)) || (rc = check_settings(s)) || (rc = check_tuning(tuning)) ||
        (rc = check_blue_noise(blue_noise, blue_noise_numel)))
        return rc;
    if (temporal_mode < SGS_TEMPORAL_RESET || temporal_mode > SGS_TEMPORAL_REPROJECT)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "invalid temporal mode " + std::to_string(temporal_mode));
    if (!out || !out->image || !out->depth || !out->alpha || !out->weight_mass)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "output pointers are NULL");
    if (temporal_mode != SGS_TEMPORAL_RESET &&
        (!history || !history->image || !history->depth || !history->alpha || !history->weight_mass))
        return fail(SGS_ERROR_INVALID_ARGUMENT, "STATIC and REPROJECT need a history");
    if (temporal_mode == SGS_TEMPORAL_REPROJECT && !view_to_previous_view)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "REPROJECT needs view_to_previous_view");
    if (s->spatial_reuse && (!k_buffer_primitives || !k_buffer_occlusion))
        return fail(SGS_ERROR_INVALID_ARGUMENT, "spatial reuse needs the k-buffer");
    if (history && temporal_mode != SGS_TEMPORAL_RESET &&
        (history->image == out->image || history->depth == out->depth || history->alpha == out->alpha ||
         history->weight_mass == out->weight_mass))
        return fail(SGS_ERROR_INVALID_ARGUMENT, "outputs must not alias the history");

    const Tuning t = resolve_tuning(tuning, gaussians->n_gaussians);
    const FbCamera cam = make_fb_camera(s->width, s->height, s->focal_x, s->focal_y, s->center_x, s->center_y,
                                        s->near_plane, s->far_plane, s->proper_antialiasing, s->supersampling_factor);
    const RenderLayout L = render_layout(s, cam, t);
    if (!workspace || workspace_bytes < L.total)
        return fail(SGS_ERROR_WORKSPACE_TOO_SMALL,
                    "workspace of " + std::to_string(L.total) + " bytes needed, got " + std::to_string(workspace_bytes));
    if (reinterpret_cast<uintptr_t>(workspace) % 256u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");

    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    char* ws = static_cast<char*>(workspace);
    uint64_t* fb = reinterpret_cast<uint64_t*>(ws + L.fb);
    if ((rc = accumulate_into(gaussians, s, cam, blue_noise, blue_noise_numel, fb, frame_idx, t, nullptr,
                              L.hybrid ? ws + L.hybrid : nullptr, stream)))
        return rc;

    const SceneArgs sc = scene_args(gaussians);
    HistoryPtrs hp = { nullptr, nullptr, nullptr, nullptr };
    if (history) hp = { history->image, history->depth, history->alpha, history->weight_mass };
    const OutputPtrs op = { out->image, out->depth, out->alpha, out->weight_mass };
    hipError_t e;
    if (!s->spatial_reuse && temporal_mode != SGS_TEMPORAL_REPROJECT) {
        e = launch_shade_resolve(sc, s->cam_position, cam, s->width, s->height, fb, hp, op, temporal_mode, stream);
        if (e != hipSuccess) return hip_fail(e, "shade_resolve");
    } else {
        float4* samples = reinterpret_cast<float4*>(ws + L.samples);
        float4* rw = s->spatial_reuse ? reinterpret_cast<float4*>(ws + L.reuse) : nullptr;
        e = launch_sample(sc, s->w2c, s->cam_position, cam, fb, samples, rw, reinterpret_cast<uint2*>(k_buffer_primitives),
                          reinterpret_cast<float4*>(k_buffer_occlusion), s->spatial_reuse != 0, stream);
        if (e != hipSuccess) return hip_fail(e, "sample");
        e = launch_resolve(cam, s->width, s->height, s->focal_x, s->focal_y, s->center_x, s->center_y, samples, rw, hp,
                           view_to_previous_view, op, s->spatial_reuse != 0, temporal_mode, stream);
        if (e != hipSuccess) return hip_fail(e, "resolve");
    }
    return SGS_SUCCESS;
}

size_t sgs_accumulate_workspace_size(const sgs_settings* settings, uint32_t n_gaussians, const sgs_tuning* tuning)
{
    if (!settings) return 0;
    const FbCamera cam = make_fb_camera(settings->width, settings->height, settings->focal_x, settings->focal_y,
                                        settings->center_x, settings->center_y, settings->near_plane,
                                        settings->far_plane, settings->proper_antialiasing,
                                        settings->supersampling_factor);
    const size_t hb = hybrid_bytes(settings, cam, resolve_tuning(tuning, n_gaussians));
    return hb ? ((hb + 255u) & ~size_t(255u)) : 0;
}

int sgs_accumulate(const sgs_gaussians* gaussians, const sgs_settings* s, const uint8_t* blue_noise,
                   uint64_t blue_noise_numel, uint64_t* framebuffer, uint32_t frame_idx, const sgs_tuning* tuning,
                   uint64_t* counters, void* workspace, size_t workspace_bytes, void* stream_ptr)
{
    int rc;
    if ((rc = check_gaussians(gaussians)) || (rc = check_settings(s)) || (rc = check_tuning(tuning)) ||
        (rc = check_blue_noise(blue_noise, blue_noise_numel)))
        return rc;
    if (!framebuffer) return fail(SGS_ERROR_INVALID_ARGUMENT, "framebuffer is NULL");
    if (reinterpret_cast<uintptr_t>(framebuffer) % 16u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "framebuffer must be 16-byte aligned");
    const Tuning t = resolve_tuning(tuning, gaussians->n_gaussians);
    const FbCamera cam = make_fb_camera(s->width, s->height, s->focal_x, s->focal_y, s->center_x, s->center_y,
                                        s->near_plane, s->far_plane, s->proper_antialiasing, s->supersampling_factor);
    const size_t need = hybrid_bytes(s, cam, t);
    if (need && (!workspace || workspace_bytes < need))
        return fail(SGS_ERROR_WORKSPACE_TOO_SMALL,
                    "workspace of " + std::to_string(need) + " bytes needed, got " + std::to_string(workspace_bytes));
    if (need && reinterpret_cast<uintptr_t>(workspace) % 256u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
    return accumulate_into(gaussians, s, cam, blue_noise, blue_noise_numel, framebuffer, frame_idx, t, counters,
                           need ? workspace : nullptr, static_cast<hipStream_t>(stream_ptr));
}

}  // extern "C"
