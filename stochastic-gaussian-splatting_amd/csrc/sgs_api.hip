// C ABI entry points of libsgs.so (declared in include/sgs.h): argument validation,
// workspace carve-up and stage dispatch. Replaces the host dispatcher and pybind11
// binding of render.cu:16-165.
#include <cmath>
#include <mutex>
#include <cstring>
#include <algorithm>
#include <string>
#include <vector>

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
    // the same IEEE single-precision operations the projection would do per Gaussian
    c.clip_left = (-0.15f * c.width - c.center_x) / c.focal_x;
    c.clip_right = (1.15f * c.width - c.center_x) / c.focal_x;
    c.clip_top = (-0.15f * c.height - c.center_y) / c.focal_y;
    c.clip_bottom = (1.15f * c.height - c.center_y) / c.focal_y;
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
    r.tiles = !(t && t->schedule == SGS_BATCH_SCATTER);
    r.max_workspace_bytes = t ? (uint64_t)t->max_workspace_mib << 20 : 0u;
    // one sample of one view: 16x16 bins (larger bins leave small frames with few workgroups)
    r.bin_log = (t && t->bin_size) ? bin_log2(t->bin_size) : 0u;
    return r;
}

}  // namespace sgs

using namespace sgs;

namespace {

// ---- single-view tile schedule (the batched renderer's tile pipeline for one view) ----
struct SingleTileLayout {
    size_t counters, ready, tile_count, tile_cursor, cleared_end, offsets, entry_buckets, entry_gaussians,
        sorted_instances, sorted_keys,
        centroids, view_keys, cov3d, pair_rect, table, queue, total;
    size_t views_bytes;
    uint32_t n_tiles, capacity, queue_capacity;
};

constexpr uint32_t kSingleQueue = 1u << 16;  // work queue of the scatter fallback

// instance capacity: about 4 + 8 P / 2^21 per Gaussian for a framebuffer of P pixels (a
// Gaussian of a given size covers more 16x16 tiles the finer the framebuffer); a frame that
// needs more falls back to the scatter schedule
uint32_t single_tile_capacity(uint64_t P, uint32_t n_gaussians)
{
    const uint64_t want = (uint64_t)n_gaussians * (4ull + ((8ull * P) >> 21));
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, 1ull << 16), 1ull << 31);
}

SingleTileLayout single_tile_layout(const FbCamera& cam, uint32_t n_gaussians, uint32_t bin_log)
{
    SingleTileLayout L;
    const uint64_t P = (uint64_t)cam.width_i * cam.height_i;
    const uint32_t bin = kRasterPixels << bin_log;
    L.n_tiles = ((cam.width_i + bin - 1u) / bin) * ((cam.height_i + bin - 1u) / bin);
    L.capacity = single_tile_capacity(P, n_gaussians);
    L.queue_capacity = kSingleQueue;
    L.views_bytes = (sizeof(ViewDesc) + 255u) & ~size_t(255u);
    Arena a;
    L.counters = a.take(sizeof(BatchCounters));
    L.ready = a.offset;
    a.offset += 4u * (size_t)L.queue_capacity;
    L.tile_count = a.offset;
    a.offset += 4u * (size_t)L.n_tiles;
    L.tile_cursor = a.offset;
    a.offset += 4u * (size_t)L.n_tiles;
    L.cleared_end = a.offset;
    a.offset = (a.offset + 255u) & ~size_t(255u);
    L.offsets = a.take(4u * ((size_t)L.n_tiles + 1u));
    L.entry_buckets = a.take(2u * (size_t)L.capacity);    // depth bucket per instance
    L.entry_gaussians = a.take(4u * (size_t)L.capacity);  // Gaussian index per instance
    L.sorted_instances = a.take(4u * (size_t)L.capacity);
    L.sorted_keys = a.take(4u * (size_t)L.capacity);
    L.centroids = a.take(16u * (((size_t)n_gaussians + 255u) / 256u));
    L.view_keys = a.take(8u);
    L.cov3d = a.take(32u * (size_t)n_gaussians);
    L.pair_rect = a.take(8u * (size_t)n_gaussians);
    L.table = a.take(L.views_bytes + sizeof(SingleViewPrefixes));
    L.queue = a.take(sizeof(BigEntry) * (size_t)L.queue_capacity);
    L.total = a.offset;
    return L;
}

bool single_tiles_allowed(const sgs_settings* s, const Tuning& t)
{
    return t.tiles && !s->use_gps && !s->hybrid_scheduling;
}

struct RenderLayout {
    size_t fb, samples, reuse, hybrid, tiles, total;
};

size_t hybrid_bytes(const sgs_settings* s, const FbCamera& cam, const Tuning& t)
{
    if (!s->hybrid_scheduling) return 0;
    return s->use_gps ? hybrid_gps_scratch_bytes(t) : hybrid_st_scratch_bytes(cam.width_i, cam.height_i, t);
}

RenderLayout render_layout(const sgs_settings* s, const FbCamera& cam, const Tuning& t, uint32_t n_gaussians)
{
    const size_t P = (size_t)cam.width_i * cam.height_i;
    Arena a;
    RenderLayout L;
    L.fb = a.take(8u * P);
    L.samples = a.take(16u * P);
    L.reuse = s->spatial_reuse ? a.take(16u * P) : 0;
    const size_t hb = hybrid_bytes(s, cam, t);
    L.hybrid = hb ? a.take(hb) : 0;
    L.tiles = single_tiles_allowed(s, t) ? a.take(single_tile_layout(cam, n_gaussians, t.bin_log).total) : 0;
    L.total = a.offset;
    return L;
}

// The tuning of one view's render: resolve_tuning, then, when the tile schedule's workspace
// would exceed the caller's cap, the scatter schedule (no binning buffers; the same
// framebuffer bit for bit)
Tuning view_tuning(const sgs_settings* s, const FbCamera& cam, const sgs_tuning* tuning, uint32_t n_gaussians)
{
    Tuning t = resolve_tuning(tuning, n_gaussians);
    if (t.max_workspace_bytes && single_tiles_allowed(s, t) &&
        render_layout(s, cam, t, n_gaussians).total > t.max_workspace_bytes)
        t.tiles = false;
    return t;
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
    if (reinterpret_cast<uintptr_t>(g->cov3d) % 16u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "cov3d must be 16-byte aligned");
    return SGS_SUCCESS;
}

int check_tuning(const sgs_tuning* t)
{
    if (t && t->tile_size && t->tile_size != 16u && t->tile_size != 32u && t->tile_size != 64u)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "tuning.tile_size must be 16, 32 or 64");
    if (t && t->schedule > SGS_BATCH_SCATTER)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "tuning.schedule must be 0, 1 (tiles) or 2 (scatter)");
    if (t && t->bin_size && bin_log2(t->bin_size) >= kBinSizes)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "tuning.bin_size must be 16, 32 or 64");
    return SGS_SUCCESS;
}

// the tile schedule for one view and one sample into fb (defined with the batched renderer)
int accumulate_tiles_single(const sgs_gaussians* g, const sgs_settings* s, const FbCamera& cam, uint64_t* fb,
                            uint32_t frame_idx, uint32_t bin_log, char* tile_ws, hipStream_t stream);

SceneArgs scene_args(const sgs_gaussians* g)
{
    return { g->means, g->scales, g->rotations, g->opacities, g->sh_coefficients, g->n_gaussians, g->total_sh_bases,
             reinterpret_cast<const float4*>(g->cov3d) };
}

// clear + accumulate into fb (tile_ws: the single-view tile section of the workspace, or NULL)
int accumulate_into(const sgs_gaussians* g, const sgs_settings* s, const FbCamera& cam, const uint8_t* blue_noise,
                    uint64_t blue_noise_numel, uint64_t* fb, uint32_t frame_idx, const Tuning& t, uint64_t* counters,
                    void* hybrid_scratch, void* tile_ws, hipStream_t stream)
{
    // the tile schedule writes every framebuffer word (no clear); blue noise and the debug
    // counters exist in the scatter kernels only
    if (tile_ws && !(blue_noise && blue_noise_numel > 0) && !counters)
        return accumulate_tiles_single(g, s, cam, fb, frame_idx, t.bin_log, static_cast<char*>(tile_ws), stream);
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

int check_render_args(const sgs_gaussians* gaussians, const sgs_settings* s, const uint8_t* blue_noise,
                      uint64_t blue_noise_numel, const int32_t* k_buffer_primitives, const float* k_buffer_occlusion,
                      const sgs_history* history, const float* view_to_previous_view, int32_t temporal_mode,
                      const sgs_history* out, const sgs_tuning* tuning)
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
    return SGS_SUCCESS;
}

// shading of each (sub)pixel's winner and the resolve into the outputs (resolve.cu:355-444):
// fused when neither spatial reuse nor REPROJECT needs the per-subpixel samples
int shade_and_resolve(const sgs_gaussians* g, const sgs_settings* s, const FbCamera& cam, const uint64_t* fb,
                      float4* samples, float4* reuse_weights, int32_t* k_buffer_primitives, float* k_buffer_occlusion,
                      const sgs_history* history, const float* view_to_previous_view, int32_t temporal_mode,
                      const sgs_history* out, hipStream_t stream)
{
    const SceneArgs sc = scene_args(g);
    HistoryPtrs hp = { nullptr, nullptr, nullptr, nullptr };
    if (history) hp = { history->image, history->depth, history->alpha, history->weight_mass };
    const OutputPtrs op = { out->image, out->depth, out->alpha, out->weight_mass };
    hipError_t e;
    if (!s->spatial_reuse && temporal_mode != SGS_TEMPORAL_REPROJECT) {
        e = launch_shade_resolve(sc, s->cam_position, cam, s->width, s->height, fb, hp, op, temporal_mode, stream);
        if (e != hipSuccess) return hip_fail(e, "shade_resolve");
        return SGS_SUCCESS;
    }
    float4* rw = s->spatial_reuse ? reuse_weights : nullptr;
    e = launch_sample(sc, s->w2c, s->cam_position, cam, fb, samples, rw, reinterpret_cast<uint2*>(k_buffer_primitives),
                      reinterpret_cast<float4*>(k_buffer_occlusion), s->spatial_reuse != 0, stream);
    if (e != hipSuccess) return hip_fail(e, "sample");
    e = launch_resolve(cam, s->width, s->height, s->focal_x, s->focal_y, s->center_x, s->center_y, samples, rw, hp,
                       view_to_previous_view, op, s->spatial_reuse != 0, temporal_mode, stream);
    if (e != hipSuccess) return hip_fail(e, "resolve");
    return SGS_SUCCESS;
}

// ---- fixed device scratch of sgs_render_noalloc ----------------------------------------
// One array per GPU, allocated with the code object when the library is loaded on that GPU
// and held until the process ends. Host code never passes the array's name to a kernel (on
// the host it is a stand-in, not a device address): each device's address is taken once,
// with that device current, through hipGetSymbolAddress and kept in g_scratch_ptr. Calls
// on one GPU share its array, so calls that overlap in time on one GPU race.
static __device__ __align__(256) unsigned char scratch[SGS_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
char* g_scratch_ptr[kMaxDevices] = {};

int device_scratch(char** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(SGS_ERROR_INVALID_ARGUMENT, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p || reinterpret_cast<uintptr_t>(p) % 256u != 0)
            return fail(SGS_ERROR_HIP, "scratch array has no 256-byte aligned device address");
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return SGS_SUCCESS;
}

struct ScratchLayout {
    size_t fb, samples, reuse, total;
};

// framebuffer words, and the per-subpixel samples / reuse weights only where the resolve reads them
ScratchLayout scratch_layout(const sgs_settings* s, const FbCamera& cam, int32_t temporal_mode)
{
    const size_t P = (size_t)cam.width_i * cam.height_i;
    Arena a;
    ScratchLayout L;
    L.fb = a.take(8u * P);
    const bool need_samples = s->spatial_reuse || temporal_mode == SGS_TEMPORAL_REPROJECT;
    L.samples = need_samples ? a.take(16u * P) : 0;
    L.reuse = s->spatial_reuse ? a.take(16u * P) : 0;
    L.total = a.offset;
    return L;
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
    return render_layout(settings, cam, view_tuning(settings, cam, tuning, n_gaussians), n_gaussians).total;
}

int sgs_render(const sgs_gaussians* gaussians, const sgs_settings* s, const uint8_t* blue_noise,
               uint64_t blue_noise_numel, int32_t* k_buffer_primitives, float* k_buffer_occlusion,
               const sgs_history* history, const float* view_to_previous_view, int32_t temporal_mode,
               const sgs_history* out, uint32_t frame_idx, const sgs_tuning* tuning, void* workspace,
               size_t workspace_bytes, void* stream_ptr)
{
    int rc;
    if ((rc = check_render_args(gaussians, s, blue_noise, blue_noise_numel, k_buffer_primitives, k_buffer_occlusion,
                                history, view_to_previous_view, temporal_mode, out, tuning)))
        return rc;

    const FbCamera cam = make_fb_camera(s->width, s->height, s->focal_x, s->focal_y, s->center_x, s->center_y,
                                        s->near_plane, s->far_plane, s->proper_antialiasing, s->supersampling_factor);
    const Tuning t = view_tuning(s, cam, tuning, gaussians->n_gaussians);
    const RenderLayout L = render_layout(s, cam, t, gaussians->n_gaussians);
    if (!workspace || workspace_bytes < L.total)
        return fail(SGS_ERROR_WORKSPACE_TOO_SMALL,
                    "workspace of " + std::to_string(L.total) + " bytes needed, got " + std::to_string(workspace_bytes));
    if (reinterpret_cast<uintptr_t>(workspace) % 256u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");

    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    char* ws = static_cast<char*>(workspace);
    uint64_t* fb = reinterpret_cast<uint64_t*>(ws + L.fb);
    if ((rc = accumulate_into(gaussians, s, cam, blue_noise, blue_noise_numel, fb, frame_idx, t, nullptr,
                              L.hybrid ? ws + L.hybrid : nullptr, L.tiles ? ws + L.tiles : nullptr, stream)))
        return rc;
    return shade_and_resolve(gaussians, s, cam, fb, reinterpret_cast<float4*>(ws + L.samples),
                             s->spatial_reuse ? reinterpret_cast<float4*>(ws + L.reuse) : nullptr, k_buffer_primitives,
                             k_buffer_occlusion, history, view_to_previous_view, temporal_mode, out, stream);
}

int sgs_render_noalloc(const sgs_gaussians* gaussians, const sgs_settings* s, const uint8_t* blue_noise,
                       uint64_t blue_noise_numel, int32_t* k_buffer_primitives, float* k_buffer_occlusion,
                       const sgs_history* history, const float* view_to_previous_view, int32_t temporal_mode,
                       const sgs_history* out, uint32_t frame_idx, const sgs_tuning* tuning, void* stream_ptr)
{
    int rc;
    if ((rc = check_render_args(gaussians, s, blue_noise, blue_noise_numel, k_buffer_primitives, k_buffer_occlusion,
                                history, view_to_previous_view, temporal_mode, out, tuning)))
        return rc;
    const FbCamera cam = make_fb_camera(s->width, s->height, s->focal_x, s->focal_y, s->center_x, s->center_y,
                                        s->near_plane, s->far_plane, s->proper_antialiasing, s->supersampling_factor);
    const ScratchLayout L = scratch_layout(s, cam, temporal_mode);
    if (L.total > SGS_NOALLOC_SCRATCH_BYTES)
        return fail(SGS_ERROR_WORKSPACE_TOO_SMALL,
                    "sgs_render_noalloc: this view needs " + std::to_string(L.total) + " bytes of scratch, more than "
                    "the library's fixed " + std::to_string((size_t)SGS_NOALLOC_SCRATCH_BYTES) +
                    " (8 bytes per framebuffer pixel, +16 with REPROJECT or spatial reuse, +16 more with spatial "
                    "reuse); use sgs_render with a workspace");
    char* ws = nullptr;
    if ((rc = device_scratch(&ws))) return rc;

    // the scatter schedule of the reference (no binning buffers): the same framebuffer as the
    // tile and hybrid schedules, bit for bit
    sgs_settings scatter = *s;
    scatter.hybrid_scheduling = 0;
    Tuning t = resolve_tuning(tuning, gaussians->n_gaussians);
    t.tiles = false;
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    uint64_t* fb = reinterpret_cast<uint64_t*>(ws + L.fb);
    if ((rc = accumulate_into(gaussians, &scatter, cam, blue_noise, blue_noise_numel, fb, frame_idx, t, nullptr,
                              nullptr, nullptr, stream)))
        return rc;
    return shade_and_resolve(gaussians, s, cam, fb, L.samples ? reinterpret_cast<float4*>(ws + L.samples) : nullptr,
                             L.reuse ? reinterpret_cast<float4*>(ws + L.reuse) : nullptr, k_buffer_primitives,
                             k_buffer_occlusion, history, view_to_previous_view, temporal_mode, out, stream);
}

int sgs_precompute_cov3d(const sgs_gaussians* gaussians, float* cov3d, void* stream)
{
    int rc;
    if ((rc = check_gaussians(gaussians))) return rc;
    if (gaussians->n_gaussians > 0 && !cov3d) return fail(SGS_ERROR_INVALID_ARGUMENT, "cov3d is NULL");
    if (reinterpret_cast<uintptr_t>(cov3d) % 16u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "cov3d must be 16-byte aligned");
    SceneArgs sc = scene_args(gaussians);
    sc.cov3d = nullptr;  // computed from the raw parameters
    const hipError_t e = launch_cov3d(sc, reinterpret_cast<float4*>(cov3d), static_cast<hipStream_t>(stream));
    return e == hipSuccess ? SGS_SUCCESS : hip_fail(e, "cov3d");
}

size_t sgs_accumulate_workspace_size(const sgs_settings* settings, uint32_t n_gaussians, const sgs_tuning* tuning)
{
    if (!settings) return 0;
    const FbCamera cam = make_fb_camera(settings->width, settings->height, settings->focal_x, settings->focal_y,
                                        settings->center_x, settings->center_y, settings->near_plane,
                                        settings->far_plane, settings->proper_antialiasing,
                                        settings->supersampling_factor);
    const Tuning t = view_tuning(settings, cam, tuning, n_gaussians);
    const size_t hb = hybrid_bytes(settings, cam, t);
    if (single_tiles_allowed(settings, t)) return single_tile_layout(cam, n_gaussians, t.bin_log).total;
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
    const FbCamera cam = make_fb_camera(s->width, s->height, s->focal_x, s->focal_y, s->center_x, s->center_y,
                                        s->near_plane, s->far_plane, s->proper_antialiasing, s->supersampling_factor);
    const Tuning t = view_tuning(s, cam, tuning, gaussians->n_gaussians);
    const bool tiles = single_tiles_allowed(s, t);
    const size_t need = tiles ? single_tile_layout(cam, gaussians->n_gaussians, t.bin_log).total : hybrid_bytes(s, cam, t);
    if (need && (!workspace || workspace_bytes < need))
        return fail(SGS_ERROR_WORKSPACE_TOO_SMALL,
                    "workspace of " + std::to_string(need) + " bytes needed, got " + std::to_string(workspace_bytes));
    if (need && reinterpret_cast<uintptr_t>(workspace) % 256u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
    return accumulate_into(gaussians, s, cam, blue_noise, blue_noise_numel, framebuffer, frame_idx, t, counters,
                           need && !tiles ? workspace : nullptr, tiles ? workspace : nullptr,
                           static_cast<hipStream_t>(stream_ptr));
}

}  // extern "C"

// ============================================================================================
// batched multi-view rendering
// ============================================================================================
namespace {

BatchTuning resolve_batch_tuning(const sgs_batch_tuning* t)
{
    BatchTuning r;
    r.sequential_threshold = (t && t->sequential_threshold) ? t->sequential_threshold : 8u;
    r.big_threshold = (t && t->big_threshold) ? t->big_threshold : 4096u;
    r.entry_fragments = (t && t->entry_fragments) ? t->entry_fragments : 2048u;
    r.queue_capacity = (t && t->queue_capacity) ? t->queue_capacity : (1u << 20);
    r.view_major = (t && t->view_major) ? 1u : 0u;
    r.blocks_per_cu = t ? t->blocks_per_cu : 0u;
    r.schedule = (t && t->schedule == SGS_BATCH_SCATTER) ? 1u : 0u;
    r.max_tile_instances = t ? t->max_tile_instances : 0u;
    r.tile_samples = t ? t->tile_samples : 0u;
    return r;
}

uint64_t view_fb_pixels(const sgs_view& v);
bool view_routed(const sgs_view& v);

// default capacity of the tile schedule's instance list (14 bytes of workspace each): per
// view, 4 + 8 P / 2^21 per Gaussian for a framebuffer of P pixels (a Gaussian covers more
// 16x16 tiles the finer the framebuffer), as for a single view
uint32_t default_tile_instances(const sgs_view* views, uint32_t n_views, uint32_t n_gaussians)
{
    uint64_t per = 0;
    for (uint32_t v = 0; v < n_views; v++)
        if (!view_routed(views[v])) per += 4ull + ((8ull * view_fb_pixels(views[v])) >> 21);
    const uint64_t want = per >= (1ull << 32) ? (1ull << 32) : per * n_gaussians;
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, 1ull << 16), 1ull << 31);
}

// batched resolve workgroups of a view: one per 16x16 tile of its image pixels
uint64_t view_resolve_blocks(const sgs_view& v)
{
    return ((uint64_t)(v.width + kResolveTile - 1u) / kResolveTile) * ((v.height + kResolveTile - 1u) / kResolveTile);
}

// device view table: ViewDesc[n], then the resolve block prefix[n+1], the bin prefix[n+1] and
// the rasteriser workgroup prefix[n+1]
size_t table_views_bytes(uint32_t n) { return ((size_t)n * sizeof(ViewDesc) + 255u) & ~size_t(255u); }
size_t table_bytes(uint32_t n) { return table_views_bytes(n) + 12u * ((size_t)n + 1u); }

uint32_t view_ss(const sgs_view& v) { return v.supersampling_factor < 1 ? 1u : (uint32_t)v.supersampling_factor; }
uint64_t view_fb_pixels(const sgs_view& v)
{
    const uint64_t ss = view_ss(v);
    return (uint64_t)v.width * ss * v.height * ss;
}
// bins of (16 << b) pixels: the view's bin_size, else 16 for framebuffers of at most
// kSmallBinMaxPixels pixels and 32 above (check_views rejects other sizes)
uint32_t view_bin_log(const sgs_view& v)
{
    if (v.bin_size) return bin_log2(v.bin_size);
    return view_fb_pixels(v) <= kSmallBinMaxPixels ? 0u : 1u;
}
// views whose samples the single-view scatter kernels accumulate (GPS, blue noise): they have
// no bins in the batch's tile schedule
bool view_routed(const sgs_view& v) { return v.use_gps || (v.blue_noise && v.blue_noise_numel > 0); }
uint32_t view_grid_w(const sgs_view& v)
{
    if (view_routed(v)) return 0u;
    const uint32_t bin = kRasterPixels << view_bin_log(v);
    return (uint32_t)((v.width * view_ss(v) + bin - 1u) / bin);
}
uint32_t view_grid_h(const sgs_view& v)
{
    if (view_routed(v)) return 0u;
    const uint32_t bin = kRasterPixels << view_bin_log(v);
    return (uint32_t)((v.height * view_ss(v) + bin - 1u) / bin);
}
// rasteriser workgroups of a view: 16x16 tiles of its bins, 4^b per bin, padded to a multiple
// of 4 so every view starts a group of four (the extra workgroups return at once)
uint64_t view_raster_groups(const sgs_view& v)
{
    return ((((uint64_t)view_grid_w(v) * view_grid_h(v)) << (2u * view_bin_log(v))) + 3u) & ~3ull;
}

// Workspace of a batch. The counters, the queue's ready flags and the tile counts and
// cursors are contiguous so one fill clears them all.
struct BatchLayout {
    size_t counters, ready, tile_count, tile_cursor, cleared_end, offsets, entry_buckets, entry_gaussians,
        sorted_instances, sorted_keys,
        centroids, cov3d, pair_rect, view_keys, queue, table, fb, samples, reuse_samples, reuse_weights,
        reuse_temp, total;
    uint64_t fb_words, n_samples, n_tiles, n_raster;
    uint32_t n_blocks, capacity;
};

BatchLayout batch_layout(const sgs_view* views, uint32_t n, uint32_t n_gaussians, const BatchTuning& t)
{
    BatchLayout L;
    L.fb_words = 0;
    L.n_samples = 0;
    L.n_tiles = 0;
    L.n_raster = 0;
    uint64_t blocks = 0;
    for (uint32_t v = 0; v < n; v++) {
        const uint64_t P = view_fb_pixels(views[v]);
        L.fb_words += P * views[v].spp;
        if (views[v].temporal_mode == SGS_TEMPORAL_REPROJECT) L.n_samples += P;
        blocks += view_resolve_blocks(views[v]);
        L.n_tiles += (uint64_t)view_grid_w(views[v]) * view_grid_h(views[v]);
        L.n_raster += view_raster_groups(views[v]);
    }
    L.n_blocks = (uint32_t)blocks;
    const bool tiles = t.schedule == 0u;
    L.capacity = !tiles ? 0u : (t.max_tile_instances ? t.max_tile_instances : default_tile_instances(views, n, n_gaussians));
    const size_t tiles_n = tiles ? (size_t)L.n_tiles : 0u;
    Arena a;
    L.counters = a.take(sizeof(BatchCounters));
    L.ready = a.offset;
    a.offset += 4u * (size_t)t.queue_capacity;
    L.tile_count = a.offset;
    a.offset += 4u * tiles_n;
    L.tile_cursor = a.offset;
    a.offset += 4u * tiles_n;
    L.cleared_end = a.offset;
    a.offset = (a.offset + 255u) & ~size_t(255u);
    L.offsets = a.take(4u * (tiles_n + 1u));
    L.entry_buckets = a.take(2u * (size_t)L.capacity);    // depth bucket per instance
    L.entry_gaussians = a.take(4u * (size_t)L.capacity);  // Gaussian index per instance
    L.sorted_instances = a.take(4u * (size_t)L.capacity);
    L.sorted_keys = a.take(4u * (size_t)L.capacity);
    const size_t n_chunks = tiles ? ((size_t)n_gaussians + 255u) / 256u : 0u;
    L.centroids = a.take(16u * n_chunks);
    L.view_keys = a.take(8u * (size_t)n);
    L.cov3d = a.take(tiles ? 32u * (size_t)n_gaussians : 0u);
    L.pair_rect = a.take(tiles ? 8u * (size_t)n_gaussians * n : 0u);
    L.queue = a.take(sizeof(BigEntry) * (size_t)t.queue_capacity);
    L.table = a.take(table_bytes(n));
    L.fb = a.take(8u * L.fb_words);
    L.samples = a.take(16u * (L.n_samples ? L.n_samples : 1u));
    // spatial-reuse views, one at a time: one sample's samples and reuse weights (16 B per
    // framebuffer pixel each) and an intermediate history (24 B per image pixel)
    uint64_t reuse_fb = 0, reuse_px = 0;
    for (uint32_t v = 0; v < n; v++)
        if (views[v].spatial_reuse) {
            reuse_fb = std::max<uint64_t>(reuse_fb, view_fb_pixels(views[v]));
            reuse_px = std::max<uint64_t>(reuse_px, (uint64_t)views[v].width * views[v].height);
        }
    L.reuse_samples = a.take(16u * reuse_fb);
    L.reuse_weights = a.take(16u * reuse_fb);
    L.reuse_temp = a.take(24u * reuse_px);
    L.total = a.offset;
    return L;
}

// resolve level of each view: 1 + the level of the earlier view whose outputs it reads
std::vector<uint32_t> view_levels(const sgs_view* views, uint32_t n)
{
    std::vector<uint32_t> level(n, 0u);
    for (uint32_t i = 0; i < n; i++) {
        if (views[i].temporal_mode == SGS_TEMPORAL_RESET) continue;
        for (uint32_t j = 0; j < i; j++)
            if (views[j].out.image == views[i].history.image) level[i] = std::max(level[i], level[j] + 1u);
    }
    return level;
}

int check_views(const sgs_view* views, uint32_t n)
{
    if (!views && n) return fail(SGS_ERROR_INVALID_ARGUMENT, "views is NULL");
    uint64_t blocks = 0;
    for (uint32_t i = 0; i < n; i++) {
        const sgs_view& v = views[i];
        const std::string at = "view " + std::to_string(i) + ": ";
        if (v.width == 0 || v.height == 0) return fail(SGS_ERROR_INVALID_ARGUMENT, at + "width and height must be > 0");
        const uint64_t ss = view_ss(v);
        if ((uint64_t)v.width * ss > 65535u || (uint64_t)v.height * ss > 65535u || view_fb_pixels(v) >= (1ull << 32))
            return fail(SGS_ERROR_INVALID_ARGUMENT, at + "framebuffer too large");
        if (v.spp < 1 || v.spp > 4096) return fail(SGS_ERROR_INVALID_ARGUMENT, at + "spp must be in [1, 4096]");
        if (v.bin_size && bin_log2(v.bin_size) >= kBinSizes)
            return fail(SGS_ERROR_INVALID_ARGUMENT, at + "bin_size must be 0, 16, 32 or 64");
        if (!(v.focal_x != 0.0f) || !(v.focal_y != 0.0f)) return fail(SGS_ERROR_INVALID_ARGUMENT, at + "focal length is 0");
        if (v.temporal_mode < SGS_TEMPORAL_RESET || v.temporal_mode > SGS_TEMPORAL_REPROJECT)
            return fail(SGS_ERROR_INVALID_ARGUMENT, at + "invalid temporal mode " + std::to_string(v.temporal_mode));
        if (!v.out.image || !v.out.depth || !v.out.alpha || !v.out.weight_mass)
            return fail(SGS_ERROR_INVALID_ARGUMENT, at + "output pointers are NULL");
        if (v.blue_noise_numel != 0 && v.blue_noise_numel != 64u * 128u * 128u)
            return fail(SGS_ERROR_INVALID_ARGUMENT, at + "blue_noise must hold 0 or 64*128*128 elements");
        if (v.blue_noise_numel != 0 && !v.blue_noise) return fail(SGS_ERROR_INVALID_ARGUMENT, at + "blue_noise is NULL");
        if (v.spatial_reuse && (!v.k_buffer_primitives || !v.k_buffer_occlusion))
            return fail(SGS_ERROR_INVALID_ARGUMENT, at + "spatial reuse needs the k-buffer");
        if (v.temporal_mode != SGS_TEMPORAL_RESET &&
            (!v.history.image || !v.history.depth || !v.history.alpha || !v.history.weight_mass))
            return fail(SGS_ERROR_INVALID_ARGUMENT, at + "STATIC and REPROJECT need a history");
        if (v.temporal_mode != SGS_TEMPORAL_RESET &&
            (v.history.image == v.out.image || v.history.depth == v.out.depth || v.history.alpha == v.out.alpha ||
             v.history.weight_mass == v.out.weight_mass))
            return fail(SGS_ERROR_INVALID_ARGUMENT, at + "outputs must not alias the history");
        blocks += view_resolve_blocks(v);
    }
    if (blocks >= (1ull << 31)) return fail(SGS_ERROR_INVALID_ARGUMENT, "too many output pixels in one batch");
    // a view may read an earlier view's outputs as its history (all four buffers); no view
    // may write buffers another view writes or reads otherwise
    for (uint32_t i = 0; i < n; i++) {
        const sgs_history& o = views[i].out;
        for (uint32_t j = 0; j < n; j++) {
            if (j == i) continue;
            const sgs_history& oj = views[j].out;
            if (o.image == oj.image || o.depth == oj.depth || o.alpha == oj.alpha || o.weight_mass == oj.weight_mass)
                return fail(SGS_ERROR_INVALID_ARGUMENT,
                            "views " + std::to_string(i) + " and " + std::to_string(j) + " write the same outputs");
            if (views[j].temporal_mode == SGS_TEMPORAL_RESET) continue;
            const sgs_history& h = views[j].history;
            const int shared = (o.image == h.image) + (o.depth == h.depth) + (o.alpha == h.alpha) +
                               (o.weight_mass == h.weight_mass);
            if (shared == 0) continue;
            if (shared != 4 || i > j)
                return fail(SGS_ERROR_INVALID_ARGUMENT,
                            "view " + std::to_string(j) + " reads buffers of view " + std::to_string(i) +
                                "; a history may only be all outputs of an earlier view");
        }
    }
    // a k-buffer is updated sample by sample in view order: views may share one (consecutive
    // frames of a stream) only in resolve order, which view_levels keeps for history chains
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = i + 1; j < n; j++)
            if (views[i].spatial_reuse && views[j].spatial_reuse &&
                (views[i].k_buffer_primitives == views[j].k_buffer_primitives) != (views[i].k_buffer_occlusion ==
                                                                                    views[j].k_buffer_occlusion))
                return fail(SGS_ERROR_INVALID_ARGUMENT,
                            "views " + std::to_string(i) + " and " + std::to_string(j) + " share half a k-buffer");
    // one rasteriser workgroup per 16x16 tile of every view's bins (at least as many as bins)
    uint64_t groups = 0;
    for (uint32_t i = 0; i < n; i++) groups += view_raster_groups(views[i]);
    if (groups >= (1ull << 31)) return fail(SGS_ERROR_INVALID_ARGUMENT, "too many framebuffer tiles in one batch");
    return SGS_SUCCESS;
}

void encode_views(const sgs_view* views, uint32_t n, void* host)
{
    ViewDesc* d = static_cast<ViewDesc*>(host);
    uint32_t* prefix = reinterpret_cast<uint32_t*>(static_cast<char*>(host) + table_views_bytes(n));
    uint32_t* tile_prefix = prefix + n + 1u;
    uint32_t* raster_prefix = tile_prefix + n + 1u;
    uint64_t fb_off = 0, smp_off = 0;
    uint32_t block = 0, tile = 0, group = 0;
    const std::vector<uint32_t> level = view_levels(views, n);
    for (uint32_t i = 0; i < n; i++) {
        const sgs_view& v = views[i];
        ViewDesc x;
        std::memset(&x, 0, sizeof x);
        x.r1 = make_float4(v.w2c[0], v.w2c[1], v.w2c[2], v.w2c[3]);
        x.r2 = make_float4(v.w2c[4], v.w2c[5], v.w2c[6], v.w2c[7]);
        x.r3 = make_float4(v.w2c[8], v.w2c[9], v.w2c[10], v.w2c[11]);
        x.cam = make_fb_camera(v.width, v.height, v.focal_x, v.focal_y, v.center_x, v.center_y, v.near_plane,
                               v.far_plane, v.proper_antialiasing, v.supersampling_factor);
        x.cam_x = v.cam_position[0];
        x.cam_y = v.cam_position[1];
        x.cam_z = v.cam_position[2];
        x.focal_x = v.focal_x;
        x.focal_y = v.focal_y;
        x.center_x = v.center_x;
        x.center_y = v.center_y;
        x.width = v.width;
        x.height = v.height;
        x.spp = v.spp;
        x.frame0 = v.frame_idx;
        x.mode = v.temporal_mode;
        x.block_begin = block;
        x.tile_begin = tile;
        x.grid_w = view_grid_w(v);
        x.grid_h = view_grid_h(v);
        x.bin_log = view_bin_log(v);
        x.raster_begin = group;
        x.level = level[i];
        x.route = view_routed(v) ? 1u : 0u;
        x.reuse = v.spatial_reuse ? 1u : 0u;
        x.fb_offset = fb_off;
        x.samples_offset = smp_off;
        x.hist = { v.history.image, v.history.depth, v.history.alpha, v.history.weight_mass };
        x.out = { v.out.image, v.out.depth, v.out.alpha, v.out.weight_mass };
        std::memcpy(x.v2p, v.view_to_previous_view, sizeof x.v2p);
        d[i] = x;
        prefix[i] = block;
        tile_prefix[i] = tile;
        raster_prefix[i] = group;
        tile += x.grid_w * x.grid_h;
        group += (uint32_t)view_raster_groups(v);
        const uint64_t P = view_fb_pixels(v);
        fb_off += P * v.spp;
        if (v.temporal_mode == SGS_TEMPORAL_REPROJECT) smp_off += P;
        block += (uint32_t)view_resolve_blocks(v);
    }
    prefix[n] = block;
    tile_prefix[n] = tile;
    raster_prefix[n] = group;
}

int accumulate_tiles_single(const sgs_gaussians* g, const sgs_settings* s, const FbCamera& cam, uint64_t* fb,
                            uint32_t frame_idx, uint32_t bin_log, char* tile_ws, hipStream_t stream)
{
    const SingleTileLayout L = single_tile_layout(cam, g->n_gaussians, bin_log);
    // view table: everything but the camera pose from the host, the pose from the device
    // pointers (no host synchronisation)
    sgs_view v;
    std::memset(&v, 0, sizeof v);
    v.width = s->width;
    v.height = s->height;
    v.focal_x = s->focal_x;
    v.focal_y = s->focal_y;
    v.center_x = s->center_x;
    v.center_y = s->center_y;
    v.near_plane = s->near_plane;
    v.far_plane = s->far_plane;
    v.proper_antialiasing = s->proper_antialiasing;
    v.supersampling_factor = s->supersampling_factor;
    v.spp = 1;
    v.frame_idx = frame_idx;
    v.temporal_mode = SGS_TEMPORAL_RESET;
    v.bin_size = kRasterPixels << bin_log;
    std::vector<char> host(table_bytes(1));
    encode_views(&v, 1, host.data());
    const ViewDesc& d = *reinterpret_cast<const ViewDesc*>(host.data());
    SingleViewPrefixes prefixes;
    std::memcpy(prefixes.w, host.data() + table_views_bytes(1), sizeof prefixes.w);
    ViewDesc* desc = reinterpret_cast<ViewDesc*>(tile_ws + L.table);
    hipError_t e = launch_single_view_table(d, L.views_bytes, prefixes, s->w2c, s->cam_position, desc, stream);
    if (e != hipSuccess) return hip_fail(e, "view table");
    BatchCounters* counters = reinterpret_cast<BatchCounters*>(tile_ws + L.counters);
    e = launch_zero_u32(reinterpret_cast<uint32_t*>(counters), (L.cleared_end - L.counters) / 4u, stream);
    if (e != hipSuccess) return hip_fail(e, "counter clear");
    TileBuffers tb;
    tb.tile_count = reinterpret_cast<uint32_t*>(tile_ws + L.tile_count);
    tb.tile_cursor = reinterpret_cast<uint32_t*>(tile_ws + L.tile_cursor);
    tb.offsets = reinterpret_cast<uint32_t*>(tile_ws + L.offsets);
    tb.entries = { reinterpret_cast<uint16_t*>(tile_ws + L.entry_buckets),
                   reinterpret_cast<uint32_t*>(tile_ws + L.entry_gaussians) };
    tb.sorted_instances = reinterpret_cast<uint32_t*>(tile_ws + L.sorted_instances);
    tb.sorted_keys = reinterpret_cast<uint32_t*>(tile_ws + L.sorted_keys);
    tb.centroids = reinterpret_cast<float4*>(tile_ws + L.centroids);
    tb.view_keys = reinterpret_cast<uint2*>(tile_ws + L.view_keys);
    tb.cov3d = reinterpret_cast<float4*>(tile_ws + L.cov3d);
    tb.pair_rect = reinterpret_cast<uint2*>(tile_ws + L.pair_rect);
    tb.compact_records = (cam.width_i + 15u) / 16u <= 255u && (cam.height_i + 15u) / 16u <= 255u ? 1u : 0u;
    tb.n_tiles = L.n_tiles;
    tb.capacity = L.capacity;
    tb.n_raster = prefixes.w[5];
    tb.spp_min = tb.spp_max = 1u;  // one view of one sample
    const SceneArgs sc = scene_args(g);
    const uint32_t* dev_prefix = reinterpret_cast<const uint32_t*>(tile_ws + L.table + L.views_bytes);
    e = launch_batch_tiles(sc, desc, 1, dev_prefix + 2, dev_prefix + 4, tb, 8u, counters, fb, nullptr, stream);
    if (e != hipSuccess) return hip_fail(e, "tile schedule");
    // instance list overflow: the scatter schedule on a cleared framebuffer, gated on the flag
    e = launch_fill_u64_if(fb, (size_t)cam.width_i * cam.height_i, kEmpty, &counters->overflow, stream);
    if (e != hipSuccess) return hip_fail(e, "framebuffer clear");
    BatchTuning bt = resolve_batch_tuning(nullptr);
    bt.queue_capacity = L.queue_capacity;
    e = launch_batch_accumulate(sc, desc, 1, fb, counters, reinterpret_cast<BigEntry*>(tile_ws + L.queue),
                                reinterpret_cast<uint32_t*>(tile_ws + L.ready), bt, &counters->overflow, stream);
    if (e != hipSuccess) return hip_fail(e, "scatter fallback");
    return SGS_SUCCESS;
}

// Views with spatial reuse at resolve level l: every sample in order, as sgs_render's loop
// would run it (resolve.cu:355-444): sample_cu updates the view's k-buffer and writes the
// sample's colours and reuse weights, resolve_cu folds them into the previous sample's result
// (the view's history for sample 0, in the view's temporal mode; STATIC after). Results
// alternate between the view's outputs and a workspace buffer so the last lands in the outputs.
int resolve_reuse_views(const SceneArgs& sc, const sgs_view* views, uint32_t n_views, const char* table,
                        const std::vector<uint32_t>& level, uint32_t l, const uint64_t* fb, char* ws,
                        const BatchLayout& L, hipStream_t stream)
{
    uint64_t fb_off = 0;
    for (uint32_t i = 0; i < n_views; i++) {
        const sgs_view& v = views[i];
        const uint64_t P = view_fb_pixels(v);
        if (v.spatial_reuse && level[i] == l) {
            const FbCamera cam = make_fb_camera(v.width, v.height, v.focal_x, v.focal_y, v.center_x, v.center_y,
                                                v.near_plane, v.far_plane, v.proper_antialiasing, v.supersampling_factor);
            const char* d = table + (size_t)i * sizeof(ViewDesc);
            const float* w2c = reinterpret_cast<const float*>(d + offsetof(ViewDesc, r1));
            const float* cam_pos = reinterpret_cast<const float*>(d + offsetof(ViewDesc, cam_x));
            const float* v2p = reinterpret_cast<const float*>(d + offsetof(ViewDesc, v2p));
            float4* smp = reinterpret_cast<float4*>(ws + L.reuse_samples);
            float4* rw = reinterpret_cast<float4*>(ws + L.reuse_weights);
            const size_t Q = (size_t)v.width * v.height;
            float* tmp = reinterpret_cast<float*>(ws + L.reuse_temp);
            const OutputPtrs out = { v.out.image, v.out.depth, v.out.alpha, v.out.weight_mass };
            const OutputPtrs temp = { tmp, tmp + 3u * Q, tmp + 4u * Q, tmp + 5u * Q };
            HistoryPtrs hist = { v.history.image, v.history.depth, v.history.alpha, v.history.weight_mass };
            for (uint32_t k = 0; k < v.spp; k++) {
                hipError_t e = launch_sample(sc, w2c, cam_pos, cam, fb + fb_off + (uint64_t)k * P, smp, rw,
                                             reinterpret_cast<uint2*>(v.k_buffer_primitives),
                                             reinterpret_cast<float4*>(v.k_buffer_occlusion), true, stream);
                if (e != hipSuccess) return hip_fail(e, "batched spatial-reuse sample");
                const OutputPtrs o = ((v.spp - 1u - k) % 2u == 0u) ? out : temp;
                const int32_t mode = k == 0 ? v.temporal_mode : SGS_TEMPORAL_STATIC;
                e = launch_resolve(cam, v.width, v.height, v.focal_x, v.focal_y, v.center_x, v.center_y, smp, rw,
                                   hist, v2p, o, true, mode, stream);
                if (e != hipSuccess) return hip_fail(e, "batched spatial-reuse resolve");
                hist = { o.image, o.depth, o.alpha, o.weight_mass };
            }
        }
        fb_off += P * v.spp;
    }
    return SGS_SUCCESS;
}

}  // namespace

extern "C" {

size_t sgs_batch_view_table_bytes(uint32_t n_views) { return table_bytes(n_views); }

int sgs_batch_encode_views(const sgs_view* views, uint32_t n_views, void* host_table)
{
    int rc;
    if ((rc = check_views(views, n_views))) return rc;
    if (!host_table) return fail(SGS_ERROR_INVALID_ARGUMENT, "host_table is NULL");
    encode_views(views, n_views, host_table);
    return SGS_SUCCESS;
}

size_t sgs_render_batch_workspace_size(const sgs_view* views, uint32_t n_views, uint32_t n_gaussians,
                                       const sgs_batch_tuning* tuning)
{
    if (!views && n_views) return 0;
    return batch_layout(views, n_views, n_gaussians, resolve_batch_tuning(tuning)).total;
}

int sgs_render_batch(const sgs_gaussians* gaussians, const sgs_view* views, uint32_t n_views,
                     const void* view_table_device, const sgs_batch_tuning* tuning, void* workspace,
                     size_t workspace_bytes, void* const* stage_events, void* stream_ptr)
{
    int rc;
    if ((rc = check_gaussians(gaussians)) || (rc = check_views(views, n_views))) return rc;
    if (n_views == 0) return SGS_SUCCESS;
    const BatchTuning t = resolve_batch_tuning(tuning);
    if (t.entry_fragments < 64u) return fail(SGS_ERROR_INVALID_ARGUMENT, "entry_fragments must be >= 64");
    if (tuning && tuning->schedule > SGS_BATCH_SCATTER)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "unknown batch schedule " + std::to_string(tuning->schedule));
    if (tuning && tuning->tile_samples && tuning->tile_samples != 8u && tuning->tile_samples != 16u)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "tile_samples must be 8 or 16");
    const BatchLayout L = batch_layout(views, n_views, gaussians->n_gaussians, t);
    if (!workspace || workspace_bytes < L.total)
        return fail(SGS_ERROR_WORKSPACE_TOO_SMALL,
                    "workspace of " + std::to_string(L.total) + " bytes needed, got " + std::to_string(workspace_bytes));
    if (reinterpret_cast<uintptr_t>(workspace) % 256u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
    if (view_table_device && reinterpret_cast<uintptr_t>(view_table_device) % 16u != 0)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "view_table_device must be 16-byte aligned");

    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    char* ws = static_cast<char*>(workspace);
    const char* table = static_cast<const char*>(view_table_device);
    hipError_t e;
    if (!table) {
        std::string host(table_bytes(n_views), '\0');
        encode_views(views, n_views, host.data());
        e = hipMemcpyAsync(ws + L.table, host.data(), host.size(), hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return hip_fail(e, "view table upload");
        table = ws + L.table;
    }
    const ViewDesc* desc = reinterpret_cast<const ViewDesc*>(table);
    const uint32_t* prefix = reinterpret_cast<const uint32_t*>(table + table_views_bytes(n_views));
    const uint32_t* tile_prefix = prefix + n_views + 1u;
    const uint32_t* raster_prefix = tile_prefix + n_views + 1u;
    BatchCounters* counters = reinterpret_cast<BatchCounters*>(ws + L.counters);
    uint32_t* ready = reinterpret_cast<uint32_t*>(ws + L.ready);
    BigEntry* queue = reinterpret_cast<BigEntry*>(ws + L.queue);
    uint64_t* fb = reinterpret_cast<uint64_t*>(ws + L.fb);
    float4* samples = reinterpret_cast<float4*>(ws + L.samples);

    auto mark = [&](int i) -> hipError_t {
        return stage_events ? hipEventRecord(static_cast<hipEvent_t>(stage_events[i]), stream) : hipSuccess;
    };
    if ((e = mark(0)) != hipSuccess) return hip_fail(e, "stage event");
    e = launch_zero_u32(reinterpret_cast<uint32_t*>(counters), (L.cleared_end - L.counters) / 4u, stream);
    if (e != hipSuccess) return hip_fail(e, "counter clear");
    const SceneArgs sc = scene_args(gaussians);
    if (t.schedule == 0u) {
        // tiles write every framebuffer word; on overflow the scatter schedule runs on a
        // cleared framebuffer instead, both gated on the device flag
        if ((e = mark(1)) != hipSuccess) return hip_fail(e, "stage event");
        TileBuffers tb;
        tb.tile_count = reinterpret_cast<uint32_t*>(ws + L.tile_count);
        tb.tile_cursor = reinterpret_cast<uint32_t*>(ws + L.tile_cursor);
        tb.offsets = reinterpret_cast<uint32_t*>(ws + L.offsets);
        tb.entries = { reinterpret_cast<uint16_t*>(ws + L.entry_buckets),
                       reinterpret_cast<uint32_t*>(ws + L.entry_gaussians) };
        tb.sorted_instances = reinterpret_cast<uint32_t*>(ws + L.sorted_instances);
        tb.sorted_keys = reinterpret_cast<uint32_t*>(ws + L.sorted_keys);
        tb.centroids = reinterpret_cast<float4*>(ws + L.centroids);
        tb.view_keys = reinterpret_cast<uint2*>(ws + L.view_keys);
        tb.cov3d = reinterpret_cast<float4*>(ws + L.cov3d);
        tb.pair_rect = reinterpret_cast<uint2*>(ws + L.pair_rect);
        tb.compact_records = 1u;
        for (uint32_t v = 0; v < n_views; v++) {
            const uint64_t ss = views[v].supersampling_factor > 1 ? (uint64_t)views[v].supersampling_factor : 1u;
            if ((views[v].width * ss + 15u) / 16u > 255u || (views[v].height * ss + 15u) / 16u > 255u)
                tb.compact_records = 0u;
        }
        tb.n_tiles = (uint32_t)L.n_tiles;
        tb.capacity = L.capacity;
        tb.n_raster = (uint32_t)L.n_raster;
        tb.spp_min = 0xffffffffu;
        tb.spp_max = 0u;
        for (uint32_t v = 0; v < n_views; v++) {
            tb.spp_min = std::min(tb.spp_min, views[v].spp);
            tb.spp_max = std::max(tb.spp_max, views[v].spp);
        }
        e = launch_batch_tiles(sc, desc, n_views, tile_prefix, raster_prefix, tb, t.tile_samples, counters, fb,
                               stage_events ? static_cast<hipEvent_t>(stage_events[2]) : nullptr, stream);
        if (e != hipSuccess) return hip_fail(e, "batched tiles");
        e = launch_fill_u64_if(fb, L.fb_words, kEmpty, &counters->overflow, stream);
        if (e != hipSuccess) return hip_fail(e, "framebuffer clear");
        e = launch_batch_accumulate(sc, desc, n_views, fb, counters, queue, ready, t, &counters->overflow, stream);
        if (e != hipSuccess) return hip_fail(e, "batched accumulation");
    } else {
        e = launch_fill_u64(fb, L.fb_words, kEmpty, stream);
        if (e != hipSuccess) return hip_fail(e, "framebuffer clear");
        if ((e = mark(1)) != hipSuccess) return hip_fail(e, "stage event");
        if ((e = mark(2)) != hipSuccess) return hip_fail(e, "stage event");
        e = launch_batch_accumulate(sc, desc, n_views, fb, counters, queue, ready, t, nullptr, stream);
        if (e != hipSuccess) return hip_fail(e, "batched accumulation");
    }
    // views with GPS or blue noise: each sample by the single-view scatter kernels, as
    // sgs_render would (cleared framebuffer, the sample's frame, its blue-noise slice)
    {
        uint64_t fb_off = 0;
        for (uint32_t i = 0; i < n_views; i++) {
            const sgs_view& v = views[i];
            const uint64_t P = view_fb_pixels(v);
            if (view_routed(v)) {
                const FbCamera cam = make_fb_camera(v.width, v.height, v.focal_x, v.focal_y, v.center_x, v.center_y,
                                                    v.near_plane, v.far_plane, v.proper_antialiasing,
                                                    v.supersampling_factor);
                const float* w2c = reinterpret_cast<const float*>(table + (size_t)i * sizeof(ViewDesc) +
                                                                  offsetof(ViewDesc, r1));
                for (uint32_t smp = 0; smp < v.spp; smp++) {
                    uint64_t* fbs = fb + fb_off + (uint64_t)smp * P;
                    const uint32_t frame = v.frame_idx + smp;
                    e = launch_fill_u64(fbs, P, kEmpty, stream);
                    if (e != hipSuccess) return hip_fail(e, "framebuffer clear");
                    if (v.use_gps) {
                        e = launch_gps_scatter(sc, w2c, cam, fbs, frame, gps_max_points(cam, v.max_gps_points),
                                               t.sequential_threshold, BinArgs{}, nullptr, stream);
                    } else {
                        // one blue-noise slice per frame, rotated by the golden ratio per pass (render.cu:77-82)
                        e = launch_st_scatter(sc, w2c, cam, fbs, frame, v.blue_noise + (size_t)(frame % 64u) * 128u * 128u,
                                              (frame / 64u) * 0x9E3779B9u, t.sequential_threshold, BinArgs{},
                                              nullptr, stream);
                    }
                    if (e != hipSuccess) return hip_fail(e, "batched view scatter");
                }
            }
            fb_off += P * v.spp;
        }
    }
    if ((e = mark(3)) != hipSuccess) return hip_fail(e, "stage event");
    // sample 0 of REPROJECT views needs the full sample buffer (3x3 neighbourhood moments)
    uint64_t fb_off = 0, smp_off = 0;
    for (uint32_t i = 0; i < n_views; i++) {
        const sgs_view& v = views[i];
        const uint64_t P = view_fb_pixels(v);
        if (v.temporal_mode == SGS_TEMPORAL_REPROJECT && !v.spatial_reuse) {
            const FbCamera cam = make_fb_camera(v.width, v.height, v.focal_x, v.focal_y, v.center_x, v.center_y,
                                                v.near_plane, v.far_plane, v.proper_antialiasing,
                                                v.supersampling_factor);
            const char* d = table + (size_t)i * sizeof(ViewDesc);
            e = launch_sample(sc, reinterpret_cast<const float*>(d + offsetof(ViewDesc, r1)),
                              reinterpret_cast<const float*>(d + offsetof(ViewDesc, cam_x)), cam, fb + fb_off,
                              samples + smp_off, nullptr, nullptr, nullptr, false, stream);
            if (e != hipSuccess) return hip_fail(e, "batched sample");
        }
        if (v.temporal_mode == SGS_TEMPORAL_REPROJECT) smp_off += P;
        fb_off += P * v.spp;
    }
    const std::vector<uint32_t> level = view_levels(views, n_views);
    const uint32_t levels = 1u + *std::max_element(level.begin(), level.end());
    for (uint32_t l = 0; l < levels; l++) {
        e = launch_batch_resolve(sc, desc, n_views, prefix, L.n_blocks, fb, samples, l, stream);
        if (e != hipSuccess) return hip_fail(e, "batched resolve");
        if ((rc = resolve_reuse_views(sc, views, n_views, table, level, l, fb, ws, L, stream))) return rc;
    }
    if ((e = mark(4)) != hipSuccess) return hip_fail(e, "stage event");
    return SGS_SUCCESS;
}

int sgs_batch_tile_stats(const void* workspace, void* stream_ptr, uint64_t* n_instances, uint64_t* visible_pairs,
                         uint32_t* overflowed)
{
    if (!workspace || !n_instances || !visible_pairs || !overflowed)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "workspace and outputs must not be NULL");
    BatchCounters c;
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    hipError_t e = hipMemcpyAsync(&c, workspace, sizeof c, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "counter read");
    *n_instances = c.total_instances;
    *visible_pairs = c.visible_pairs;
    *overflowed = c.overflow;
    return SGS_SUCCESS;
}

}  // extern "C"

// ---- scene loading (sgs_scene.hip) ----
namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr uint64_t kMortonMax = 0x7fffffffull;  // the radix sort counts items in int

struct MortonLayout {
    size_t bounds, codes_in, codes_out, index_in, temp, temp_bytes, total;
};

MortonLayout morton_layout(uint64_t n)
{
    MortonLayout L;
    L.bounds = 0;
    L.codes_in = align256(L.bounds + 6 * sizeof(uint32_t));
    L.codes_out = align256(L.codes_in + n * sizeof(uint64_t));  // used when the caller wants no codes
    L.index_in = align256(L.codes_out + n * sizeof(uint64_t));
    L.temp = align256(L.index_in + n * sizeof(uint32_t));
    L.temp_bytes = sgs::morton_sort_temp_bytes(n);
    L.total = align256(L.temp + L.temp_bytes);
    return L;
}

bool valid_bases(uint32_t b) { return b == 1 || b == 4 || b == 9 || b == 16; }

int check_scene_out(const sgs_scene_arrays* o)
{
    if (!o) return fail(SGS_ERROR_INVALID_ARGUMENT, "output arrays are NULL");
    if (o->n_gaussians && (!o->means || !o->scales || !o->rotations || !o->opacities || !o->sh_coefficients))
        return fail(SGS_ERROR_INVALID_ARGUMENT, "an output array pointer is NULL");
    if (!valid_bases(o->total_sh_bases))
        return fail(SGS_ERROR_INVALID_ARGUMENT, "total_sh_bases must be 1, 4, 9 or 16 (SH degree 0-3)");
    return SGS_SUCCESS;
}

sgs::SceneOut scene_out(const sgs_scene_arrays* o)
{
    return {o->means, o->scales, o->rotations, o->opacities, o->sh_coefficients, o->n_gaussians, o->total_sh_bases};
}

}  // namespace

extern "C" {

int sgs_ply_decode(const void* records, uint64_t n_records, const sgs_ply_layout* layout, uint64_t first,
                   const sgs_scene_arrays* out, void* stream)
{
    if (int rc = check_scene_out(out)) return rc;
    if (!layout) return fail(SGS_ERROR_INVALID_ARGUMENT, "layout is NULL");
    if (layout->total_sh_bases != out->total_sh_bases)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "layout and output disagree on total_sh_bases");
    if (first > out->n_gaussians || n_records > out->n_gaussians - first)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "records first.." + std::to_string(first + n_records) +
                                                    " exceed the " + std::to_string(out->n_gaussians) + " output rows");
    if (n_records == 0) return SGS_SUCCESS;
    if (!records) return fail(SGS_ERROR_INVALID_ARGUMENT, "records is NULL");
    sgs::PlyDecodeTable t;
    t.stride = layout->stride;
    t.big_endian = layout->big_endian ? 1u : 0u;
    const uint32_t fields = 11u + 3u * layout->total_sh_bases;
    for (uint32_t f = 0; f < SGS_PLY_MAX_FIELDS; ++f) {
        t.offset[f] = 0;
        t.type[f] = SGS_SCALAR_FLOAT32;
        if (f >= fields) continue;
        const uint32_t type = layout->type[f];
        if (type < SGS_SCALAR_INT8 || type > SGS_SCALAR_FLOAT64)
            return fail(SGS_ERROR_INVALID_ARGUMENT, "field " + std::to_string(f) + ": unknown scalar type " +
                                                        std::to_string(type));
        static const uint32_t size[] = {0, 1, 1, 2, 2, 4, 4, 4, 8};
        if ((uint64_t)layout->offset[f] + size[type] > layout->stride)
            return fail(SGS_ERROR_INVALID_ARGUMENT, "field " + std::to_string(f) + " lies outside the record");
        t.offset[f] = layout->offset[f];
        t.type[f] = (uint8_t)type;
    }
    const hipError_t e =
        sgs::launch_ply_decode(records, n_records, t, first, scene_out(out), static_cast<hipStream_t>(stream));
    return e == hipSuccess ? SGS_SUCCESS : hip_fail(e, "ply decode");
}

size_t sgs_morton_order_workspace_size(uint64_t n) { return n > kMortonMax ? 0 : morton_layout(n).total; }

int sgs_morton_order(const float* means, uint64_t n, uint32_t* order, uint64_t* sorted_codes, void* workspace,
                     size_t workspace_bytes, void* stream)
{
    if (n > kMortonMax) return fail(SGS_ERROR_INVALID_ARGUMENT, "at most 2^31 - 1 means");
    if (n == 0) return SGS_SUCCESS;
    if (!means || !order) return fail(SGS_ERROR_INVALID_ARGUMENT, "means and order must not be NULL");
    const MortonLayout L = morton_layout(n);
    if (L.temp_bytes == 0) return fail(SGS_ERROR_HIP, "radix sort size query failed");
    if (!workspace || workspace_bytes < L.total)
        return fail(SGS_ERROR_WORKSPACE_TOO_SMALL, "workspace needs " + std::to_string(L.total) + " bytes");
    if ((uintptr_t)workspace & 255) return fail(SGS_ERROR_INVALID_ARGUMENT, "workspace must be 256-byte aligned");
    char* w = static_cast<char*>(workspace);
    uint64_t* codes_out = sorted_codes ? sorted_codes : reinterpret_cast<uint64_t*>(w + L.codes_out);
    sgs::MortonBuffers b;
    b.bounds = reinterpret_cast<uint32_t*>(w + L.bounds);
    b.codes_in = reinterpret_cast<uint64_t*>(w + L.codes_in);
    b.index_in = reinterpret_cast<uint32_t*>(w + L.index_in);
    b.temp = w + L.temp;
    b.temp_bytes = L.temp_bytes;
    const hipError_t e = sgs::launch_morton_order(means, n, order, codes_out, b, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? SGS_SUCCESS : hip_fail(e, "morton order");
}

int sgs_gather_gaussians(const sgs_gaussians* src, const uint32_t* order, const sgs_scene_arrays* dst, void* stream)
{
    if (int rc = check_scene_out(dst)) return rc;
    if (!src) return fail(SGS_ERROR_INVALID_ARGUMENT, "src is NULL");
    if (src->total_sh_bases != dst->total_sh_bases)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "src and dst disagree on total_sh_bases");
    if (dst->n_gaussians > src->n_gaussians)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "dst has more rows than src");
    if (dst->n_gaussians == 0) return SGS_SUCCESS;
    if (!order || !src->means || !src->scales || !src->rotations || !src->opacities || !src->sh_coefficients)
        return fail(SGS_ERROR_INVALID_ARGUMENT, "a source pointer or order is NULL");
    const sgs::SceneArgs sc{src->means, src->scales, src->rotations, src->opacities, src->sh_coefficients,
                            src->n_gaussians, src->total_sh_bases};
    const hipError_t e = sgs::launch_gather_rows(sc, order, scene_out(dst), static_cast<hipStream_t>(stream));
    return e == hipSuccess ? SGS_SUCCESS : hip_fail(e, "gather");
}

}  // extern "C"
