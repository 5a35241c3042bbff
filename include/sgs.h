/*
 * C ABI of the MI355X-native stochastic Gaussian splatting renderer (libsgs.so).
 *
 * Plain pointers and sizes only. Every pointer named "device" is HIP device memory of
 * the current device; the library never allocates: the caller passes outputs, persistent
 * state (k-buffer, history) and a workspace of at least the size the matching
 * *_workspace_size function reports. All work is enqueued on the caller's stream
 * (a hipStream_t passed as void*; NULL is the null stream) and no entry point
 * synchronises with the host.
 *
 * Every entry point returns SGS_SUCCESS (0) or a negative sgs_status; sgs_last_error()
 * then describes the failure (thread-local).
 *
 * Reference interface cited as file:line of fhahlbohm/stochastic-gaussian-splatting,
 * StochasticGSCudaBackend/StochasticGSCudaBackend/.
 */
#ifndef SGS_H
#define SGS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* k-buffer entries per (sub)pixel: resolve.h:9, bound as _C.k_buffer_size (render.cu:164). */
#define SGS_K_BUFFER_SIZE 4

typedef enum sgs_status {
    SGS_SUCCESS = 0,
    SGS_ERROR_INVALID_ARGUMENT = -1,
    SGS_ERROR_WORKSPACE_TOO_SMALL = -2,
    SGS_ERROR_HIP = -3,
} sgs_status;

/* TemporalMode (render.py:44-47, resolve.h:11-15). */
typedef enum sgs_temporal_mode {
    SGS_TEMPORAL_RESET = 0,
    SGS_TEMPORAL_STATIC = 1,
    SGS_TEMPORAL_REPROJECT = 2,
} sgs_temporal_mode;

/* Gaussian parameters (render.py:73-77, render.cu:17-21): device pointers to contiguous
 * float32 arrays. rotations must be 16-byte aligned. */
typedef struct sgs_gaussians {
    const float* means;           /* (N,3) world positions */
    const float* scales;          /* (N,3) log scales */
    const float* rotations;       /* (N,4) quaternion (r,i,j,k), unnormalized */
    const float* opacities;       /* (N,) logits */
    const float* sh_coefficients; /* (N,B,3) */
    uint32_t n_gaussians;
    uint32_t total_sh_bases;      /* B in {1,4,9,16} (render.cu:150) */
    const float* cov3d;           /* optional (N,8) cache of the view-independent projection terms
                                     from sgs_precompute_cov3d, 16-byte aligned; NULL: computed
                                     from scales, rotations and opacities in every call. It must
                                     be recomputed when those change. */
} sgs_gaussians;

/* Settings (render.py:8-42) in its as_tuple() order; w2c (4,4 row-major) and
 * cam_position (3,) are device pointers like the reference's CUDA tensors. */
typedef struct sgs_settings {
    const float* w2c;
    const float* cam_position;
    uint32_t width;
    uint32_t height;
    float focal_x;
    float focal_y;
    float center_x;
    float center_y;
    float near_plane;
    float far_plane;
    int32_t proper_antialiasing;
    int32_t supersampling_factor; /* values < 1 mean 1 (render.cu:52) */
    int32_t use_gps;
    int32_t hybrid_scheduling;
    int32_t spatial_reuse;
} sgs_settings;

/* History (render.py:49-53): device pointers, colour/depth premultiplied by alpha. */
typedef struct sgs_history {
    float* image;       /* (H,W,3) */
    float* depth;       /* (H,W,1) */
    float* alpha;       /* (H,W,1) */
    float* weight_mass; /* (H,W,1) */
} sgs_history;

/* Scheduling knobs, compile-time constants in the reference (scatter_utils.cuh:11,
 * accumulate_hybrid.cu:19-29, accumulate_hybrid_gps.cu:18-22). A zero field takes its
 * default; a NULL sgs_tuning* takes all defaults. None of them changes a result bit. */
typedef struct sgs_tuning {
    uint32_t sequential_threshold;    /* fragments/points one lane handles alone (default 8) */
    uint32_t binning_pixel_threshold; /* hybrid ST: AABB area above which a Gaussian is binned (512) */
    uint32_t binning_point_threshold; /* hybrid GPS: points above which a Gaussian is chunked (512) */
    uint32_t max_binned_gaussians;    /* binning capacity (min(N, 2^20)); overflow is scattered */
    uint32_t tile_size;               /* hybrid ST LDS tile edge: 16, 32 or 64 (32) */
    uint32_t max_tile_instances;      /* tile-instance capacity (2^24); overflow is scattered */
    uint32_t max_gps_points;          /* GPS points per Gaussian; 0 = rn(W*H/(2 ss^2)), the
                                         reference's cap (accumulate_scatter.cu:128). This one
                                         does change results when it binds. */
    uint32_t schedule;                /* stochastic transparency without hybrid scheduling: 0 or
                                         SGS_BATCH_TILES (1) bins the view into framebuffer bins
                                         and rasterises their 16x16 tiles in LDS, the batched
                                         renderer's tile schedule for one view and one sample;
                                         SGS_BATCH_SCATTER (2) scatters with global atomicMin as
                                         the reference does (accumulate_scatter.cu:19-77). Calls with blue
                                         noise or debug counters always scatter. Identical
                                         framebuffers either way. */
    uint32_t bin_size;                /* tile schedule: edge in pixels of the framebuffer bins
                                         whose depth-sorted instance lists the 16x16 raster tiles
                                         read: 16, 32 or 64 (default 16 for one view) */
    uint32_t max_workspace_mib;       /* cap, in MiB, on the workspace of the tile schedule: the
                                         tile schedule keeps, per (bin, Gaussian) instance, 14
                                         bytes (a depth-bucket sort of each bin's list, which the
                                         sorting-free reference does not have) plus 48 bytes per
                                         Gaussian; a view whose tile workspace would exceed the
                                         cap takes the scatter schedule instead, which needs only
                                         the framebuffer and per-subpixel samples, as the
                                         reference's (README.md:6). 0: no cap. Identical results
                                         either way. */
} sgs_tuning;

/* The load-time precompute the reference suggests (README.md:139-142): per Gaussian the
 * sigmoid opacity and the 3D covariance R diag(exp(2 s)) R^T of the normalised quaternion
 * (gs_utils.h:78-110), with log(255 opacity), as (N,8) float32 into the device array cov3d
 * (16-byte aligned; a Gaussian culled in every view, by opacity below 1/255 or a degenerate
 * quaternion, gets opacity -1). Pass it as sgs_gaussians.cov3d: every projection then reads
 * these 32 bytes instead of recomputing them from 32 bytes of raw parameters, with the same
 * floating-point operations, so renders are bit-identical with and without it. */
int sgs_precompute_cov3d(const sgs_gaussians* gaussians, float* cov3d, void* stream);

/* Replaces _C.k_buffer_size (render.cu:164). */
int sgs_k_buffer_size(void);

/* Thread-local description of the last failure of an sgs_* call on this thread. */
const char* sgs_last_error(void);

/* Bytes of device workspace sgs_render needs for these settings. */
size_t sgs_render_workspace_size(const sgs_settings* settings, uint32_t n_gaussians, const sgs_tuning* tuning);

/* Replaces _C.render (render.cu:16-45, 162-165; Python side render.py:72-95).
 * One stochastic sample of one view: clear the packed framebuffer, accumulate the
 * closest accepted fragment per (sub)pixel, shade it, resolve with the history.
 *   blue_noise        device (64,128,128) uint8 volume or NULL/0 elements for white noise
 *   k_buffer_*        device (K,P,2) int32 / (K,P,4) float32, updated in place; may be
 *                     NULL unless settings->spatial_reuse
 *   history           read in STATIC/REPROJECT modes (may be NULL in RESET)
 *   view_to_previous_view device (4,4) float32, read in REPROJECT mode
 *   out               fresh (H,W,*) outputs; must not alias history
 *   frame_idx         the sample's RNG frame (the reference's process-global counter,
 *                     render.cu:56-57, made explicit)
 * On return (stream order) the first 8*P bytes of the workspace hold the packed
 * framebuffer (float_bits(depth) << 32 | index, all ones = empty) of this sample. */
int sgs_render(const sgs_gaussians* gaussians, const sgs_settings* settings,
               const uint8_t* blue_noise, uint64_t blue_noise_numel,
               int32_t* k_buffer_primitives, float* k_buffer_occlusion,
               const sgs_history* history, const float* view_to_previous_view, int32_t temporal_mode,
               const sgs_history* out, uint32_t frame_idx, const sgs_tuning* tuning,
               void* workspace, size_t workspace_bytes, void* stream);

/* No counterpart in the reference. sgs_render for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same arguments and outputs as sgs_render without the workspace. It calls no hipMalloc,
 * hipFree or allocating library routine; its framebuffer (and the per-subpixel samples and
 * reuse weights REPROJECT and spatial reuse read) live in a fixed device array of
 * SGS_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs on until the process
 * ends. A view needing more (8 bytes per framebuffer pixel, +16 with REPROJECT or spatial
 * reuse, +16 more with spatial reuse: 1080p with spatial reuse fits, and 4K in RESET or STATIC) returns
 * SGS_ERROR_WORKSPACE_TOO_SMALL before anything is enqueued. It accumulates with the scatter
 * schedule (accumulate_scatter.cu:19-77, no binning buffers), whose framebuffer equals the
 * tile and hybrid schedules' bit for bit. Every call on one GPU shares that GPU's array:
 * calls that overlap in time on one GPU (other streams or threads) race; calls ordered on one
 * stream do not. */
#define SGS_NOALLOC_SCRATCH_BYTES (128u * 1024u * 1024u)
int sgs_render_noalloc(const sgs_gaussians* gaussians, const sgs_settings* settings,
                       const uint8_t* blue_noise, uint64_t blue_noise_numel,
                       int32_t* k_buffer_primitives, float* k_buffer_occlusion,
                       const sgs_history* history, const float* view_to_previous_view, int32_t temporal_mode,
                       const sgs_history* out, uint32_t frame_idx, const sgs_tuning* tuning, void* stream);

/* No reference counterpart (the reference only binds render): the accumulation stage
 * of sgs_render alone, into a caller framebuffer of (H*ss)*(W*ss) uint64 words that it
 * clears first. counters, if not NULL, is a device uint64[4] that receives
 * {fragments tested, atomics issued, Gaussians projected, Gaussians binned}; counting
 * adds a few instructions per fragment, so timed runs pass NULL. */
size_t sgs_accumulate_workspace_size(const sgs_settings* settings, uint32_t n_gaussians, const sgs_tuning* tuning);
int sgs_accumulate(const sgs_gaussians* gaussians, const sgs_settings* settings,
                   const uint8_t* blue_noise, uint64_t blue_noise_numel,
                   uint64_t* framebuffer, uint32_t frame_idx, const sgs_tuning* tuning,
                   uint64_t* counters, void* workspace, size_t workspace_bytes, void* stream);


/* ---------------------------------------------------------------------------------------
 * Batched multi-view rendering (no reference counterpart: the reference renders one view
 * per call from Python). Every view renders spp samples; view v's result equals spp
 * successive sgs_render calls with frames frame_idx .. frame_idx+spp-1, the first with
 * temporal_mode and the rest STATIC (identical framebuffer winners and outputs). Views of
 * stochastic transparency with white noise (hybrid scheduling or not: identical
 * framebuffers) accumulate all their samples in the batch's shared schedule; views with
 * GPS or blue noise accumulate each sample with the single-view scatter kernels into the
 * batch's framebuffers; views with spatial reuse are resolved sample by sample with their
 * k-buffer (the per-sample order the k-buffer's updates need), the others in one pass.
 * --------------------------------------------------------------------------------------- */
typedef struct sgs_view {
    float w2c[16];                    /* host copy, row-major world->view */
    float cam_position[3];            /* host copy */
    uint32_t width;
    uint32_t height;
    float focal_x;
    float focal_y;
    float center_x;
    float center_y;
    float near_plane;
    float far_plane;
    int32_t proper_antialiasing;
    int32_t supersampling_factor;
    uint32_t spp;                     /* samples of this view, >= 1 */
    uint32_t frame_idx;               /* RNG frame of sample 0 */
    int32_t temporal_mode;            /* mode of sample 0 */
    float view_to_previous_view[16];  /* host copy, read in REPROJECT mode */
    sgs_history history;              /* device, read unless RESET */
    sgs_history out;                  /* device outputs (H,W,*) */
    uint32_t bin_size;                /* tile schedule: edge in pixels of the framebuffer bins
                                         this view's instances are listed and depth-sorted per,
                                         read by the bins' 16x16 raster tiles: 16, 32 or 64; 0 =
                                         16 for framebuffers (W ss x H ss) of at most 2^19
                                         pixels, else 32. No result bit depends on it. */
    /* the single-view features of sgs_settings / sgs_render, per view (all zero: stochastic
     * transparency, white noise, no spatial reuse) */
    int32_t use_gps;
    int32_t hybrid_scheduling;
    int32_t spatial_reuse;
    uint32_t max_gps_points;          /* as sgs_tuning.max_gps_points (0: the reference's cap) */
    const uint8_t* blue_noise;        /* device (64,128,128) uint8 volume, or NULL */
    uint64_t blue_noise_numel;        /* 0 or 64*128*128 */
    int32_t* k_buffer_primitives;     /* spatial reuse: device (K,P,2) int32, updated in place */
    float* k_buffer_occlusion;        /* spatial reuse: device (K,P,4) float32 */
} sgs_view;

/* Scheduling of the batched accumulation; zero fields take defaults. None of them changes
 * a result bit. */
typedef struct sgs_batch_tuning {
    uint32_t sequential_threshold;    /* fragments one lane handles alone (8) */
    uint32_t big_threshold;           /* fragments above which a Gaussian is split into
                                         work-queue entries (4096) */
    uint32_t entry_fragments;         /* fragments per queue entry (2048) */
    uint32_t queue_capacity;          /* queue entries (2^20); overflow is scattered in place */
    uint32_t view_major;              /* 1: items ordered view by view; 0: Gaussian chunk by
                                         chunk across all views (default) */
    uint32_t blocks_per_cu;           /* persistent workgroups per CU (0: occupancy limit) */
    uint32_t schedule;                /* SGS_BATCH_TILES (default) or SGS_BATCH_SCATTER */
    uint32_t max_tile_instances;      /* capacity of the tile schedule's (view, tile, Gaussian)
                                         list, 14 bytes of workspace each (default per Gaussian
                                         and view 4 + 8 P / 2^21 for a framebuffer of P pixels,
                                         within [2^16, 2^31]); a batch that needs more runs the
                                         scatter schedule */
    uint32_t tile_samples;            /* samples held in LDS per pass over the instances of a
                                         view of several samples: 8 (passes over whole 16x16
                                         tiles) or 16 (passes over half tiles, two workgroups
                                         per tile); 0 (default): 8 for views of up to 8
                                         samples, 16 above */
} sgs_batch_tuning;

/* SGS_BATCH_TILES: every (view, Gaussian) is binned into the framebuffer bins (the view's
 * bin_size) its bounding box overlaps and each bin's list is depth-sorted; one workgroup per
 * 16x16 tile of a bin keeps the tile's words of up to 16 samples in LDS and writes them out
 * once. SGS_BATCH_SCATTER: one persistent launch over (Gaussian chunk, view) items with 64-bit
 * atomicMin into the framebuffer in device memory and a work queue for large Gaussians. */
#define SGS_BATCH_TILES 1u
#define SGS_BATCH_SCATTER 2u

/* Device workspace needed by sgs_render_batch for these views and n_gaussians Gaussians. */
size_t sgs_render_batch_workspace_size(const sgs_view* views, uint32_t n_views, uint32_t n_gaussians,
                                       const sgs_batch_tuning* tuning);

/* Bytes of the encoded view table, and the encoding itself into host memory (for callers
 * that upload it once, e.g. before capturing the render into a graph). */
size_t sgs_batch_view_table_bytes(uint32_t n_views);
int sgs_batch_encode_views(const sgs_view* views, uint32_t n_views, void* host_table);

/* Renders all views. A view's history may be the four output buffers of an earlier view of
 * the same batch (consecutive frames of one camera stream): such views are resolved after
 * the view they read, in stream order; any other overlap between one view's outputs and
 * another view's outputs or history is rejected. view_table_device: the encoded table already
 * in device memory, or
 * NULL to let the library copy it into the workspace (a pageable host-to-device copy,
 * not capturable into a graph). stage_events: NULL, or 5 hipEvent_t (as void*) recorded
 * on the stream before the counter and framebuffer clears, before the accumulation
 * (binning, for the tile schedule), before the accumulation's main kernel (the tile
 * rasteriser or the persistent scatter), after the accumulation and after the resolve, for
 * timing the stages. */
int sgs_render_batch(const sgs_gaussians* gaussians, const sgs_view* views, uint32_t n_views,
                     const void* view_table_device, const sgs_batch_tuning* tuning,
                     void* workspace, size_t workspace_bytes, void* const* stage_events, void* stream);

/* After sgs_render_batch on this workspace (waits for the stream): the number of
 * (view, tile, Gaussian) instances the tile schedule binned, the (view, Gaussian) pairs that
 * touched at least one tile, and whether the instances exceeded max_tile_instances (then
 * the scatter schedule produced the result). Lets a caller size max_tile_instances from a
 * previous frame. */
int sgs_batch_tile_stats(const void* workspace, void* stream, uint64_t* n_instances, uint64_t* visible_pairs,
                         uint32_t* overflowed);

/* ---------------------------------------------------------------------------------------
 * Scene loading: 3DGS PLY vertex records -> renderer arrays in Morton order. Replaces the
 * model loader StochasticGSModel.build (Model.py:22-79, one directory above the package):
 * plyfile column reads + torch.as_tensor(float32) + the SH re-layout (Model.py:24-75), and
 * morton_encode(means) + argsort + row indexing of every array (Model.py:32-34,43,53,58,
 * 69,74). The host reads the file (and its header); the device decodes and reorders.
 * --------------------------------------------------------------------------------------- */

/* Scalar types of PLY properties (char/uchar/short/ushort/int/uint/float/double). */
typedef enum sgs_scalar_type {
    SGS_SCALAR_INT8 = 1,
    SGS_SCALAR_UINT8 = 2,
    SGS_SCALAR_INT16 = 3,
    SGS_SCALAR_UINT16 = 4,
    SGS_SCALAR_INT32 = 5,
    SGS_SCALAR_UINT32 = 6,
    SGS_SCALAR_FLOAT32 = 7,
    SGS_SCALAR_FLOAT64 = 8,
} sgs_scalar_type;

/* Where the loader finds each field in one vertex record. Field order: x, y, z, scale_0..2,
 * rot_0..3, opacity, f_dc_0..2, f_rest_0..f_rest_{3(B-1)-1} (11 + 3B fields; the names of
 * Model.py:26-73). */
#define SGS_PLY_MAX_FIELDS 59
typedef struct sgs_ply_layout {
    uint32_t stride;          /* bytes per vertex record */
    uint32_t big_endian;      /* 1 for format binary_big_endian */
    uint32_t total_sh_bases;  /* B = 1 + (number of f_rest_* properties) / 3, in {1,4,9,16} */
    uint32_t offset[SGS_PLY_MAX_FIELDS]; /* byte offset of field f within a record */
    uint32_t type[SGS_PLY_MAX_FIELDS];   /* its sgs_scalar_type */
} sgs_ply_layout;

/* Writable renderer arrays (same layout as sgs_gaussians, device memory). */
typedef struct sgs_scene_arrays {
    float* means;
    float* scales;
    float* rotations;
    float* opacities;
    float* sh_coefficients;
    uint64_t n_gaussians;
    uint32_t total_sh_bases;
} sgs_scene_arrays;

/* Decodes n_records vertex records (device copy of the file's records first ..
 * first+n_records-1) into rows first.. of out: every field converted to float32 as
 * numpy's astype(float32) does, SH re-laid from the file's channel-major f_rest_k,
 * k = c*(B-1) + (b-1), to sh[row][b][c] (Model.py:61-75). Rows outside the range are not
 * touched, so a file can be decoded in chunks. */
int sgs_ply_decode(const void* records, uint64_t n_records, const sgs_ply_layout* layout, uint64_t first,
                   const sgs_scene_arrays* out, void* stream);

/* Morton order of n means (device (n,3) float32), n < 2^31. Code: the bounding box of the
 * finite coordinates, its largest side as the cube edge, 2^21 cells per axis, x in bit 3k,
 * y in 3k+1, z in 3k+2 (63 bits); order = stable sort of the codes (equal codes keep their
 * index order). Replaces morton_encode + torch.argsort (Model.py:32-33); the NeRFICG
 * morton_encode it calls is not in the reference tree, so this code is this library's own
 * definition. order: device (n,) uint32; sorted_codes: device (n,) uint64 or NULL. */
size_t sgs_morton_order_workspace_size(uint64_t n);
int sgs_morton_order(const float* means, uint64_t n, uint32_t* order, uint64_t* sorted_codes, void* workspace,
                     size_t workspace_bytes, void* stream);

/* dst row i = src row order[i] for all five arrays (Model.py:34,43,53,58,69,74: array[order]).
 * src and dst must not overlap; dst->n_gaussians rows are written (a row whose index is not
 * below src->n_gaussians is left untouched). */
int sgs_gather_gaussians(const sgs_gaussians* src, const uint32_t* order, const sgs_scene_arrays* dst, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SGS_H */
