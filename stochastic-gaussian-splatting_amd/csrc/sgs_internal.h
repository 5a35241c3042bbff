// Host-side declarations shared by the translation units of libsgs.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "sgs.h"
#include "sgs_device.h"

namespace sgs {

// Tile schedule geometry: each view's instances are binned per framebuffer bin of
// (16 << b) x (16 << b) pixels, b = 0 .. kBinSizes - 1 chosen per view; the rasteriser keeps
// 16x16-pixel tiles in LDS, 4^b of them per bin, each reading its bin's depth-sorted list.
constexpr uint32_t kRasterPixels = 16;
constexpr uint32_t kRasterLog2 = 4;
constexpr uint32_t kBinSizes = 3;
// default bin of a view: 16 pixels for framebuffers of at most this many pixels (few bins
// with long lists otherwise, each read by four tiles), 32 pixels above
constexpr uint64_t kSmallBinMaxPixels = 1ull << 19;
// Batched resolve: one workgroup per 16x16 tile of image pixels, so a winner covering a
// pixel's neighbours above and below is shaded from the same CU's caches.
constexpr uint32_t kResolveTile = 16;
// bin_size (16, 32 or 64 pixels) -> b; kBinSizes for any other value
inline uint32_t bin_log2(uint32_t bin_size)
{
    return bin_size == 16u ? 0u : bin_size == 32u ? 1u : bin_size == 64u ? 2u : kBinSizes;
}

// Gaussian arrays as the kernels read them.
struct SceneArgs {
    const float* means;
    const float* scales;
    const float* rotations;
    const float* opacities;
    const float* sh;
    uint32_t n;
    uint32_t sh_bases;
    const float4* cov3d;  // cached view-independent projection terms (store_cov3d layout), or NULL
};

// EWA projection of Gaussian g from its raw parameters or, when the scene carries it, from
// the cached view-independent terms (bit-identical either way)
__device__ __forceinline__ bool project_scene(const SceneArgs& sc, uint32_t g, const float4 r1, const float4 r2,
                                              const float4 r3, const FbCamera& cam, Projected& o)
{
    if (sc.cov3d) return project_cached(sc.means, sc.cov3d, g, r1, r2, r3, cam, o);
    return project_gaussian(sc.means, sc.scales, sc.rotations, sc.opacities, g, r1, r2, r3, cam, o);
}

// Resolved scheduling knobs (zero fields of sgs_tuning replaced by defaults).
struct Tuning {
    uint32_t sequential_threshold;
    uint32_t binning_pixel_threshold;
    uint32_t binning_point_threshold;
    uint32_t max_binned;
    uint32_t tile_size;
    uint32_t max_tile_instances;
    uint32_t max_gps_points;  // 0 = reference heuristic
    bool tiles;               // single view: tile schedule when the settings allow it
    uint32_t bin_log;         // single-view tile schedule bins of (16 << bin_log) pixels
    uint64_t max_workspace_bytes;  // single view: scatter when the tile workspace exceeds it (0: no cap)
};

// One binned large Gaussian of the hybrid ST path (64 bytes, 16-byte aligned).
struct alignas(16) BinnedST {
    uint64_t fragment;
    float mean_x, mean_y;          // mean2d - 0.5 (scatter pixel convention)
    float conic_x, conic_y, conic_z, opacity;
    float power_threshold;
    uint32_t x0x1;                 // x0 | x1 << 16 (AABB, framebuffer pixels)
    uint32_t y0y1;                 // y0 | y1 << 16
    uint32_t pad[5];
};
static_assert(sizeof(BinnedST) == 64, "BinnedST layout");

// One binned Gaussian of the hybrid GPS path (64 bytes).
struct alignas(16) BinnedGPS {
    uint64_t fragment;
    float mean_x, mean_y;
    float L0, L1, L2, power_threshold;
    float conic_x, conic_y, conic_z, opacity_rcp;
    float opacity_dilog;
    uint32_t n_points;
    uint32_t seed_x, seed_y;       // make_seed(index, frame) before the count draw
};
static_assert(sizeof(BinnedGPS) == 64, "BinnedGPS layout");

// Device counters at the start of the hybrid scratch area.
struct HybridCounters {
    uint32_t n_binned;             // may exceed the capacity; min() is taken by readers
    uint32_t n_instances;          // hybrid ST tile instances (total, may exceed capacity)
    uint32_t overflow;             // non-zero: instance list overflowed, binned set is scattered
    uint32_t pad[61];
};

// Binning outputs of an accumulation kernel (all null when not binning).
struct BinArgs {
    HybridCounters* counters;
    void* binned;                  // BinnedST* or BinnedGPS*
    uint32_t capacity;
    uint32_t threshold;
};

struct HistoryPtrs {
    const float* image;
    const float* depth;
    const float* alpha;
    const float* weight_mass;
};
struct OutputPtrs {
    float* image;
    float* depth;
    float* alpha;
    float* weight_mass;
};

// One view of a batched render as the kernels read it (device table, 16-byte aligned).
struct alignas(16) ViewDesc {
    float4 r1, r2, r3;          // world-to-view rows
    FbCamera cam;               // framebuffer (subpixel) camera
    float cam_x, cam_y, cam_z;  // camera position
    float focal_x, focal_y, center_x, center_y;  // image pixels
    uint32_t width, height;     // image pixels
    uint32_t spp;               // samples rendered for this view
    uint32_t frame0;            // RNG frame of sample 0; sample s uses frame0 + s
    int32_t mode;               // temporal mode of sample 0; later samples are STATIC
    uint32_t block_begin;       // first resolve block of this view
    uint32_t tile_begin;        // first bin of this view in the batch's bin numbering
    uint32_t grid_w, grid_h;    // bins per row / column of this view's framebuffer
    uint32_t bin_log;           // bins of (16 << bin_log) pixels
    uint32_t raster_begin;      // first rasteriser workgroup (16x16 tile) of this view
    uint32_t level;             // resolve order: 1 + the level of the view whose output is this view's history
    uint32_t route;             // 1: samples accumulated one by one by the single-view scatter kernels
                                // (GPS, blue noise); the batch's binning, rasteriser and scatter skip it
    uint32_t reuse;             // 1: spatial reuse, resolved sample by sample (the batched resolve skips it)
    uint64_t fb_offset;         // first framebuffer word of sample 0 (samples are P words apart)
    uint64_t samples_offset;    // REPROJECT views: first float4 of the sample-0 buffer
    HistoryPtrs hist;
    OutputPtrs out;
    float v2p[16];
};

// Work-queue state of the persistent batched accumulation.
struct BatchCounters {
    uint32_t next_item;         // phase A: next (Gaussian chunk, view) item
    uint32_t done_items;        // phase A items finished (release after their queue writes)
    uint32_t produced;          // queue slots reserved
    uint32_t consumed;          // queue slots claimed by phase B waves
    uint32_t n_instances;       // (view, tile, Gaussian) instances of the tile schedule
    uint32_t overflow;          // instances exceed the capacity: the scatter schedule runs instead
    unsigned long long total_instances;  // 64-bit instance total of the counting pass
    unsigned long long visible_pairs;    // (view, Gaussian) pairs with at least one tile
    uint32_t order_count[17];   // bins per list-length class (bit length, 16 and above merged)
    uint32_t order_cursor[17];  // placement cursors of the sort's bin order, per class
    uint32_t pad[20];
};
static_assert(sizeof(BatchCounters) == 256, "BatchCounters layout");

// A fragment range of a large Gaussian in one view, deferred to phase B (64 bytes).
struct alignas(16) BigEntry {
    uint64_t fragment;
    float mean_x, mean_y;       // mean2d - 0.5
    float ce_a, ce_b, ce_c, opacity;
    uint32_t x0, y0, w, view;   // AABB origin and width, view index
    uint32_t f0, f1, pad0, pad1;  // fragment range [f0, f1) in row-major AABB order
};
static_assert(sizeof(BigEntry) == 64, "BigEntry layout");

struct BatchTuning {
    uint32_t schedule;          // 0: LDS tiles, 1: scatter with global atomics
    uint32_t max_tile_instances;
    uint32_t tile_samples;      // samples per LDS pass of the tile rasteriser (0: 8 or 16 by view, 8, 16)
    uint32_t sequential_threshold;
    uint32_t big_threshold;     // fragments above which a Gaussian goes to the queue
    uint32_t entry_fragments;   // fragments per queue entry
    uint32_t queue_capacity;
    uint32_t view_major;        // 0: chunk-major items (Gaussian parameters reused across views)
    uint32_t blocks_per_cu;     // persistent grid = CUs x blocks_per_cu (0: occupancy limit)
};

// Workspace carve-up. Every sub-buffer is 256-byte aligned.
struct Arena {
    size_t offset = 0;
    size_t take(size_t bytes)
    {
        const size_t at = offset;
        offset += (bytes + 255u) & ~size_t(255u);
        return at;
    }
};

FbCamera make_fb_camera(uint32_t width, uint32_t height, float focal_x, float focal_y, float center_x, float center_y,
                        float near_plane, float far_plane, int32_t proper_antialiasing, int32_t supersampling_factor);
uint32_t gps_max_points(const FbCamera& cam, uint32_t override_points);

// ---- launchers (sgs_accumulate.hip) ----
hipError_t launch_fill_u64(uint64_t* p, size_t n, uint64_t value, hipStream_t stream);
hipError_t launch_zero_u32(uint32_t* p, size_t n, hipStream_t stream);
hipError_t launch_st_scatter(const SceneArgs& sc, const float* w2c, const FbCamera& cam, uint64_t* fb, uint32_t frame,
                             const uint8_t* blue_noise_slice, uint32_t blue_noise_rotation, uint32_t seq_threshold,
                             const BinArgs& bin, uint64_t* counters, hipStream_t stream);
hipError_t launch_gps_scatter(const SceneArgs& sc, const float* w2c, const FbCamera& cam, uint64_t* fb, uint32_t frame,
                              uint32_t max_points, uint32_t seq_threshold, const BinArgs& bin, uint64_t* counters,
                              hipStream_t stream);

// ---- launchers (sgs_batch.hip) ----
hipError_t launch_batch_accumulate(const SceneArgs& sc, const ViewDesc* views, uint32_t n_views, uint64_t* fb,
                                   BatchCounters* counters, BigEntry* queue, uint32_t* ready, const BatchTuning& t,
                                   const uint32_t* run_if_set, hipStream_t stream);
// tile schedule: binning (count, scan, fill) and per-tile LDS rasterisation of every view's
// samples; the scatter schedule above runs instead (gated on counters->overflow) when the
// instance list overflows
// Instances of the tile schedule, (view, bin, Gaussian) each: the depth bucket of the per-bin
// sort (2 bytes, < 1024) and the Gaussian index (4 bytes) in separate arrays, so the sort's
// counting pass reads only the buckets.
struct EntryArrays {
    uint16_t* bucket;
    uint32_t* gaussian;
    __device__ __forceinline__ void store(uint32_t slot, uint32_t b, uint32_t g) const
    {
        bucket[slot] = (uint16_t)b;
        gaussian[slot] = g;
    }
};

struct TileBuffers {
    uint32_t* tile_count;       // [n_tiles] instances per tile (zeroed by the caller); after the scan, the sort's bin order
    uint32_t* tile_cursor;      // [n_tiles] fill cursors (zeroed by the caller)
    uint32_t* offsets;          // [n_tiles + 1] exclusive scan of tile_count
    EntryArrays entries;        // [capacity] depth bucket and Gaussian index of each (view, bin, Gaussian) instance
    uint32_t* sorted_instances; // [capacity] each tile's instances, nearest depth bucket first
    uint32_t* sorted_keys;      // [capacity] depth bucket floor at each tile batch start (begin + 256 m)
    float4* centroids;          // [n_chunks] mean position of each chunk of 256 Gaussians
    float4* cov3d;              // [2 * n] view-independent projection terms of each Gaussian
    uint2* pair_rect;           // [n_views * n] tile rectangle of each (view, Gaussian) pair
    uint32_t compact_records;   // 1: 4-byte records (every view's 16x16-tile grid within 255 x 255)
    uint2* view_keys;           // [n_views] depth-key bucketing of each view (kmin, shift)
    uint32_t n_tiles, capacity;  // bins of the batch; instance capacity
    uint32_t n_raster;           // rasteriser workgroups: 16x16 tiles of every view's bins
    uint32_t spp_min, spp_max;   // fewest and most samples of the batch's views
};
// tile_prefix / raster_prefix: first bin / rasteriser workgroup of each view, n_views + 1 each
hipError_t launch_batch_tiles(const SceneArgs& sc, const ViewDesc* views, uint32_t n_views, const uint32_t* tile_prefix,
                              const uint32_t* raster_prefix, const TileBuffers& tb, uint32_t layers,
                              BatchCounters* counters, uint64_t* fb, hipEvent_t raster_event, hipStream_t stream);
// the view-independent projection terms of every Gaussian (store_cov3d layout, 2 float4 each)
hipError_t launch_cov3d(const SceneArgs& sc, float4* cache, hipStream_t stream);
hipError_t launch_fill_u64_if(uint64_t* p, size_t n, uint64_t value, const uint32_t* run_if_set, hipStream_t stream);
// device view table of a single view whose world-to-view matrix and camera position are in
// device memory: `desc` with its rows and position taken from w2c / cam_position, then the
// prefixes (resolve blocks, bins, rasteriser workgroups: {0, count} each), the table layout
// of the batched renderer for one view, whose ViewDesc part is views_bytes long
struct SingleViewPrefixes {
    uint32_t w[6];
};
hipError_t launch_single_view_table(const ViewDesc& desc, size_t views_bytes, const SingleViewPrefixes& prefixes,
                                    const float* w2c, const float* cam_position, void* table, hipStream_t stream);

// ---- launchers (sgs_scene.hip) ----
struct SceneOut {
    float* means;
    float* scales;
    float* rotations;
    float* opacities;
    float* sh;
    uint64_t n;
    uint32_t sh_bases;
};
struct PlyDecodeTable {
    uint32_t stride, big_endian;
    uint32_t offset[SGS_PLY_MAX_FIELDS];
    uint8_t type[SGS_PLY_MAX_FIELDS];
};
struct MortonBuffers {
    uint32_t* bounds;     // [6] ordered-key minima and maxima
    uint64_t* codes_in;   // [n]
    uint32_t* index_in;   // [n]
    void* temp;           // radix-sort temporary storage
    size_t temp_bytes;
};
hipError_t launch_ply_decode(const void* records, uint64_t n_records, const PlyDecodeTable& table, uint64_t first,
                             const SceneOut& out, hipStream_t stream);
hipError_t launch_gather_rows(const SceneArgs& src, const uint32_t* order, const SceneOut& dst, hipStream_t stream);
size_t morton_sort_temp_bytes(uint64_t n);
hipError_t launch_morton_order(const float* means, uint64_t n, uint32_t* order, uint64_t* sorted_codes,
                               const MortonBuffers& ws, hipStream_t stream);

// ---- launchers (sgs_hybrid.hip) ----
size_t hybrid_st_scratch_bytes(uint32_t fb_width, uint32_t fb_height, const Tuning& t);
size_t hybrid_gps_scratch_bytes(const Tuning& t);
hipError_t run_hybrid_st(const SceneArgs& sc, const float* w2c, const FbCamera& cam, uint64_t* fb, uint32_t frame,
                         const uint8_t* bn_slice, uint32_t bn_rotation, const Tuning& t, void* scratch,
                         uint64_t* counters, hipStream_t stream);
hipError_t run_hybrid_gps(const SceneArgs& sc, const float* w2c, const FbCamera& cam, uint64_t* fb, uint32_t frame,
                          uint32_t max_points, const Tuning& t, void* scratch, uint64_t* counters,
                          hipStream_t stream);

// ---- launchers (sgs_resolve.hip) ----
hipError_t launch_batch_resolve(const SceneArgs& sc, const ViewDesc* views, uint32_t n_views,
                                const uint32_t* block_prefix, uint32_t n_blocks, const uint64_t* fb,
                                const float4* samples, uint32_t level, hipStream_t stream);
hipError_t launch_sample(const SceneArgs& sc, const float* w2c, const float* cam_position, const FbCamera& cam,
                         const uint64_t* fb, float4* samples, float4* reuse_weights, uint2* k_prim, float4* k_occ,
                         bool spatial_reuse, hipStream_t stream);
hipError_t launch_resolve(const FbCamera& cam, uint32_t width, uint32_t height, float focal_x, float focal_y,
                          float center_x, float center_y, const float4* samples, const float4* reuse_weights,
                          const HistoryPtrs& hist, const float* view_to_previous_view, OutputPtrs out,
                          bool spatial_reuse, int32_t temporal_mode, hipStream_t stream);
// sample + resolve fused (no spatial reuse, RESET/STATIC): reads the framebuffer directly
hipError_t launch_shade_resolve(const SceneArgs& sc, const float* cam_position, const FbCamera& cam, uint32_t width,
                                uint32_t height, const uint64_t* fb, const HistoryPtrs& hist, OutputPtrs out,
                                int32_t temporal_mode, hipStream_t stream);

}  // namespace sgs
