// pfh_kernels.h - launch interface of the gfx950 kernels (pfh_kernels.hip).
#ifndef PFH_KERNELS_H
#define PFH_KERNELS_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pfhip.h"

namespace pfh {

// Per-batch device counters. Written by the kernels, read by the host once
// per frame (pfh_end_scene) and by later kernels of the same batch.
// Tile rows per column segment of propagate's work list.
constexpr uint32_t kPropSegmentRows = 16;
// tile_status bits: the tile is linked into its framebuffer tile's list (a
// backdrop or an alpha tile), it occludes the tiles below it (a backdrop, no
// alpha tile and a path that writes z), it has an alpha tile
constexpr uint8_t kTileLinked = 1, kTileOccluder = 2, kTileAlpha = 4;

struct BatchCounters {
    uint32_t microlines;      // total microlines diced (may exceed capacity)
    uint32_t fills;           // total fills binned (may exceed capacity)
    uint32_t alpha;           // alpha tiles allocated by this batch
    uint32_t alpha_start;     // frame-global id of the batch's first alpha tile
    uint32_t microline_count; // min(microlines, capacity)
    uint32_t fill_count;      // min(fills, capacity)
    uint32_t overflow;        // bit 0 microlines, bit 1 fills, bit 2 mask pages
    uint32_t error;           // bit 0 bad segment/point index, bit 1 bad path index,
                              // bit 2 tile index outside the batch
    uint32_t frame_alpha_end; // frame alpha count after this batch
    uint32_t pad[7];
};

// Frame-level device state shared by the batches of one frame.
struct FrameState {
    uint32_t alpha_count;     // Renderer core alpha_tile_count (gpu/renderer.rs:146, 357)
    uint32_t pad[15];
};

struct DeviceScene {
    const float2 *points;
    const uint2 *indices;
    uint32_t point_count, index_count;
};

// Everything the kernels of one batch need.
struct BatchArgs {
    // inputs
    DeviceScene scene;
    const uint4 *dice_md;          // pfh_dice_metadata
    const uint4 *tile_path_info;   // pfh_tile_path_info
    const int32_t *prop_md;        // pfh_propagate_metadata, 12 words per path
    const int32_t *backdrops_init; // pfh_backdrop_info, 3 words per column
    uint32_t path_count, tile_count, segment_count, column_count;
    float xform[6];                // m00 m01 m10 m11 tx ty
    // clip batch (propagate / fill)
    const int32_t *clip_prop_md;
    const uint2 *clip_tiles;       // the clip batch's tile words 2-3
    uint32_t clip_path_count, clip_tile_count;
    // working buffers
    int32_t *backdrops;            // working copy, 3 words per column
    uint32_t *seg_offsets;         // segment_count + 1
    uint4 *microlines;             // microline_capacity
    uint32_t microline_capacity;
    uint32_t *ml_fill_offsets;     // microline_capacity + 1
    uint2 *fill_words;             // Fill words 0-1 (the 8.8 ends), fill_capacity
    uint32_t *fill_links;          // Fill word 2 (the list link), fill_capacity
    uint32_t *fill_tiles;          // tile index per fill
    uint32_t *path_fills;          // path_count + 1: first fill of each path
    uint32_t fill_capacity;
    uint2 *sorted_fills;           // fill words 0-1 sorted by (tile, fill index), fill_capacity
    uint2 *tile_runs;              // per tile with fills: its last run (first sorted index, count)
    uint2 *run_prev;               // at a run's first sorted index: the previous run or (0, 0)
    uint2 *tiles;                  // tile_count: words 2-3 (words 0 and 1 are kept in
                                   // tile_next and tile_first)
    uint32_t *tile_next;           // tile word 0 of linked tiles (k_sort_tiles; propagate for
                                   // linked tiles outside the framebuffer)
    uint32_t *tile_first;          // tile word 1 of tiles with fills (k_fill_link)
    uint8_t *tile_flags;           // tile_count: nonzero when bin gave the tile a fill (2 once
                                   // k_fill_link has written its head)
    uint32_t *tile_delta;          // tile_count: bin's backdrop deltas << 24 (tile word 2's top byte)
    uint8_t *tile_status;          // tile_count, from propagate: kTileLinked | kTileOccluder |
                                   // kTileAlpha, what k_sort_tiles reads of words 2-3
    uint32_t *col_alpha_offsets;   // column_count + 1
    uint32_t prop_seg_rows;        // segments (16 tile rows each) of the tallest column
    int32_t *seg_delta;            // [segment][column]: backdrop delta sum of the segment
    uint32_t *seg_alpha;           // [segment][column]: alpha tiles the segment allocates
    uint2 *alpha_tiles;            // tile_count
    uint2 *alpha_words;            // per alpha tile: its tile's words 2-3 (k_fill reads
                                   // them here, beside alpha_tiles, not at the tile)
    int32_t *zbuffer;              // 8 + fb tiles
    int32_t *first_tile_map;       // fb tiles
    uint32_t *row_paths;           // fb_height_tiles * path_count
    uint32_t *row_counts;          // fb_height_tiles
    uint4 *list_stats;             // per fb tile: linked, survivors, survivor alpha tiles
    uint32_t *scan_partials;
    BatchCounters *counters;
    FrameState *frame;
    // frame
    uint32_t *mask;                // 64 words per alpha tile
    uint32_t mask_capacity;        // alpha tiles
    const float4 *lut;             // padded area LUT (259 rows of 260 texels, byte / 255)
    const uint32_t *lut_sat;       // its saturated cells per row (pfh_lut_saturation)
    float *coverage;              // optional: 256 floats per alpha tile, the signed
                                   // coverage sum before abs/clamp (adjoint checkpoint)
    uint32_t *tile_paths;          // optional: the batch path of every tile (adjoint)
    uint8_t *cov_state;            // optional: per pixel 2 bits (the adjoint's gate
                                   // 0 < |C| < 1 and no limiting clip; C < 0), one byte
                                   // per lane, 64 B per alpha tile
    int32_t states_only;           // recompute pass: write cov_state only (no mask)
    int32_t fb_width_tiles, fb_height_tiles;
    // launch shape (pfh_launch_shape)
    uint32_t grid_cap;             // workgroups of the grid-stride kernels, at most
    uint32_t fill_grid;            // workgroups of k_fill, at most
};

struct TileArgs {
    const uint2 *tiles;            // tile words 2-3
    const uint32_t *tile_next;     // tile word 0 (the sorted list link)
    uint32_t tile_count;
    uint32_t mask_capacity;        // alpha tiles held by `mask`
    const int32_t *first_tile_map;
    const uint32_t *mask;
    const float4 *metadata;        // 10 float4 (f16-rounded texels) per paint, the
                                   // TextureMetadataEntry layout of gpu/renderer.rs:702-775
    uint32_t paint_count;
    // colour texture 0 of the batch (RGBA8 words, rows of color_width), or null
    const uint32_t *color_texture;
    int32_t color_width, color_height;
    uint32_t color_flags;          // TextureSamplingFlags
    float fb_width, fb_height;     // the draw viewport in pixels (uFramebufferSize)
    uint32_t *framebuffer;         // fb_width_tiles*16 x fb_height_tiles*16 RGBA8 words
    int32_t fb_width_tiles, fb_height_tiles;
    int32_t load;                  // 1: LOAD_ACTION_LOAD, 0: clear
    float clear[4];
    int32_t plain_paints;          // every paint's control word is zero (solid src-over)
    // with gradients on: the tile of every list slot, recorded as the
    // composite walks the lists (slot offsets: exclusive scan of the lists'
    // lengths), so that the adjoint reads them instead of walking again;
    // slots a walk does not reach hold ~0. Null otherwise.
    uint32_t *list_tiles;
    const uint32_t *list_offsets;  // fb tiles + 1
    uint32_t list_capacity;
};

// The adjoint of composite (k_tile) for one draw batch drawn over a clear
// colour: per framebuffer tile, the layers of the sorted list.
struct TileGradArgs {
    TileArgs f;                    // the forward composite's inputs
    const float *coverage;         // fill checkpoint, 256 per alpha tile, or
    const uint8_t *cov_state;      // the fill's 2-bit states (BatchArgs::cov_state)
    const uint2 *alpha_tiles;      // the batch's alpha tiles (tile, clip alpha id)
    uint32_t alpha_start;          // frame-global id of the batch's first alpha tile
    uint32_t alpha_count;
    const float *grad_image;       // dL/d(image): width x height RGBA floats, rows top down
    uint32_t width, height;
    const uint32_t *list_offsets;  // per fb tile: first slot of its list (exclusive scan)
    uint32_t slot_capacity;
    uint32_t *list_tiles;          // scratch: the tile of each list slot
    int32_t lists_recorded;        // list_tiles holds the forward composite's slots
    float *trans;                  // scratch: 256 floats per list slot
    const uint32_t *tile_paths;    // the batch path of every tile (k_bound)
    const uint4 *dice_md;          // the batch's paths: draw path id in .x
    uint32_t path_count;
    float *grad_cov;               // out: dL/d(signed coverage sum), 256 per alpha tile
    uint32_t *grad_stamp;          // out: epoch of each alpha tile grad_cov holds
    uint32_t epoch;
    float *grad_colors;            // out, accumulated: RGBA per draw path
    uint32_t color_count;          // draw paths held by grad_colors
};

// The adjoint of dice + coverage: per microline of the batch, the derivative
// of the pixel coverage sums it crosses with respect to its end points,
// carried to the scene's points.
struct MicrolineGradArgs {
    BatchArgs b;
    uint32_t microline_count;
    const float *grad_cov;         // dL/d(signed coverage sum), 256 per alpha tile
    const uint32_t *grad_stamp;    // alpha tiles whose stamp is not epoch read as zero
    uint32_t epoch;
    uint32_t alpha_capacity;
    float *grad_points;            // out, accumulated: 2 floats per scene point
};

// Every kernel launch of a frame, for per-kernel timing.
enum KernelId {
    K_BATCH_INIT, K_DICE_COUNT, K_DICE_SCAN, K_DICE_EMIT, K_BOUND, K_BIN_COUNT, K_BIN_SCAN,
    K_BIN_EMIT, K_FILL_LINK, K_PROPAGATE_COUNT, K_PROPAGATE_SCAN, K_PROPAGATE_EMIT,
    K_ROW_PATHS, K_SORT_TILES, K_FILL, K_TILE, K_COUNT
};
extern const char *const kKernelNames[K_COUNT];

// Brackets kernel launches with events on the launch stream.
class Profiler {
  public:
    virtual ~Profiler() {}
    virtual bool wants(int kernel) const = 0;
    virtual void begin(hipStream_t s, int kernel) = 0;
    virtual void end(hipStream_t s, int kernel) = 0;
};

void launch_batch_init(const BatchArgs &a, hipStream_t s, Profiler *p);
void launch_dice(const BatchArgs &a, hipStream_t s, Profiler *p);
void launch_bound(const BatchArgs &a, hipStream_t s, Profiler *p);
void launch_bin_count(const BatchArgs &a, hipStream_t s, Profiler *p);  // count, scan
void launch_bin_emit(const BatchArgs &a, hipStream_t s, Profiler *p);
void launch_fill_link(const BatchArgs &a, hipStream_t s, Profiler *p);  // per-tile fill lists
void launch_propagate(const BatchArgs &a, hipStream_t s, Profiler *p);
void launch_row_paths(const BatchArgs &a, hipStream_t s, Profiler *p);
void launch_sort(const BatchArgs &a, hipStream_t s, Profiler *p);
void launch_fill(const BatchArgs &a, hipStream_t s, Profiler *p);
void launch_tile(const TileArgs &a, hipStream_t s, Profiler *p);
void launch_frame_reset(FrameState *frame, hipStream_t s);
void launch_fill_zero(float *dst, size_t count, hipStream_t s);
void launch_list_lengths(const uint4 *list_stats, uint32_t *lengths, uint32_t fb_tiles,
                         hipStream_t s);
void launch_tile_backward(const TileGradArgs &a, hipStream_t s);
void launch_microline_backward(const MicrolineGradArgs &a, hipStream_t s);
void launch_framebuffer_to_float(const uint32_t *fb, uint32_t pitch_words, uint32_t width,
                                 uint32_t height, float *dst, hipStream_t s);
// Exclusive scan of n u32 values in place; out[n] receives the total.
void launch_scan(uint32_t *data, uint32_t n, uint32_t *partials, hipStream_t s);
void launch_lut_expand(const uint8_t *rgba, float4 *lut, hipStream_t s);
// binary16 metadata texels to floats (exact)
void launch_metadata_expand(const uint16_t *halves, float4 *metadata, uint32_t texels, hipStream_t s);

// The D3D9 level (gpu/d3d9/renderer.rs): mask raster, clip combine and
// tile composite. The mask holds 256 binary16 values per alpha tile.
struct D3D9FillArgs {
    const uint2 *fills;        // Fill words 0-1, grouped by alpha tile in fill order
    const uint4 *jobs;         // per job: alpha tile, first fill, end fill, 0
    uint32_t job_count;
    uint16_t *mask;
    uint32_t alpha_capacity;   // alpha tiles the mask holds
    const float4 *lut;
};
struct D3D9ClipArgs {
    const int4 *clips;         // Clip: dest, dest backdrop, src, src backdrop
    uint32_t clip_count;
    uint16_t *mask;
    uint32_t alpha_capacity;
};
struct D3D9TileArgs {
    const uint4 *tiles;        // TileObjectPrimitive words
    const uint32_t *list_offsets;  // per framebuffer tile, plus the end
    const uint32_t *list;      // tile indices per framebuffer tile, in draw order
    const uint16_t *mask;
    uint32_t alpha_capacity;
    const float4 *metadata;    // 10 texels per paint; texel 2 is the base colour
    uint32_t paint_count;
    uint32_t *framebuffer;
    int32_t fb_width_tiles, fb_height_tiles;
    int32_t load;
    float clear[4];
};
void launch_d3d9_fill(const D3D9FillArgs &a, hipStream_t s);
void launch_d3d9_clip(const D3D9ClipArgs &a, hipStream_t s);
void launch_d3d9_tile(const D3D9TileArgs &a, hipStream_t s);

// Scratch words needed by the scans of a batch.
size_t scan_partials_words(uint32_t max_elements);

}  // namespace pfh

#endif
