/*
 * pfhip.h - C ABI of pathfinder_amd, an MI355X (gfx950) implementation of
 * Pathfinder's D3D11-level compute rasterization path:
 *   bound -> dice -> bin -> propagate -> sort -> fill -> tile.
 *
 * Two levels, mirroring the reference:
 *
 *  1. The renderer seam. The reference drives its D3D11 level through
 *     RendererD3D11 (renderer/src/gpu/d3d11/renderer.rs) with four calls made
 *     by Renderer::render_command / end_scene (renderer/src/gpu/renderer.rs:
 *     350-460): upload_scene, prepare_tiles, prepare_and_draw_tiles,
 *     end_frame. pfh_context_* replaces RendererD3D11 plus the RendererCore
 *     state it uses (area LUT, mask pages, texture metadata, stats).
 *
 *  2. The scene front-end. pfh_scene_* mirrors Scene / SceneBuilder
 *     (renderer/src/scene.rs, renderer/src/builder.rs) and produces the same
 *     RenderCommand stream (renderer/src/gpu_data.rs:37-105) that the seam
 *     consumes, so a caller can hand over a whole scene.
 *
 * Conventions follow the reference's own C API (c/src/lib.rs): opaque
 * handles created by a create call and released by the matching destroy call, row-major
 * POD structs, caller-owned input arrays borrowed for the duration of a call.
 * Unlike the reference (which panics), every fallible call returns a
 * pfh_status; pfh_last_error() gives the message. A context is used from one
 * thread at a time, like Renderer (&mut self).
 *
 * All wire structs below are byte-identical to the #[repr(C)] structs of
 * renderer/src/gpu_data.rs, so a Rust binding can pass `slice.as_ptr()`.
 */
#ifndef PFHIP_H
#define PFHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFHIP_ABI_VERSION 2

typedef enum pfh_status {
    PFH_OK = 0,
    PFH_ERROR_INVALID_ARGUMENT = 1,
    PFH_ERROR_HIP = 2,          /* a HIP runtime call failed */
    PFH_ERROR_OUT_OF_MEMORY = 3,
    PFH_ERROR_STATE = 4,        /* call out of order (e.g. draw before prepare) */
    PFH_ERROR_NO_DEVICE = 5,    /* no gfx950 device visible */
    PFH_ERROR_LIMIT = 6         /* a reference format limit would be exceeded */
} pfh_status;

/* ------------------------------------------------------------ geometry */

/* c/src/lib.rs:134-177 */
typedef struct pfh_vector2f { float x, y; } pfh_vector2f;
typedef struct pfh_rect_f { pfh_vector2f origin, lower_right; } pfh_rect_f;
typedef struct pfh_color_u { uint8_t r, g, b, a; } pfh_color_u;
typedef struct pfh_color_f { float r, g, b, a; } pfh_color_f;
/* PFTransform2F: row-major 2x2 matrix + translation. */
typedef struct pfh_transform2f {
    float m00, m01, m10, m11;
    float tx, ty;
} pfh_transform2f;

/* ------------------------------------------------------------ wire formats
 * renderer/src/gpu_data.rs */

/* SegmentIndicesD3D11 (gpu_data.rs:178-181); flags builder.rs:47-48 */
#define PFH_CURVE_IS_QUADRATIC 0x80000000u
#define PFH_CURVE_IS_CUBIC 0x40000000u
typedef struct pfh_segment_indices { uint32_t first_point_index, flags; } pfh_segment_indices;

/* DiceMetadataD3D11 (gpu_data.rs:329-334) */
typedef struct pfh_dice_metadata {
    uint32_t global_path_id, first_global_segment_index, first_batch_segment_index, pad;
} pfh_dice_metadata;

/* TilePathInfoD3D11 (gpu_data.rs:299-309) */
typedef struct pfh_tile_path_info {
    int16_t tile_min_x, tile_min_y, tile_max_x, tile_max_y;
    uint32_t first_tile_index;
    uint16_t color;
    uint8_t ctrl;
    int8_t backdrop;
} pfh_tile_path_info;

/* PropagateMetadataD3D11 (gpu_data.rs:314-325); clip_path_index ~0 = none */
typedef struct pfh_propagate_metadata {
    int32_t tile_rect[4]; /* min_x, min_y, max_x, max_y (RectI) */
    uint32_t tile_offset, path_index, z_write, clip_path_index;
    uint32_t backdrop_offset, pad0, pad1, pad2;
} pfh_propagate_metadata;

/* BackdropInfoD3D11 (gpu_data.rs:409-414) */
typedef struct pfh_backdrop_info {
    int32_t initial_backdrop, tile_x_offset;
    uint32_t path_index;
} pfh_backdrop_info;

/* TileD3D11 (gpu_data.rs:279-288) */
typedef struct pfh_tile {
    int32_t next_tile_id, first_fill_id;
    int16_t alpha_tile_id_lo;
    int8_t alpha_tile_id_hi, backdrop_delta;
    uint16_t color;
    uint8_t ctrl;
    int8_t backdrop;
} pfh_tile;

/* Fill (gpu_data.rs:356-363): 8.8 fixed point within the tile + list link */
typedef struct pfh_fill { uint16_t from_x, from_y, to_x, to_y; uint32_t link; } pfh_fill;

/* MicrolineD3D11 (gpu_data.rs:418-428) */
typedef struct pfh_microline {
    int16_t from_x_px, from_y_px, to_x_px, to_y_px;
    uint8_t from_x_subpx, from_y_subpx, to_x_subpx, to_y_subpx;
    uint32_t path_index;
} pfh_microline;

/* AlphaTileD3D11 (gpu_data.rs:292-295) */
typedef struct pfh_alpha_tile { uint32_t alpha_tile_index, clip_tile_index; } pfh_alpha_tile;

/* SegmentsD3D11 (gpu_data.rs:171-174) */
typedef struct pfh_segments {
    const pfh_vector2f *points;
    uint32_t point_count;
    const pfh_segment_indices *indices;
    uint32_t index_count;
} pfh_segments;

/* PathSource (gpu_data.rs:143-148) */
#define PFH_PATH_SOURCE_DRAW 0u
#define PFH_PATH_SOURCE_CLIP 1u

/* TileBatchDataD3D11 + PrepareTilesInfoD3D11 (gpu_data.rs:123-168) and the
 * clip batch id of ClippedPathInfo. Arrays have path_count entries except
 * backdrops (backdrop_count = sum of tile-rect widths). */
typedef struct pfh_tile_batch {
    uint32_t batch_id;
    uint32_t path_source;
    uint32_t path_count, tile_count, segment_count;
    pfh_transform2f transform;
    const pfh_dice_metadata *dice_metadata;
    const pfh_tile_path_info *tile_path_info;
    const pfh_propagate_metadata *propagate_metadata;
    const pfh_backdrop_info *backdrops;
    uint32_t backdrop_count;
    int32_t clip_batch_id; /* -1: no clipped paths */
    /* DrawTileBatchD3D11.color_texture (gpu_data.rs; TileBatchTexture): the
     * colour texture page of a draw batch (-1: none) and its sampling flags */
    int32_t color_texture_page;
    uint32_t color_texture_sampling_flags;
} pfh_tile_batch;

/* TextureSamplingFlags (gpu/src/lib.rs:523-528) */
#define PFH_SAMPLING_REPEAT_U 0x01u
#define PFH_SAMPLING_REPEAT_V 0x02u
#define PFH_SAMPLING_NEAREST_MIN 0x04u
#define PFH_SAMPLING_NEAREST_MAG 0x08u

/* ColorCombineMode (gpu/blend.rs:204-212) */
#define PFH_COMBINE_NONE 0
#define PFH_COMBINE_SRC_IN 1
#define PFH_COMBINE_DEST_IN 2

/* Filter (renderer/src/paint.rs, effects.rs PatternFilter) and its parameters:
 *   PFH_FILTER_RADIAL_GRADIENT: line from x, y, line to x, y, radii r0, r1,
 *     uv origin x, y (Filter::RadialGradient {line, radii, uv_origin});
 *   PFH_FILTER_BLUR: direction (0 = x, 1 = y), sigma;
 *   PFH_FILTER_COLOR_MATRIX: the five columns of ColorMatrix (20 floats).
 * The text filter needs the gamma LUT and font rendering, which are not on
 * this path: entries with it are refused. */
#define PFH_FILTER_NONE 0
#define PFH_FILTER_RADIAL_GRADIENT 1
#define PFH_FILTER_BLUR 2
#define PFH_FILTER_COLOR_MATRIX 3
typedef struct pfh_filter {
    int32_t kind;
    float params[20];
} pfh_filter;

/* TextureMetadataEntry (gpu_data.rs): per paint, the transform from device
 * pixels to colour texture coordinates (color_0_transform), how the colour
 * texture combines with the base colour, the base colour, the filter, and
 * the blend mode (BlendMode::to_composite_ctrl picks the composite op). */
typedef struct pfh_texture_metadata_entry {
    pfh_transform2f color_0_transform;
    int32_t color_0_combine_mode;  /* PFH_COMBINE_* */
    pfh_color_u base_color;
    pfh_filter filter;
    int32_t blend_mode;            /* PFH_BLEND_MODE_* */
} pfh_texture_metadata_entry;

/* ------------------------------------------------------------ context */

typedef struct pfh_context pfh_context;

/* RendererLevel (gpu/options.rs:41-46). A context serves the seams of both
 * levels; the level names the one its owner builds scenes for. */
#define PFH_RENDERER_LEVEL_D3D9 1
#define PFH_RENDERER_LEVEL_D3D11 2

typedef struct pfh_context_options {
    int32_t device;               /* HIP device ordinal */
    int32_t level;                /* PFH_RENDERER_LEVEL_D3D9 or _D3D11 */
    const uint8_t *area_lut_rgba; /* 256x256 RGBA8 area LUT; NULL = built-in copy */
    uint32_t initial_microline_capacity; /* 0 = 16384 (d3d11/renderer.rs:43) */
    uint32_t initial_fill_capacity;      /* 0 = 16384 (d3d11/renderer.rs:44) */
    uint32_t enable_timing;       /* start with per-kernel timing on (PFH_PROFILE_ALL) */
} pfh_context_options;

/* Renderer::new (gpu/renderer.rs:181-340) + RendererD3D11::new. */
pfh_status pfh_context_create(const pfh_context_options *options, pfh_context **out);
void pfh_context_destroy(pfh_context *ctx);
const char *pfh_last_error(const pfh_context *ctx);

/* RendererOptions {dest, background_color} (gpu/options.rs:25-33) for one frame.
 * Renderer::begin_scene (gpu/renderer.rs:350-358). */
typedef struct pfh_frame_options {
    int32_t width, height;        /* destination framebuffer in pixels */
    int32_t has_background;       /* 1: first draw clears to background_color */
    pfh_color_f background_color;
} pfh_frame_options;
pfh_status pfh_begin_scene(pfh_context *ctx, const pfh_frame_options *frame);

/* UploadTextureMetadata: one solid base colour per paint (PaintMetadata
 * base_color, paint.rs:323-331; encoded as f16 per gpu/renderer.rs:702-775). */
pfh_status pfh_upload_texture_metadata(pfh_context *ctx, const pfh_color_u *base_colors,
                                       uint32_t paint_count);

/* UploadTextureMetadata (gpu/renderer.rs:702-775, compute_filter_params
 * :967-1049): each entry becomes ten RGBA16F texels, rounded to binary16 as
 * the reference rounds them (the composite reads these values). */
pfh_status pfh_upload_texture_metadata_entries(pfh_context *ctx,
                                               const pfh_texture_metadata_entry *entries,
                                               uint32_t count);

/* AllocateTexturePage / UploadTexelData (gpu_data.rs:37-105; gpu/renderer.rs
 * allocate_pattern_texture_page, upload_texel_data): RGBA8 colour texture
 * pages, zero when allocated; texels fill `rect` (origin, lower right) row by
 * row. Pages live until the context is destroyed or reallocated. */
typedef struct pfh_rect_i { int32_t min_x, min_y, max_x, max_y; } pfh_rect_i;
pfh_status pfh_allocate_texture_page(pfh_context *ctx, uint32_t page_id, int32_t width,
                                     int32_t height);
pfh_status pfh_upload_texel_data(pfh_context *ctx, uint32_t page_id, const pfh_rect_i *rect,
                                 const pfh_color_u *texels);

/* RendererD3D11::upload_scene (d3d11/renderer.rs:236-241). */
pfh_status pfh_upload_scene(pfh_context *ctx, const pfh_segments *draw_segments,
                            const pfh_segments *clip_segments);

/* RendererD3D11::prepare_tiles (d3d11/renderer.rs:427-542): PrepareClipTilesD3D11. */
pfh_status pfh_prepare_tiles(pfh_context *ctx, const pfh_tile_batch *batch);

/* RendererD3D11::prepare_and_draw_tiles (d3d11/renderer.rs:414-424): DrawTilesD3D11. */
pfh_status pfh_prepare_and_draw_tiles(pfh_context *ctx, const pfh_tile_batch *batch);

/* Renderer::end_scene (gpu/renderer.rs:426-460). Waits for the frame, grows
 * any buffer that overflowed and re-renders the frame if one did. Batch
 * buffers stay readable until the next pfh_begin_scene (the reference frees
 * them here, d3d11/renderer.rs:784-795). */
pfh_status pfh_end_scene(pfh_context *ctx);

/* Re-runs every batch of the last scene from device-resident inputs (no host
 * uploads, no host synchronisation). Used to time frames. */
pfh_status pfh_rerender_scene(pfh_context *ctx);

/* No counterpart in the reference: with graphs enabled, pfh_rerender_scene
 * captures the frame's launches on the context's three streams into a HIP
 * graph on its first call and replays the graph afterwards (per-kernel
 * timing keeps the launch-by-launch path). The graph is dropped at
 * pfh_begin_scene, when buffers grow, and when gradients are switched. */
pfh_status pfh_set_graphs(pfh_context *ctx, int32_t enable);

/* No counterpart in the reference (its dispatch sizes are fixed by the GL
 * driver): the launch shape of this context's frames. grid_cap bounds the
 * workgroups of the grid-stride kernels (batch init, dice, bound, bin,
 * propagate), fill_grid those of the fill kernel (one wavefront per 64
 * alpha tiles at a time). Defaults 16384 and 8192. */
typedef struct pfh_launch_shape {
    uint32_t grid_cap;
    uint32_t fill_grid;
} pfh_launch_shape;
pfh_status pfh_set_launch_shape(pfh_context *ctx, const pfh_launch_shape *shape);
pfh_status pfh_get_launch_shape(pfh_context *ctx, pfh_launch_shape *shape);

/* The autotuner: re-renders the last scene `frames` times per candidate
 * (1024, 2048, 4096, 8192, 16384 workgroups), first for fill_grid, then for
 * grid_cap, timing whole frames with events on the context stream; keeps
 * the fastest shape in the context and returns it in `chosen` (may be
 * NULL). It then times eager frames against replays of a captured graph of
 * the frame at that shape and keeps the faster for pfh_rerender_scene
 * (pfh_get_graphs reports the choice). Results do not depend on either. */
pfh_status pfh_autotune(pfh_context *ctx, uint32_t frames, pfh_launch_shape *chosen);
pfh_status pfh_get_graphs(pfh_context *ctx, int32_t *enabled);
pfh_status pfh_synchronize(pfh_context *ctx);

/* Reads the destination framebuffer, RGBA8, top row = scene row 0
 * (tile.cs.glsl:67-69 writes scene row y to image row H - y; this call
 * returns scene rows). stride in bytes (0 = width * 4). */
pfh_status pfh_read_framebuffer(pfh_context *ctx, uint8_t *rgba, size_t stride);

/* Debug readbacks of one batch of the last scene, in the reference layouts. */
typedef enum pfh_debug_kind {
    PFH_DEBUG_MICROLINES = 0,     /* pfh_microline[microline count] */
    PFH_DEBUG_FILLS = 1,          /* pfh_fill[fill count] */
    PFH_DEBUG_TILES = 2,          /* pfh_tile[tile count] */
    PFH_DEBUG_Z_BUFFER = 3,       /* int32[8 + fb tiles] (8-word header) */
    PFH_DEBUG_FIRST_TILE_MAP = 4, /* int32[fb tiles] */
    PFH_DEBUG_ALPHA_TILES = 5,    /* pfh_alpha_tile[batch alpha tile count] */
    PFH_DEBUG_BACKDROPS = 6,      /* pfh_backdrop_info[backdrop count], after bin */
    PFH_DEBUG_ALPHA_RANGE = 7,    /* uint32[2]: first and end alpha tile id */
    PFH_DEBUG_COVERAGE = 8,       /* float[256 per alpha tile of the batch]: the coverage
                                     checkpoint (pfh_set_gradients), pixel (x, y) of a
                                     tile at [(y & 3) * 64 + (y >> 2) * 16 + x] */
    PFH_DEBUG_COVERAGE_STATES = 9 /* uint8[64 per alpha tile of the batch]: the adjoint's
                                     states (PFH_CHECKPOINT_STATES / _RECOMPUTE), byte
                                     (y >> 2) * 16 + x, bits 2 (y & 3) (gate) and
                                     2 (y & 3) + 1 (C < 0) */
} pfh_debug_kind;
pfh_status pfh_read_debug(pfh_context *ctx, uint32_t batch_id, pfh_debug_kind kind, void *dst,
                          size_t capacity, size_t *out_len);
/* The frame's alpha mask: 64 RGBA8 texels per alpha tile, texel (x, y) of
 * tile `id` at word id*64 + y*16 + x (the reference stores the same texel at
 * (16*(id & 255) + x, 4*(id >> 8) + y) of a 4096-wide RGBA8 texture). */
pfh_status pfh_read_mask(pfh_context *ctx, uint32_t *dst, size_t capacity_words,
                         size_t *out_words);

/* RenderStats (gpu/perf.rs:20-44) plus the path-specific counts of the last
 * scene: microlines, columns, tiles linked into framebuffer-tile lists before
 * the z-buffer cull, tiles composited after it and how many of those have an
 * alpha mask, and alpha tiles that are clipped. */
typedef struct pfh_render_stats {
    uint64_t path_count, fill_count, alpha_tile_count, total_tile_count, drawcall_count;
    uint64_t microline_count, segment_count, column_count;
    uint64_t linked_tile_count, composited_tile_count, composited_alpha_tile_count;
    uint64_t clipped_alpha_tile_count;
    uint64_t gpu_bytes_allocated;
    uint32_t overflow_replays;
    uint32_t pad;
} pfh_render_stats;
pfh_status pfh_get_stats(pfh_context *ctx, pfh_render_stats *stats);

/* Per-kernel device timing: HIP events recorded around kernel launches on the
 * stream each kernel runs on (the reference times draw calls with GL timer queries per
 * TimeCategory, gpu/perf.rs:78-219). PFH_PROFILE_ONE times only `kernel`.
 * PFH_PROFILE_SERIAL times every kernel alone: frames re-rendered by
 * pfh_rerender_scene run every stage of every batch in order on the context
 * stream instead of overlapping batches and stages on the side streams (the
 * results are the same). */
#define PFH_PROFILE_OFF 0
#define PFH_PROFILE_ALL 1
#define PFH_PROFILE_ONE 2
#define PFH_PROFILE_SERIAL 3
typedef struct pfh_kernel_time {
    const char *name;
    uint64_t launches;
    double total_ms;
} pfh_kernel_time;
uint32_t pfh_profile_kernel_count(void);
const char *pfh_profile_kernel_name(uint32_t kernel);
pfh_status pfh_profile_set(pfh_context *ctx, int32_t mode, int32_t kernel);
pfh_status pfh_profile_reset(pfh_context *ctx);
/* Waits for the stream and returns totals since the last reset, one entry per kernel. */
pfh_status pfh_profile_read(pfh_context *ctx, pfh_kernel_time *out, uint32_t capacity,
                            uint32_t *count);

/* ------------------------------------------------------------ gradients */

/* No counterpart in the reference: the adjoint of fill + composite for
 * inverse rendering. The gradient is that of the continuous model the
 * kernels evaluate: per path and pixel the signed integral C of the winding
 * number over the pixel (exact area; the LUT coverage sum approximates it),
 * mask clamp(|C|, 0, 1) (min with the clip texel) and the src-over
 * composite; 8.8 fixed point, unorm8 masks and pixels and f16 paints pass
 * gradients through unchanged. Clip paths receive no gradient.
 *
 * pfh_set_gradients: before a frame, makes the fill kernel checkpoint C for
 * every alpha-tile pixel (256 floats per alpha tile, sized with the mask). */
pfh_status pfh_set_gradients(pfh_context *ctx, int32_t enable);

/* What the forward keeps for pfh_backward (no counterpart in the reference),
 * set outside a scene:
 *   PFH_CHECKPOINT_COVERAGE  the signed coverage sum C of every alpha-tile
 *                            pixel, 1 KiB per alpha tile (the debug read of
 *                            C needs it);
 *   PFH_CHECKPOINT_STATES    what the adjoint uses of C, 2 bits per pixel
 *                            (0 < |C| < 1 with no limiting clip; C < 0),
 *                            64 B per alpha tile, written by the same fill
 *                            (the default: smallest and fastest);
 *   PFH_CHECKPOINT_RECOMPUTE nothing: pfh_backward runs the draw batch's fill
 *                            again to write those states (64 B per alpha
 *                            tile of scratch) before the adjoint.
 * The three give the same gradients bit for bit. */
#define PFH_CHECKPOINT_COVERAGE 0
#define PFH_CHECKPOINT_STATES 1
#define PFH_CHECKPOINT_RECOMPUTE 2
pfh_status pfh_set_gradient_checkpoint(pfh_context *ctx, int32_t mode);

/* The last frame's image as floats (byte / 255), width x height x 4, rows
 * top down, into device memory `dst`. */
pfh_status pfh_framebuffer_to_float(pfh_context *ctx, float *dst);

/* Backward pass of the last frame, which must hold one draw batch composited
 * over a clear colour and have been rendered with gradients enabled.
 * grad_image (device): dL/d(image), width x height x 4 floats as
 * pfh_framebuffer_to_float lays them out. Accumulates (+=) into device
 * buffers: grad_points, 2 floats per point of the uploaded draw segments
 * (pfh_segments.points order, which repeats each contour's first point at
 * its end), and grad_colors, 4 floats (RGBA of the path's paint) per draw
 * path in push order; path_count bounds it. Waits for the result.
 * grad_image, grad_points and grad_colors must be ordinary (coarse-grained)
 * device memory from hipMalloc on the context's device, as torch allocates
 * it: the gradients are accumulated with hardware float atomics, which are
 * not guaranteed on fine-grained or host-mapped memory. */
pfh_status pfh_backward(pfh_context *ctx, const float *grad_image, float *grad_points,
                        uint32_t point_count, float *grad_colors, uint32_t path_count);

/* No counterpart in the reference: pfh_backward for callers that cannot
 * allocate device memory (code under HIP graph capture, tight inverse-
 * rendering loops, harnesses that forbid hipMalloc). Same arguments, checks
 * and results as pfh_backward; it allocates nothing: its scratch (list
 * offsets and their scan, per-slot transmittances, dL/dC per alpha tile and
 * its stamps) lives in one static device array of PFH_NOALLOC_SCRATCH_BYTES
 * per device, which each device holds until the process ends. Frames whose
 * scratch would not fit (framebuffer tiles, composited list slots and alpha
 * tiles; e.g. 512 x 512 with ~40k alpha tiles) return PFH_ERROR_LIMIT. Every
 * call on a device shares that device's array, so calls that overlap on one
 * GPU (two threads, two contexts) race; the array keeps the last call's
 * contents, and each call clears what must start at zero. */
#define PFH_NOALLOC_SCRATCH_BYTES (64u << 20)
pfh_status pfh_backward_noalloc(pfh_context *ctx, const float *grad_image, float *grad_points,
                                uint32_t point_count, float *grad_colors, uint32_t path_count);

/* ------------------------------------------------------------ scene front-end */

typedef struct pfh_scene pfh_scene;
typedef struct pfh_command_list pfh_command_list;

/* FillRule (content/src/fill.rs:18-23) */
#define PFH_FILL_RULE_WINDING 0
#define PFH_FILL_RULE_EVEN_ODD 1
/* BlendMode (content/src/effects.rs:99-160). SrcOver keeps the value 0; the
 * others follow the reference's declaration order. At the D3D11 level a draw
 * path's blend mode decides its tile rect (destructive modes use the view
 * box, builder.rs:430-436) and whether it writes the z-buffer (opaque paint
 * and SrcOver or Clear, effects.rs:202-233, builder.rs:83); the composite
 * itself is src-over for every mode, as in the reference, whose paint
 * metadata always carries SrcOver (paint.rs:573). */
#define PFH_BLEND_MODE_SRC_OVER 0
#define PFH_BLEND_MODE_CLEAR 1
#define PFH_BLEND_MODE_COPY 2
#define PFH_BLEND_MODE_SRC_IN 3
#define PFH_BLEND_MODE_SRC_OUT 4
#define PFH_BLEND_MODE_SRC_ATOP 5
#define PFH_BLEND_MODE_DEST_IN 6
#define PFH_BLEND_MODE_DEST_OUT 7
#define PFH_BLEND_MODE_DEST_OVER 8
#define PFH_BLEND_MODE_DEST_ATOP 9
#define PFH_BLEND_MODE_XOR 10
#define PFH_BLEND_MODE_LIGHTER 11
#define PFH_BLEND_MODE_DARKEN 12
#define PFH_BLEND_MODE_LIGHTEN 13
#define PFH_BLEND_MODE_MULTIPLY 14
#define PFH_BLEND_MODE_SCREEN 15
#define PFH_BLEND_MODE_HARD_LIGHT 16
#define PFH_BLEND_MODE_OVERLAY 17
#define PFH_BLEND_MODE_COLOR_DODGE 18
#define PFH_BLEND_MODE_COLOR_BURN 19
#define PFH_BLEND_MODE_SOFT_LIGHT 20
#define PFH_BLEND_MODE_DIFFERENCE 21
#define PFH_BLEND_MODE_EXCLUSION 22
#define PFH_BLEND_MODE_HUE 23
#define PFH_BLEND_MODE_SATURATION 24
#define PFH_BLEND_MODE_COLOR 25
#define PFH_BLEND_MODE_LUMINOSITY 26
#define PFH_BLEND_MODE_COUNT 27
/* PointFlags (content/src/outline.rs:50-60) */
#define PFH_POINT_CONTROL_0 0x01u
#define PFH_POINT_CONTROL_1 0x02u

/* Outline (content/src/outline.rs:34-60) as flat arrays: contour k spans
 * points [contour_ends[k-1], contour_ends[k]). Bounds are the union of every
 * point, control points included (outline.rs:560-568). */
typedef struct pfh_outline {
    const pfh_vector2f *points;
    const uint8_t *flags;
    uint32_t point_count;
    const uint32_t *contour_ends;
    uint32_t contour_count;
} pfh_outline;

pfh_scene *pfh_scene_create(void);
void pfh_scene_destroy(pfh_scene *scene);
/* Scene::set_view_box (scene.rs:222). */
void pfh_scene_set_view_box(pfh_scene *scene, const pfh_rect_f *view_box);
/* Scene::push_paint (scene.rs:191, palette dedup paint.rs:381-390). Returns the PaintId. */
uint32_t pfh_scene_push_paint(pfh_scene *scene, const pfh_color_u *color);
/* ------------------------------------------------------------ D3D9 level
 * RendererD3D9 (gpu/d3d9/renderer.rs:150-472): fills from the CPU tiler are
 * rasterized into a binary16 RGBA mask (each fragment added to its texel in
 * fill order, the sum rounded to binary16), clip tiles combine masks, and the
 * batch's tiles composite src-over into the RGBA8 target one tile at a time
 * (the target rounded to 8 bits after every tile). Solid paints, SrcOver. */

/* Fill (gpu_data.rs:356-363) as the raster level uses it: the line in 8.8
 * fixed point within the tile, and the alpha tile it belongs to. */
typedef struct pfh_fill_d3d9 {
    uint16_t from_x, from_y, to_x, to_y;
    uint32_t alpha_tile_index;
} pfh_fill_d3d9;

/* TileObjectPrimitive (gpu_data.rs:266-275). alpha_tile_id 0xFFFFFFFF: a
 * solid tile. ctrl: 1 nonzero, 2 even-odd (the mask's fill rule). */
typedef struct pfh_tile_object_d3d9 {
    int16_t tile_x, tile_y;
    uint32_t alpha_tile_id;
    uint32_t path_id;
    uint16_t color;
    uint8_t ctrl;
    int8_t backdrop;
} pfh_tile_object_d3d9;

/* Clip (gpu_data.rs:378-383): dest = min(|dest + dest_backdrop|, |src + src_backdrop|). */
typedef struct pfh_clip_d3d9 {
    uint32_t dest_tile_id;
    int32_t dest_backdrop;
    uint32_t src_tile_id;
    int32_t src_backdrop;
} pfh_clip_d3d9;

/* DrawTileBatchD3D9 (gpu_data.rs:227-237) for solid paints and SrcOver: the
 * tiles in draw order, the clip tiles, and z_buffer_data (DenseTileMap<i32>:
 * the rect in tiles and its values row by row; a tile whose path id is below
 * its z value is not drawn). */
typedef struct pfh_draw_tile_batch_d3d9 {
    const pfh_tile_object_d3d9 *tiles;
    uint32_t tile_count;
    const pfh_clip_d3d9 *clips;
    uint32_t clip_count;
    const int32_t *z_buffer;
    pfh_rect_i z_buffer_rect;
} pfh_draw_tile_batch_d3d9;

/* RendererD3D9::add_fills (:150-174): queues fills; alpha tiles 0 ..
 * max(alpha_tile_index) get mask storage. */
pfh_status pfh_d3d9_add_fills(pfh_context *ctx, const pfh_fill_d3d9 *fills, uint32_t count);
/* draw_buffered_fills (:176-184): rasterizes the queued fills into the mask. */
pfh_status pfh_d3d9_flush_fills(pfh_context *ctx);
/* clip_tiles + draw_tiles (:254-440) for one batch, inside a scene. */
pfh_status pfh_d3d9_draw_tiles(pfh_context *ctx, const pfh_draw_tile_batch_d3d9 *batch);

/* Gradient and pattern paints (Paint::from_gradient / from_pattern,
 * paint.rs:128-146; content/src/gradient.rs, content/src/pattern.rs). The
 * build renders each gradient into a row of a 256x256 colour texture page and
 * gives each image its own page (paint.rs:476-560, 789-846), and the paint's
 * TextureMetadataEntry maps device pixels to those texels (:603-638). */
typedef struct pfh_color_stop {
    float offset;
    pfh_color_u color;
} pfh_color_stop;

#define PFH_GRADIENT_LINEAR 0
#define PFH_GRADIENT_RADIAL 1
#define PFH_GRADIENT_WRAP_CLAMP 0
#define PFH_GRADIENT_WRAP_REPEAT 1

typedef struct pfh_gradient {
    int32_t geometry;             /* PFH_GRADIENT_LINEAR or PFH_GRADIENT_RADIAL */
    pfh_vector2f from, to;        /* the gradient line (radial: between the centres) */
    float radius0, radius1;       /* radial only */
    pfh_transform2f transform;    /* radial only: GradientGeometry::Radial.transform */
    const pfh_color_stop *stops;  /* in Gradient::add order; kept stably sorted by offset */
    uint32_t stop_count;
    int32_t wrap;                 /* PFH_GRADIENT_WRAP_* */
} pfh_gradient;

#define PFH_PATTERN_REPEAT_X 0x1u
#define PFH_PATTERN_REPEAT_Y 0x2u
#define PFH_PATTERN_NO_SMOOTHING 0x4u

typedef struct pfh_pattern {
    int32_t width, height;        /* Image size */
    const pfh_color_u *pixels;    /* width * height, row by row */
    pfh_transform2f transform;    /* Pattern::transform */
    uint32_t flags;               /* PFH_PATTERN_* */
    pfh_filter filter;            /* NONE, BLUR (params: direction != 0 for y, sigma)
                                     or COLOR_MATRIX */
} pfh_pattern;

/* PaintCompositeOp (paint.rs:87-93): how the overlay meets the base colour. */
#define PFH_PAINT_COMPOSITE_SRC_IN 0
#define PFH_PAINT_COMPOSITE_DEST_IN 1

/* Push a gradient or pattern paint with a base colour (Paint::from_gradient
 * and from_pattern use opaque white and SrcIn). Returns the PaintId, deduped
 * like pfh_scene_push_paint, or 0xFFFFFFFF for an invalid description
 * (unknown geometry, a pattern filter other than blur or colour matrix, an
 * empty or oversized image). */
uint32_t pfh_scene_push_gradient_paint(pfh_scene *scene, const pfh_color_u *base_color,
                                       const pfh_gradient *gradient, int32_t composite_op);
uint32_t pfh_scene_push_pattern_paint(pfh_scene *scene, const pfh_color_u *base_color,
                                      const pfh_pattern *pattern, int32_t composite_op);

/* In-place updates for callers that re-render one structure with moving
 * points and changing colours (inverse rendering); no counterpart in the
 * reference. set_draw_points replaces the coordinates of every draw path's
 * points (count must equal the points pushed) and recomputes the bounds;
 * set_paint_colors replaces the paint table with these solid colours, paint
 * i being colours[i] even where two are equal. */
pfh_status pfh_scene_set_draw_points(pfh_scene *scene, const pfh_vector2f *points,
                                     uint32_t count);
pfh_status pfh_scene_set_paint_colors(pfh_scene *scene, const pfh_color_u *colors,
                                      uint32_t count);

/* Scene::push_clip_path (scene.rs:98). clip_path = -1 for none. Returns ClipPathId. */
uint32_t pfh_scene_push_clip_path(pfh_scene *scene, const pfh_outline *outline, int32_t fill_rule,
                                  int32_t clip_path);
/* Scene::push_draw_path (scene.rs:77) with DrawPath::{set_clip_path,set_fill_rule,
 * set_blend_mode}. Returns DrawPathId. */
uint32_t pfh_scene_push_draw_path(pfh_scene *scene, const pfh_outline *outline, uint32_t paint,
                                  int32_t fill_rule, int32_t clip_path, int32_t blend_mode);
/* Bulk variants: path i owns contours [path_contour_ends[i-1], path_contour_ends[i]);
 * contour_ends index the shared points array. Per-path arrays may be NULL
 * (defaults: winding, no clip, src-over). Returns the first id. */
uint32_t pfh_scene_push_draw_paths(pfh_scene *scene, const pfh_vector2f *points,
                                   const uint8_t *flags, uint32_t point_count,
                                   const uint32_t *contour_ends, uint32_t contour_count,
                                   const uint32_t *path_contour_ends, uint32_t path_count,
                                   const uint32_t *paints, const int32_t *fill_rules,
                                   const int32_t *clip_paths, const int32_t *blend_modes);
uint32_t pfh_scene_push_clip_paths(pfh_scene *scene, const pfh_vector2f *points,
                                   const uint8_t *flags, uint32_t point_count,
                                   const uint32_t *contour_ends, uint32_t contour_count,
                                   const uint32_t *path_contour_ends, uint32_t path_count,
                                   const int32_t *fill_rules, const int32_t *clip_paths);
uint32_t pfh_scene_draw_path_count(const pfh_scene *scene);

/* The scene's paths as flat arrays (the layout pfh_scene_push_*_paths takes:
 * contour_ends are absolute point indices, path i owns contours
 * [path_contour_ends[i-1], path_contour_ends[i])). Pointers stay valid until
 * the scene is modified or exported again. */
typedef struct pfh_path_arrays {
    const pfh_vector2f *points;
    const uint8_t *flags;
    uint32_t point_count;
    const uint32_t *contour_ends;
    uint32_t contour_count;
    const uint32_t *path_contour_ends;
    uint32_t path_count;
    const uint32_t *paints;       /* draw paths only */
    const int32_t *fill_rules;
    const int32_t *clip_paths;
    const int32_t *blend_modes;   /* draw paths only */
} pfh_path_arrays;
pfh_status pfh_scene_path_arrays(pfh_scene *scene, uint32_t path_source, pfh_path_arrays *out);
pfh_status pfh_scene_paints(const pfh_scene *scene, const pfh_color_u **colors, uint32_t *count);
pfh_status pfh_scene_get_view_box(const pfh_scene *scene, pfh_rect_f *out);

/* SVG ingest: SVGScene::from_tree (svg/src/lib.rs:65-290) over this library's
 * own reader (paths, basic shapes, groups, transforms, solid fills and
 * strokes; gradients, clip paths, masks, text and images are skipped and
 * counted). Returns NULL on a malformed document, like
 * PFSVGSceneCreateWithMemory (c/src/lib.rs:804-815). info may be NULL. */
typedef struct pfh_svg_info {
    float view_box[4];            /* x, y, width, height */
    uint32_t fills, strokes;      /* draw paths pushed */
    uint32_t skipped_elements;    /* elements (and their subtrees) not supported */
    uint32_t unsupported_paints;  /* url() or unknown paints (drawn as none) */
} pfh_svg_info;
pfh_scene *pfh_svg_scene_create(const char *bytes, size_t byte_len, pfh_svg_info *info);

/* Stroking: OutlineStrokeToFill (content/src/stroke.rs) with StrokeStyle
 * {line_width, line_cap, line_join (miter limit)}. The outline is stroked in
 * its own space, then transformed (transform may be NULL); closed[k] != 0
 * joins contour k's end to its start (NULL: every contour closed). The
 * result is a fill outline for the nonzero rule. */
#define PFH_LINE_CAP_BUTT 0
#define PFH_LINE_CAP_SQUARE 1
#define PFH_LINE_CAP_ROUND 2
#define PFH_LINE_JOIN_MITER 0
#define PFH_LINE_JOIN_BEVEL 1
#define PFH_LINE_JOIN_ROUND 2
typedef struct pfh_stroke_style {
    float line_width;
    int32_t line_cap;
    int32_t line_join;
    float miter_limit;  /* for PFH_LINE_JOIN_MITER; <= 0 means 4 */
} pfh_stroke_style;
typedef struct pfh_outline_buffer pfh_outline_buffer;
pfh_status pfh_stroke_outline(const pfh_outline *outline, const uint8_t *closed,
                              const pfh_stroke_style *style, const pfh_transform2f *transform,
                              pfh_outline_buffer **out);
const pfh_outline *pfh_outline_buffer_view(const pfh_outline_buffer *buffer);
void pfh_outline_buffer_destroy(pfh_outline_buffer *buffer);

/* BuildOptions (options.rs:53-61): a 2D transform only. */
typedef struct pfh_build_options {
    pfh_transform2f transform;
} pfh_build_options;

/* SceneBuilder::build (builder.rs:148-222) into a command list. */
pfh_status pfh_scene_build(pfh_scene *scene, const pfh_build_options *options,
                           pfh_command_list **out);
void pfh_command_list_destroy(pfh_command_list *list);

/* RenderCommand kinds (gpu_data.rs:37-105) produced for the D3D11 level. */
typedef enum pfh_command_kind {
    PFH_COMMAND_START = 0,
    PFH_COMMAND_UPLOAD_TEXTURE_METADATA = 1,
    PFH_COMMAND_UPLOAD_SCENE_D3D11 = 2,
    PFH_COMMAND_PREPARE_CLIP_TILES_D3D11 = 3,
    PFH_COMMAND_DRAW_TILES_D3D11 = 4,
    PFH_COMMAND_FINISH = 5,
    PFH_COMMAND_ALLOCATE_TEXTURE_PAGE = 6,
    PFH_COMMAND_UPLOAD_TEXEL_DATA = 7,
    PFH_COMMAND_ADD_FILLS_D3D9 = 8,
    PFH_COMMAND_FLUSH_FILLS_D3D9 = 9,
    PFH_COMMAND_DRAW_TILES_D3D9 = 10
} pfh_command_kind;
uint32_t pfh_command_list_len(const pfh_command_list *list);
int32_t pfh_command_kind_at(const pfh_command_list *list, uint32_t index);
/* For PREPARE_CLIP_TILES_D3D11 / DRAW_TILES_D3D11: the batch (pointers stay
 * valid while the list lives). */
pfh_status pfh_command_batch(const pfh_command_list *list, uint32_t index, pfh_tile_batch *out);
/* For UPLOAD_SCENE_D3D11. */
pfh_status pfh_command_segments(const pfh_command_list *list, uint32_t index,
                                pfh_segments *draw, pfh_segments *clip);
/* For UPLOAD_TEXTURE_METADATA: base colours, one per paint. */
pfh_status pfh_command_paints(const pfh_command_list *list, uint32_t index,
                              const pfh_color_u **colors, uint32_t *count);
/* For UPLOAD_TEXTURE_METADATA: the TextureMetadataEntry of every paint
 * (create_texture_metadata, paint.rs:640-659). */
pfh_status pfh_command_texture_metadata(const pfh_command_list *list, uint32_t index,
                                        const pfh_texture_metadata_entry **entries,
                                        uint32_t *count);
/* For ALLOCATE_TEXTURE_PAGE (gpu_data.rs:45-49). */
pfh_status pfh_command_texture_page(const pfh_command_list *list, uint32_t index,
                                    uint32_t *page_id, int32_t *width, int32_t *height);
/* For UPLOAD_TEXEL_DATA (gpu_data.rs:50-54): the page, the rect and its
 * texels, row by row. */
pfh_status pfh_command_texel_data(const pfh_command_list *list, uint32_t index,
                                  uint32_t *page_id, pfh_rect_i *rect,
                                  const pfh_color_u **texels);

/* For ADD_FILLS_D3D9: one path's fills. */
pfh_status pfh_command_fills_d3d9(const pfh_command_list *list, uint32_t index,
                                  const pfh_fill_d3d9 **fills, uint32_t *count);
/* For DRAW_TILES_D3D9 (pointers stay valid while the list lives). */
pfh_status pfh_command_batch_d3d9(const pfh_command_list *list, uint32_t index,
                                  pfh_draw_tile_batch_d3d9 *out);

/* SceneBuilder::build for the D3D9 level (builder.rs:148-222 with
 * build_paths_on_cpu :224-318 and the D3D9 arm of
 * build_tile_batches_for_draw_path_display_item :998-1041): the tiler runs
 * here, on the calling thread and up to 15 more, and the list holds
 * ADD_FILLS_D3D9 per path with fills (clip paths, then draw paths, alpha
 * tiles numbered path by path), FLUSH_FILLS_D3D9 and one DRAW_TILES_D3D9.
 * Fails with PFH_ERROR_INVALID_ARGUMENT for a paint with a colour texture or
 * a blend mode other than SrcOver. A clip path's own clip path is ignored,
 * as the reference's tiler gets no built clip paths for clip paths. */
pfh_status pfh_scene_build_d3d9(pfh_scene *scene, const pfh_build_options *options,
                                pfh_command_list **out);

/* Renderer::render_command for every command of the list, between
 * pfh_begin_scene and pfh_end_scene (the caller brackets it). */
pfh_status pfh_render_commands(pfh_context *ctx, const pfh_command_list *list);

/* Scene::build_and_render (scene.rs:369-378): begin, build, render, end. */
pfh_status pfh_scene_build_and_render(pfh_scene *scene, pfh_context *ctx,
                                      const pfh_frame_options *frame,
                                      const pfh_build_options *options);
/* The same at the D3D9 level (pfh_scene_build_d3d9). */
pfh_status pfh_scene_build_and_render_d3d9(pfh_scene *scene, pfh_context *ctx,
                                           const pfh_frame_options *frame,
                                           const pfh_build_options *options);

#ifdef __cplusplus
}
#endif

#endif /* PFHIP_H */
