/*
 * pf_oracle.h - CPU oracle for the D3D11-level compute rasterization path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product (pathfinder_amd/libpfhip.so) never links or calls it.
 *
 * It restates, with sequential-invocation semantics (invocations run in global
 * id order, so every atomic returns the value a sequential execution gives),
 * the seven compute shaders of the reference:
 *   shaders/d3d11/dice.cs.glsl, bound.cs.glsl, bin.cs.glsl, propagate.cs.glsl,
 *   sort.cs.glsl, fill.cs.glsl (+ fill_compute.inc.glsl, fill_area.inc.glsl),
 *   tile.cs.glsl (+ tile_fragment.inc.glsl:519-614, tile_vertex.inc.glsl:15-46)
 * driven the way renderer/src/gpu/d3d11/renderer.rs drives them
 * (prepare_tiles :427-542, draw_tiles :711-782).
 *
 * Word layouts are the reference's (renderer/src/gpu_data.rs):
 *   microline 4 x u32, fill 3 x u32, tile 4 x u32, alpha tile 2 x u32,
 *   z-buffer 8-word header + one i32 per framebuffer tile.
 * The alpha mask is kept linear by alpha-tile id: texel (lx, ly) of alpha
 * tile `id` is word id*64 + ly*16 + lx (RGBA8, r in the low byte); the
 * framebuffer is RGBA8 in scene rows (row y of the scene at row y).
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* What to read back from a batch. */
enum {
    ORC_MICROLINES = 0,     /* u32[4] per microline */
    ORC_FILLS = 1,          /* u32[3] per fill */
    ORC_TILES = 2,          /* u32[4] per tile */
    ORC_ZBUFFER = 3,        /* i32[8 + fb tiles] */
    ORC_FIRST_TILE_MAP = 4, /* i32[fb tiles] */
    ORC_ALPHA_TILES = 5,    /* u32[2] per alpha tile of the batch */
    ORC_BACKDROPS = 6,      /* i32[3] per column, after bin */
    ORC_ALPHA_RANGE = 7     /* u32[2]: first, end alpha tile id of the batch */
};

/* Semantic switches that the reference leaves undefined or gets wrong. */
enum {
    /* propagate.cs.glsl:121-122 reads clip metadata with a stride of two
     * uvec4; the struct is three uvec4 (gpu_data.rs:314-325). 0 = stride 3. */
    ORC_OPT_CLIP_STRIDE_LITERAL = 1,
    /* Planted mistakes, for tests that show a pin catches them (never set
     * outside those tests): 1 = twice dice's segment count, 2 = half of it
     * (rounded up), 3 = the last mix of sampleCurve with its ends swapped. */
    ORC_OPT_PLANT_DICE = 2,
    /* Planted paint mistakes, likewise: 1 = colour transform transposed,
     * 2 = bilinear weights of the two axes swapped, 3 = overlay's operands
     * unswapped, 4 = blur sum without the pair factor 2, 5 = hue taken from
     * the destination, 6 = radial gradient always on the larger root,
     * 7 = colour matrix transposed, 8 = combine DestIn as SrcIn. */
    ORC_OPT_PLANT_PAINT = 3
};

/* Per-stage wall-clock seconds (steady clock) accumulated since
 * orc_stage_reset, for the CPU baseline: the seven D3D11 stages, then the
 * D3D9 level's tiling, mask raster and composite. */
enum {
    ORC_STAGE_DICE, ORC_STAGE_BOUND, ORC_STAGE_BIN, ORC_STAGE_PROPAGATE, ORC_STAGE_FILL,
    ORC_STAGE_SORT, ORC_STAGE_TILE, ORC_STAGE_D3D9_TILER, ORC_STAGE_D3D9_RASTER,
    ORC_STAGE_D3D9_COMPOSITE, ORC_STAGE_COUNT
};
extern double orc_stage_seconds[ORC_STAGE_COUNT];
double orc_now(void);
void orc_stage_reset(void);
void orc_stage_times(double *out);

orc_ctx *orc_create(const uint8_t *area_lut_rgba /* 256*256*4 */);
void orc_destroy(orc_ctx *ctx);

/* Renderer::begin_scene (gpu/renderer.rs:350-358): resets the frame-global
 * alpha-tile count. has_background != 0 makes the first draw batch clear to
 * background_rgba (clear_color_for_draw_operation, :1190-1211). */
void orc_begin_frame(orc_ctx *ctx, int32_t width, int32_t height, int32_t has_background,
                     const float background_rgba[4]);

/* UploadTextureMetadata: one RGBA8 base colour per paint. Converted like
 * ColorU::to_f32 (color/src/lib.rs:70-73) and f16::from_f32
 * (gpu/renderer.rs:712-727). */
void orc_upload_paints(orc_ctx *ctx, const uint8_t *rgba8, uint32_t paint_count);

/* TextureMetadataEntry (renderer/src/gpu_data.rs) as the seam takes it;
 * byte-identical to pfh_texture_metadata_entry. filter_kind: 0 none,
 * 1 radial gradient (params: line from x, y, to x, y, radii r0, r1, uv
 * origin x, y), 2 blur (params: direction 0 x / 1 y, sigma), 3 colour
 * matrix (params: the five columns). combine: 0 none, 1 src-in, 2 dest-in.
 * blend_mode: PFH_BLEND_MODE_* numbering (SrcOver 0). */
typedef struct orc_metadata_entry {
    float transform[6];  /* m00, m01, m10, m11, tx, ty: device pixel -> colour UV */
    int32_t combine;
    uint8_t base_color[4];
    int32_t filter_kind;
    float filter_params[20];
    int32_t blend_mode;
} orc_metadata_entry;

/* UploadTextureMetadata (gpu/renderer.rs:702-775 with compute_filter_params
 * :967-1049): each entry becomes 10 RGBA16F texels. Returns 0, or -1 for an
 * unknown filter. */
int orc_upload_metadata(orc_ctx *ctx, const orc_metadata_entry *entries, uint32_t count);

/* AllocateTexturePage / UploadTexelData (gpu_data.rs:37-105): RGBA8 colour
 * texture pages, zeroed when allocated. */
void orc_allocate_page(orc_ctx *ctx, uint32_t page, int32_t width, int32_t height);
int orc_upload_texels(orc_ctx *ctx, uint32_t page, int32_t x, int32_t y, int32_t width,
                      int32_t height, const uint8_t *rgba8);

/* The colour texture of a draw batch (DrawTileBatchD3D11.color_texture:
 * TileBatchTexture {page, sampling_flags}); page < 0 for none. Sampling
 * flags: 1 repeat u, 2 repeat v, 4 nearest min, 8 nearest mag. */
int orc_set_batch_color_texture(orc_ctx *ctx, uint32_t batch_id, int32_t page,
                                uint32_t sampling_flags);

/* RendererD3D11::upload_scene (d3d11/renderer.rs:236-241). indices are
 * (first_point_index, flags) pairs (SegmentIndicesD3D11). */
void orc_upload_scene(orc_ctx *ctx,
                      const float *draw_points, uint32_t draw_point_count,
                      const uint32_t *draw_indices, uint32_t draw_segment_count,
                      const float *clip_points, uint32_t clip_point_count,
                      const uint32_t *clip_indices, uint32_t clip_segment_count);

/* prepare_tiles (d3d11/renderer.rs:427-542) for one TileBatchDataD3D11.
 * transform = {m00, m01, m10, m11, tx, ty} (row-major, like PFTransform2F).
 * dice_md: u32[4]/path, tile_path_info: u32[4]/path, prop_md: i32[12]/path,
 * backdrops: i32[3]/column. Returns 0 on success. */
int orc_prepare_batch(orc_ctx *ctx, uint32_t batch_id, uint32_t path_source,
                      uint32_t path_count, uint32_t tile_count, uint32_t segment_count,
                      const float transform[6],
                      const uint32_t *dice_md, const uint32_t *tile_path_info,
                      const int32_t *prop_md, const int32_t *backdrops,
                      uint32_t column_count, int32_t clip_batch_id);

/* draw_tiles (d3d11/renderer.rs:711-782) for a prepared batch. */
int orc_draw_batch(orc_ctx *ctx, uint32_t batch_id);

/* RendererD3D11::end_frame: drops every batch. */
void orc_end_frame(orc_ctx *ctx);

void orc_set_option(orc_ctx *ctx, int option, int value);

/* Element count of `what` for a batch (see the enum above), or -1. */
int64_t orc_count(orc_ctx *ctx, uint32_t batch_id, int what);
/* Copies `what` of a batch into dst (capacity in bytes). Returns bytes copied or -1. */
int64_t orc_read(orc_ctx *ctx, uint32_t batch_id, int what, void *dst, size_t cap);
/* Frame-global alpha mask, 64 RGBA8 words per alpha tile. */
uint32_t orc_alpha_tile_count(orc_ctx *ctx);
int64_t orc_read_mask(orc_ctx *ctx, uint32_t *dst, size_t cap_words);
/* Framebuffer, RGBA8, width*height words in scene rows. */
int64_t orc_read_framebuffer(orc_ctx *ctx, uint32_t *dst, size_t cap_words);

/* Threads for the order-independent loops (bound, sort, fill, tile). */
void orc_set_threads(int threads);

/* Helpers exposed for closed-form tests. */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
/* fill_area.inc.glsl computeCoverage for one (from, to) pair; out[4]. */
void orc_compute_coverage(const orc_ctx *ctx, float fx, float fy, float tx, float ty, float out[4]);

/* ------------------------------------------------------------ D3D9 level
 * (pf_d3d9.c): the CPU tiler and mask raster of the reference's D3D9 level,
 * a second whole-frame opinion that agrees with the D3D11 path only to a
 * tolerance. Paths are given as segments packed like SegmentsD3D11 (points,
 * {first point, flags} per segment) with the end of each path's segments. */
typedef struct orc_d3d9_paths {
    const float *points;         /* x, y per point */
    const uint32_t *indices;     /* first point, flags per segment */
    const uint32_t *path_seg_end;/* per path: one past its last segment */
    const uint32_t *paint;       /* per path (draw paths) */
    const int32_t *fill_rule;    /* per path: 0 nonzero, 1 even-odd */
    const int32_t *clip;         /* per path: clip path index or -1 */
    uint32_t n_paths;
} orc_d3d9_paths;

typedef struct orc_d3d9_scene {
    float view_box[4];           /* min x, min y, max x, max y */
    float transform[6];          /* m00 m01 m10 m11 tx ty */
    int32_t width, height, has_background;
    float background[4];
    const uint8_t *paints_rgba;
    uint32_t n_paints;
    orc_d3d9_paths draw, clip;
} orc_d3d9_scene;

/* Renders the scene into fb (RGBA8, ceil(w/16)*16 x ceil(h/16)*16 words,
 * scene rows) with the LUT of ctx. Returns 0 on success. */
int orc_d3d9_render(const orc_ctx *ctx, const orc_d3d9_scene *scene, uint32_t *fb, int threads);

/* The render commands of the D3D9 level before any raster: every fill
 * (Fill: LineSegmentU16 from x, y, to x, y, then the alpha tile: 3 words),
 * the draw batch's tiles (TileObjectPrimitive: 4 words), its clip tiles
 * (Clip: 4 words) and its z-buffer (the view box's tile rect, row by row).
 * Arrays are malloc'd; release them with orc_d3d9_built_free. */
typedef struct orc_d3d9_built {
    uint32_t *fills;
    uint32_t n_fills;
    uint32_t *tiles;
    uint32_t n_tiles;
    int32_t *clips;
    uint32_t n_clips;
    int32_t *z;
    int32_t z_rect[4];
    uint32_t n_alpha;
} orc_d3d9_built;
int orc_d3d9_build(const orc_d3d9_scene *scene, int threads, orc_d3d9_built *out);
void orc_d3d9_built_free(orc_d3d9_built *built);

#ifdef __cplusplus
}
#endif

#endif
