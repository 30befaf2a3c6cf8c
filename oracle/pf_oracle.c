/*
 * pf_oracle.c - CPU oracle for the D3D11-level compute rasterization path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see pf_oracle.h.
 *
 * Each stage below restates one compute shader of the reference with
 * sequential-invocation semantics: invocation i runs to completion before
 * invocation i+1 starts, so atomicAdd/atomicExchange/atomicMax return what a
 * sequential execution returns. That fixes the orders the reference leaves
 * to the GPU (microline order, fill order and per-tile fill lists, alpha-tile
 * ids, pre-sort tile lists).
 *
 * Floating point is evaluated in IEEE binary32 with no contraction
 * (built with -ffp-contract=off), and every GLSL built-in whose result the
 * GLSL spec leaves to the implementation is pinned to one formula:
 *   mix(a, b, t)          = a + t * (b - a)
 *   round(x)              = round half to even (Mesa lowers GLSL round() to
 *                           round_even)
 *   length(v)             = sqrt(v.x * v.x + v.y * v.y), sqrt correctly rounded
 *   uint(x) for x < 0     = 0 (saturating, like v_cvt_u32_f32)
 *   texture() on the LUT  = bilinear, CLAMP_TO_EDGE, fp32 weights a, b, each
 *                           lerp fused as llvmpipe's lp_build_lerp emits it:
 *                           lerp(w, x, y) = fma(w, y - x, x);
 *                           lerp(b, lerp(a, T00, T10), lerp(a, T01, T11))
 *   unorm8 load           = c / 255
 *   unorm8 store          = floor(clamp(x, 0, 1) * 255 + 0.5)
 * These choices are the contract the HIP kernels are held to bit for bit.
 *
 * Decisions where the reference is undefined or defective:
 *   - clip metadata is read with stride 3 (propagate.cs.glsl:121-122 uses 2;
 *     the struct is 3 uvec4, gpu_data.rs:314-325). ORC_OPT_CLIP_STRIDE_LITERAL
 *     switches to the literal stride for tests that pin the defect.
 *   - tiles outside the framebuffer tile grid are neither linked into the
 *     first-tile map nor written to the z-buffer (the reference indexes out of
 *     bounds for clip paths, which it does not intersect with the view box,
 *     builder.rs:1207).
 *   - the whole z-buffer starts at zero (the reference uploads tile_area zero
 *     words at offset 0, d3d11/renderer.rs:562-566, leaving the last eight z
 *     words stale); the header words the upload does zero are zeroed here too.
 *   - the framebuffer is kept in scene rows; tile.cs.glsl:67-69 stores row y
 *     at image row H - y (dropping scene row 0).
 */
#include "pf_oracle.h"

#include <math.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>

/* ORC_OPT_PLANT_PAINT (0: none); read by the paint functions */
static int g_plant_paint = 0;
#endif

#define TILE_SIZE 16
#define FLAG_QUAD 0x80000000u
#define FLAG_CUBIC 0x40000000u
#define TOLERANCE 0.25f
#define MICROLINE_LENGTH 16.0f
#define MAX_ITERATIONS 1024u
#define MAX_BATCHES 4096

typedef struct { float x, y; } v2;

typedef struct {
    int used;
    uint32_t batch_id, path_source, path_count, tile_count, segment_count, column_count;
    int32_t clip_batch_id;
    float xform[6];
    uint32_t *dice_md;   /* 4 per path */
    uint32_t *tpi;       /* 4 per path */
    int32_t *prop_md;    /* 12 per path */
    int32_t *backdrops;  /* 3 per column */
    uint32_t *microlines; uint32_t n_microlines, cap_microlines;
    uint32_t *fills; uint32_t n_fills, cap_fills;
    uint32_t *tiles;     /* 4 per tile */
    int32_t *zbuffer;    /* 8 + fb tiles */
    int32_t *ftm;        /* fb tiles */
    uint32_t *alpha;     /* 2 per alpha tile */
    uint32_t n_alpha, alpha_start;
    int32_t color_page;        /* colour texture 0 of a draw batch, -1 none */
    uint32_t color_flags;      /* its TextureSamplingFlags */
} orc_batch;

typedef struct {
    int32_t width, height;     /* 0: not allocated */
    uint8_t *rgba;
} orc_page;

struct orc_ctx {
    float lut[256 * 256 * 4]; /* texel (u, v) channel c at (v * 256 + u) * 4 + c */
    int32_t width, height, fbw, fbh;
    int has_background, dest_dirty;
    float background[4];
    uint16_t *paints; uint32_t paint_count;   /* f16 rgba per paint */
    float *metadata; uint32_t metadata_count; /* 40 f16-rounded floats per entry */
    orc_page *pages; uint32_t page_count;
    float *draw_points; uint32_t n_draw_points;
    uint32_t *draw_indices; uint32_t n_draw_segments;
    float *clip_points; uint32_t n_clip_points;
    uint32_t *clip_indices; uint32_t n_clip_segments;
    uint32_t *mask; uint32_t mask_cap_tiles;  /* 64 words per alpha tile */
    uint32_t alpha_tile_count;
    uint32_t *fb; /* fbw*16 x fbh*16 RGBA8 words */
    orc_batch batches[MAX_BATCHES];
    int opt_clip_stride_literal;
    int opt_plant_dice;
};

static int g_threads = 1;
void orc_set_threads(int threads) { g_threads = threads < 1 ? 1 : threads; }

/* Wall-clock seconds per stage since the last orc_stage_reset (CPU baseline
 * timing): ORC_STAGE_* in pf_oracle.h. */
double orc_stage_seconds[ORC_STAGE_COUNT];
double orc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
void orc_stage_reset(void) {
    for (int i = 0; i < ORC_STAGE_COUNT; i++) orc_stage_seconds[i] = 0.0;
}
void orc_stage_times(double *out) {
    for (int i = 0; i < ORC_STAGE_COUNT; i++) out[i] = orc_stage_seconds[i];
}
#define TIMED(stage, expr)                                     \
    do {                                                       \
        double t0_ = orc_now();                                \
        expr;                                                  \
        orc_stage_seconds[stage] += orc_now() - t0_;           \
    } while (0)

/* ---------------------------------------------------------------- helpers */

static inline float mixf(float a, float b, float t) { return a + t * (b - a); }
static inline float minf_glsl(float x, float y) { return y < x ? y : x; }
static inline float maxf_glsl(float x, float y) { return x < y ? y : x; }
static inline float clampf_glsl(float x, float lo, float hi) {
    return minf_glsl(maxf_glsl(x, lo), hi);
}
static inline uint32_t f2u_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}
static inline int32_t f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x < -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}
static inline uint32_t unorm8(float x) {
    float c = clampf_glsl(x, 0.0f, 1.0f);
    return (uint32_t)floorf(c * 255.0f + 0.5f);
}

uint16_t orc_f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu, mant = x & 0x7fffffu;
    if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
    int e = (int)exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        uint32_t m = mant | 0x800000u;
        int shift = 14 - e;
        uint32_t h = m >> shift, rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = sign | ((uint32_t)e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)h;
}

float orc_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = ((uint32_t)h >> 10) & 0x1fu, mant = (uint32_t)h & 0x3ffu, x;
    if (exp == 0) {
        if (mant == 0) {
            x = sign;
        } else {
            int e = -1;
            do { mant <<= 1; e++; } while (!(mant & 0x400u));
            x = sign | ((uint32_t)(127 - 15 - e) << 23) | ((mant & 0x3ffu) << 13);
        }
    } else if (exp == 0x1fu) {
        x = sign | 0x7f800000u | (mant << 13);
    } else {
        x = sign | ((exp + 112u) << 23) | (mant << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

static orc_batch *find_batch(orc_ctx *ctx, uint32_t batch_id) {
    for (int i = 0; i < MAX_BATCHES; i++)
        if (ctx->batches[i].used && ctx->batches[i].batch_id == batch_id) return &ctx->batches[i];
    return NULL;
}

static void free_batch(orc_batch *b) {
    free(b->dice_md); free(b->tpi); free(b->prop_md); free(b->backdrops);
    free(b->microlines); free(b->fills); free(b->tiles); free(b->zbuffer); free(b->ftm);
    free(b->alpha);
    memset(b, 0, sizeof(*b));
}

/* ---------------------------------------------------------------- context */

orc_ctx *orc_create(const uint8_t *lut) {
    orc_ctx *ctx = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    if (!ctx) return NULL;
    for (int i = 0; i < 256 * 256 * 4; i++) ctx->lut[i] = (float)lut[i] / 255.0f;
    return ctx;
}

void orc_destroy(orc_ctx *ctx) {
    if (!ctx) return;
    orc_end_frame(ctx);
    free(ctx->metadata);
    for (uint32_t i = 0; i < ctx->page_count; i++) free(ctx->pages[i].rgba);
    free(ctx->pages);
    free(ctx->paints); free(ctx->draw_points); free(ctx->draw_indices);
    free(ctx->clip_points); free(ctx->clip_indices); free(ctx->mask); free(ctx->fb);
    free(ctx);
}

void orc_set_option(orc_ctx *ctx, int option, int value) {
    if (option == ORC_OPT_CLIP_STRIDE_LITERAL) ctx->opt_clip_stride_literal = value;
    if (option == ORC_OPT_PLANT_DICE) ctx->opt_plant_dice = value;
    if (option == ORC_OPT_PLANT_PAINT) g_plant_paint = value;
}

void orc_begin_frame(orc_ctx *ctx, int32_t width, int32_t height, int32_t has_background,
                     const float bg[4]) {
    int32_t fbw = (width + TILE_SIZE - 1) / TILE_SIZE, fbh = (height + TILE_SIZE - 1) / TILE_SIZE;
    if (fbw != ctx->fbw || fbh != ctx->fbh || !ctx->fb) {
        free(ctx->fb);
        ctx->fb = (uint32_t *)calloc((size_t)fbw * fbh * 256, 4);
    }
    ctx->width = width; ctx->height = height; ctx->fbw = fbw; ctx->fbh = fbh;
    ctx->has_background = has_background;
    for (int i = 0; i < 4; i++) ctx->background[i] = bg ? bg[i] : 0.0f;
    ctx->dest_dirty = 0;
    ctx->alpha_tile_count = 0;
}

void orc_upload_paints(orc_ctx *ctx, const uint8_t *rgba8, uint32_t n) {
    free(ctx->paints);
    ctx->paints = (uint16_t *)malloc((size_t)(n ? n : 1) * 4 * sizeof(uint16_t));
    ctx->paint_count = n;
    const float inv255 = 1.0f / 255.0f; /* ColorU::to_f32 multiplies by splat(1/255) */
    for (uint32_t i = 0; i < 4 * n; i++) ctx->paints[i] = orc_f32_to_f16((float)rgba8[i] * inv255);
    /* a solid paint's TextureMetadataEntry (paint.rs:641-659): identity colour
     * transform, no combine, no filter, SrcOver */
    orc_metadata_entry *e = (orc_metadata_entry *)calloc(n ? n : 1, sizeof(orc_metadata_entry));
    for (uint32_t i = 0; i < n; i++) {
        e[i].transform[0] = 1.0f;
        e[i].transform[3] = 1.0f;
        memcpy(e[i].base_color, rgba8 + 4 * (size_t)i, 4);
    }
    orc_upload_metadata(ctx, e, n);
    free(e);
}

/* ---------------------------------------------------------------- paints */

/* BlendMode::to_composite_ctrl (gpu/blend.rs:172-202), PFH_BLEND_MODE_* numbering. */
static int composite_ctrl(int32_t blend_mode) {
    static const int table[27] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, /* SrcOver .. Lighter */
                                  4, 5, 1, 2, 8, 3, 6, 7, 9, 0xa, 0xb, 0xc, 0xd, 0xe, 0xf};
    return blend_mode >= 0 && blend_mode < 27 ? table[blend_mode] : 0;
}

int orc_upload_metadata(orc_ctx *ctx, const orc_metadata_entry *entries, uint32_t n) {
    float *m = (float *)calloc((size_t)(n ? n : 1) * 40, sizeof(float));
    const float inv255 = 1.0f / 255.0f;
    for (uint32_t i = 0; i < n; i++) {
        const orc_metadata_entry *e = &entries[i];
        float t[40];
        memset(t, 0, sizeof(t));
        /* texels 0-1: the colour transform (m11 m21 m12 m22 | m13 m23) */
        t[0] = e->transform[0]; t[1] = e->transform[2]; t[2] = e->transform[1];
        if (g_plant_paint == 1) { t[1] = e->transform[1]; t[2] = e->transform[2]; }
        t[3] = e->transform[3]; t[4] = e->transform[4]; t[5] = e->transform[5];
        for (int c = 0; c < 4; c++) t[8 + c] = (float)e->base_color[c] * inv255;
        /* compute_filter_params (gpu/renderer.rs:967-1049) */
        int ctrl = (composite_ctrl(e->blend_mode) << 10) | ((e->combine & 3) << 8);
        const float *p = e->filter_params;
        switch (e->filter_kind) {
        case 0: break;
        case 1: /* RadialGradient: p0 = line.from, line.vector; p1 = radii, uv_origin */
            t[12] = p[0]; t[13] = p[1]; t[14] = p[2] - p[0]; t[15] = p[3] - p[1];
            t[16] = p[4]; t[17] = p[5]; t[18] = p[6]; t[19] = p[7];
            ctrl |= 1 << 4;
            break;
        case 2: { /* Blur */
            const float sqrt_2_pi_inv = 0.3989422804014327f;
            float sigma_inv = 1.0f / p[1];
            float gx = sqrt_2_pi_inv * sigma_inv;
            float gy = expf(-0.5f * sigma_inv * sigma_inv);
            float gz = gy * gy;
            int y_dir = p[0] != 0.0f;
            t[12] = y_dir ? 0.0f : 1.0f; t[13] = y_dir ? 1.0f : 0.0f;
            t[14] = ceilf(1.5f * p[1]) * 2.0f;
            t[16] = gx; t[17] = gy; t[18] = gz;
            ctrl |= 3 << 4;
            break;
        }
        case 3: /* ColorMatrix: p0..p4 = the five columns */
            for (int k = 0; k < 20; k++) t[12 + k] = p[k];
            ctrl |= 4 << 4;
            break;
        default:
            free(m);
            return -1;
        }
        t[32] = (float)ctrl;
        for (int k = 0; k < 40; k++) m[(size_t)i * 40 + k] = orc_f16_to_f32(orc_f32_to_f16(t[k]));
    }
    free(ctx->metadata);
    ctx->metadata = m;
    ctx->metadata_count = n;
    return 0;
}

void orc_allocate_page(orc_ctx *ctx, uint32_t page, int32_t width, int32_t height) {
    if (page >= ctx->page_count) {
        uint32_t n = page + 1;
        ctx->pages = (orc_page *)realloc(ctx->pages, n * sizeof(orc_page));
        memset(ctx->pages + ctx->page_count, 0, (n - ctx->page_count) * sizeof(orc_page));
        ctx->page_count = n;
    }
    orc_page *pg = &ctx->pages[page];
    free(pg->rgba);
    pg->width = width > 0 ? width : 0;
    pg->height = height > 0 ? height : 0;
    pg->rgba = (uint8_t *)calloc((size_t)(pg->width ? pg->width : 1) * (pg->height ? pg->height : 1), 4);
}

int orc_upload_texels(orc_ctx *ctx, uint32_t page, int32_t x, int32_t y, int32_t w, int32_t h,
                      const uint8_t *rgba8) {
    if (page >= ctx->page_count || !ctx->pages[page].rgba) return -1;
    orc_page *pg = &ctx->pages[page];
    if (x < 0 || y < 0 || w < 0 || h < 0 || x + w > pg->width || y + h > pg->height) return -1;
    for (int32_t r = 0; r < h; r++)
        memcpy(pg->rgba + ((size_t)(y + r) * pg->width + x) * 4, rgba8 + (size_t)r * w * 4,
               (size_t)w * 4);
    return 0;
}

int orc_set_batch_color_texture(orc_ctx *ctx, uint32_t batch_id, int32_t page, uint32_t flags) {
    orc_batch *b = find_batch(ctx, batch_id);
    if (!b) return -1;
    b->color_page = page;
    b->color_flags = flags;
    return 0;
}

typedef struct {
    const orc_page *page;  /* NULL: no texture (samples are 0) */
    uint32_t flags;
} orc_tex;

static inline float lerp_fma(float w, float a, float b) { return fmaf(w, b - a, a); }
static inline float unorm8_load(uint8_t c) { return (float)c / 255.0f; }

/* floor(x) as an index, saturated to +-1e9 first (NaN to -1e9) so that a
 * wild coordinate converts the same way on every target */
static inline int32_t floor_index(float x) {
    return (int32_t)fminf(fmaxf(floorf(x), -1e9f), 1e9f);
}

static inline int32_t wrap_index(int32_t i, int32_t n, int repeat) {
    if (repeat) {
        int32_t m = i % n;
        return m < 0 ? m + n : m;
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

/* texture(colorTexture, uv) on an RGBA8 page: GL_LINEAR (or GL_NEAREST with
 * the nearest flags), REPEAT or CLAMP_TO_EDGE per axis; the bilinear form is
 * the one the area LUT samples with (lerp(w, a, b) = fma(w, b - a, a)). */
static void tex_sample(const orc_tex *t, float u, float v, float out[4]) {
    for (int c = 0; c < 4; c++) out[c] = 0.0f;
    if (!t->page || !t->page->width || !t->page->height) return;
    const orc_page *pg = t->page;
    int ru = (t->flags & 1) != 0, rv = (t->flags & 2) != 0;
    float W = (float)pg->width, H = (float)pg->height;
    if (t->flags & 8) {
        int32_t i = wrap_index(floor_index(u * W), pg->width, ru);
        int32_t j = wrap_index(floor_index(v * H), pg->height, rv);
        const uint8_t *p = pg->rgba + ((size_t)j * pg->width + i) * 4;
        for (int c = 0; c < 4; c++) out[c] = unorm8_load(p[c]);
        return;
    }
    float sx = u * W - 0.5f, sy = v * H - 0.5f;
    float fx = floorf(sx), fy = floorf(sy);
    float a = sx - fx, b = sy - fy;
    if (g_plant_paint == 2) { float w = a; a = b; b = w; }
    int32_t ix = floor_index(sx), iy = floor_index(sy);
    int32_t i0 = wrap_index(ix, pg->width, ru), i1 = wrap_index(ix + 1, pg->width, ru);
    int32_t j0 = wrap_index(iy, pg->height, rv), j1 = wrap_index(iy + 1, pg->height, rv);
    const uint8_t *p00 = pg->rgba + ((size_t)j0 * pg->width + i0) * 4;
    const uint8_t *p10 = pg->rgba + ((size_t)j0 * pg->width + i1) * 4;
    const uint8_t *p01 = pg->rgba + ((size_t)j1 * pg->width + i0) * 4;
    const uint8_t *p11 = pg->rgba + ((size_t)j1 * pg->width + i1) * 4;
    for (int c = 0; c < 4; c++) {
        float A = lerp_fma(a, unorm8_load(p00[c]), unorm8_load(p10[c]));
        float B = lerp_fma(a, unorm8_load(p01[c]), unorm8_load(p11[c]));
        out[c] = lerp_fma(b, A, B);
    }
}

#define FRAC_6_PI_F 1.9098593171027443f
#define FRAC_PI_3_F 1.0471975511965976f

static inline float comp_div(float n, float d) { return d != 0.0f ? n / d : 0.0f; }
static inline float glsl_mod(float x, float y) { return x - y * floorf(x / y); }

static void rgb_to_hsl(const float rgb[3], float hsl[3]) {
    float v = maxf_glsl(maxf_glsl(rgb[0], rgb[1]), rgb[2]);
    float xmin = minf_glsl(minf_glsl(rgb[0], rgb[1]), rgb[2]);
    float c = v - xmin, l = mixf(xmin, v, 0.5f);
    float tx, ty, tz;
    if (rgb[0] == v) { tx = 0.0f; ty = rgb[1]; tz = rgb[2]; }
    else if (rgb[1] == v) { tx = 2.0f; ty = rgb[2]; tz = rgb[0]; }
    else { tx = 4.0f; ty = rgb[0]; tz = rgb[1]; }
    hsl[0] = FRAC_PI_3_F * comp_div((tx * c + ty) - tz, c);
    hsl[1] = comp_div(c, v);
    hsl[2] = l;
}

static void hsl_to_rgb(const float hsl[3], float rgb[3]) {
    float a = hsl[1] * minf_glsl(hsl[2], 1.0f - hsl[2]);
    static const float k0[3] = {0.0f, 8.0f, 4.0f};
    for (int i = 0; i < 3; i++) {
        float ks = glsl_mod(k0[i] + hsl[0] * FRAC_6_PI_F, 12.0f);
        float m = minf_glsl(ks - 3.0f, 9.0f - ks);
        rgb[i] = hsl[2] - clampf_glsl(m, -1.0f, 1.0f) * a;
    }
}

static inline float screen1(float d, float s) { return (d + s) - d * s; }
static inline float hard_light1(float d, float s) {
    return s <= 0.5f ? (d * 2.0f) * s : screen1(d, 2.0f * s - 1.0f);
}
static inline float dodge1(float d, float s) {
    return d == 0.0f ? 0.0f : (s == 1.0f ? 1.0f : d / (1.0f - s));
}

/* compositeRGB (tile_fragment.inc.glsl:465-501). */
static void composite_rgb(const float d[3], const float s[3], int op, float out[3]) {
    if (op >= 0xc) {
        float dh[3], sh[3], mix3[3];
        rgb_to_hsl(d, dh);
        rgb_to_hsl(s, sh);
        switch (op) {
        case 0xc: mix3[0] = g_plant_paint == 5 ? dh[0] : sh[0]; mix3[1] = dh[1]; mix3[2] = dh[2]; break;
        case 0xd: mix3[0] = dh[0]; mix3[1] = sh[1]; mix3[2] = dh[2]; break;
        case 0xe: mix3[0] = sh[0]; mix3[1] = sh[1]; mix3[2] = dh[2]; break;
        default:  mix3[0] = dh[0]; mix3[1] = dh[1]; mix3[2] = sh[2]; break;
        }
        hsl_to_rgb(mix3, out);
        return;
    }
    for (int i = 0; i < 3; i++) {
        float x = d[i], y = s[i], r;
        switch (op) {
        case 0x1: r = x * y; break;
        case 0x2: r = screen1(x, y); break;
        case 0x3: r = g_plant_paint == 3 ? hard_light1(x, y) : hard_light1(y, x); break;
        case 0x4: r = minf_glsl(x, y); break;
        case 0x5: r = maxf_glsl(x, y); break;
        case 0x6: r = dodge1(x, y); break;
        case 0x7: r = 1.0f - dodge1(1.0f - x, 1.0f - y); break;
        case 0x8: r = hard_light1(x, y); break;
        case 0x9: {
            float dark = x <= 0.25f ? ((16.0f * x - 12.0f) * x + 4.0f) * x : sqrtf(x);
            float factor = y <= 0.5f ? x * (1.0f - x) : dark - x;
            r = x + (y * 2.0f - 1.0f) * factor;
            break;
        }
        case 0xa: r = fabsf(x - y); break;
        case 0xb: r = (x + y) - (2.0f * x) * y; break;
        default: r = y; break;
        }
        out[i] = r;
    }
}

/* calculateColor (tile_fragment.inc.glsl:567-614) for one pixel and layer,
 * after the mask: the premultiplied source colour. m is the entry's 40
 * metadata values; (fx, fy) the pixel centre. */
static void calculate_color(const orc_ctx *ctx, const orc_tex *tex, const float *m, float fx,
                            float fy, float mask_alpha, float out[4]) {
    float color[4] = {m[8], m[9], m[10], m[11]};
    int ctrl = (int)m[32];
    int combine = (ctrl >> 8) & 0x3;
    if (combine != 0) {
        /* computeTileVaryings: mat2(texel 0) * fragCoord + texel 1 */
        float u = (m[0] * fx + m[2] * fy) + m[4];
        float v = (m[1] * fx + m[3] * fy) + m[5];
        int filter = (ctrl >> 4) & 0xf;
        float c0[4];
        if (filter == 1) {
            /* filterRadialGradient (:280-306) */
            float dpx = u - m[12], dpy = v - m[13], dcx = m[14], dcy = m[15];
            float dr = m[17] - m[16];
            float a = (dcx * dcx + dcy * dcy) - dr * dr;
            float b = (dpx * dcx + dpy * dcy) + m[16] * dr;
            float c = (dpx * dpx + dpy * dpy) - m[16] * m[16];
            float discrim = b * b - a * c;
            for (int k = 0; k < 4; k++) c0[k] = 0.0f;
            if (discrim != 0.0f) {
                float sq = sqrtf(discrim);
                float t0 = (sq + b) / a, t1 = (-sq + b) / a;
                if (t0 > t1) { float w = t0; t0 = t1; t1 = w; }
                float t = t0 >= 0.0f && g_plant_paint != 6 ? t0 : t1;
                tex_sample(tex, m[18] + t, m[19] + 0.0f, c0);
            }
        } else if (filter == 3) {
            /* filterBlur (:313-351) */
            float W = tex->page ? (float)tex->page->width : 0.0f;
            float H = tex->page ? (float)tex->page->height : 0.0f;
            float ox = m[12] / W, oy = m[13] / H;
            int support = (int)m[14];
            float gx = m[16], gy = m[17], gz = m[18];
            float sum = gx;
            tex_sample(tex, u, v, c0);
            for (int k = 0; k < 4; k++) c0[k] *= gx;
            { float nx = gx * gy, ny = gy * gz; gx = nx; gy = ny; }
            for (int i = 1; i <= support; i += 2) {
                float partial = gx;
                { float nx = gx * gy, ny = gy * gz; gx = nx; gy = ny; }
                partial += gx;
                float k = (float)i + gx / partial;
                float sx = ox * k, sy = oy * k;
                float lo[4], hi[4];
                tex_sample(tex, u - sx, v - sy, lo);
                tex_sample(tex, u + sx, v + sy, hi);
                for (int q = 0; q < 4; q++) c0[q] += (lo[q] + hi[q]) * partial;
                sum += (g_plant_paint == 4 ? 1.0f : 2.0f) * partial;
                { float nx = gx * gy, ny = gy * gz; gx = nx; gy = ny; }
            }
            for (int k = 0; k < 4; k++) c0[k] /= sum;
        } else if (filter == 4) {
            /* filterColorMatrix (:353-363): mat4(p0..p3) * src + p4 */
            float src[4];
            tex_sample(tex, u, v, src);
            for (int k = 0; k < 4; k++)
                c0[k] = g_plant_paint == 7
                            ? (((m[12 + 4 * k] * src[0] + m[13 + 4 * k] * src[1]) +
                                m[14 + 4 * k] * src[2]) + m[15 + 4 * k] * src[3]) + m[28 + k]
                            : (((m[12 + k] * src[0] + m[16 + k] * src[1]) + m[20 + k] * src[2]) +
                               m[24 + k] * src[3]) + m[28 + k];
        } else {
            tex_sample(tex, u, v, c0);
        }
        /* combineColor0 (:81-89) */
        if (combine == 1 || (combine == 2 && g_plant_paint == 8)) {
            float a = c0[3] * color[3];
            color[0] = c0[0]; color[1] = c0[1]; color[2] = c0[2]; color[3] = a;
        } else if (combine == 2) {
            color[3] = c0[3] * color[3];
        }
    }
    color[3] *= mask_alpha;
    int op = (ctrl >> 10) & 0xf;
    if (op != 0) {
        /* composite (:508-523): the destination is sampled from colour
         * texture 0 at fragCoord / framebufferSize, the texture the D3D11
         * tile shader binds there (d3d11/tile.cs.glsl:135-152) */
        float dest[4], blended[3];
        tex_sample(tex, fx / (float)ctx->width, fy / (float)ctx->height, dest);
        composite_rgb(dest, color, op, blended);
        float sa = color[3];
        for (int k = 0; k < 3; k++)
            color[k] = ((sa * (1.0f - dest[3])) * color[k] + (sa * dest[3]) * blended[k]) +
                       (1.0f - sa) * dest[k];
        color[3] = 1.0f;
    }
    for (int k = 0; k < 3; k++) color[k] *= color[3];
    for (int k = 0; k < 4; k++) out[k] = color[k];
}

static void *dup(const void *src, size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (p && bytes) memcpy(p, src, bytes);
    return p;
}

void orc_upload_scene(orc_ctx *ctx, const float *dp, uint32_t ndp, const uint32_t *di, uint32_t nds,
                      const float *cp, uint32_t ncp, const uint32_t *ci, uint32_t ncs) {
    free(ctx->draw_points); free(ctx->draw_indices); free(ctx->clip_points); free(ctx->clip_indices);
    ctx->draw_points = (float *)dup(dp, (size_t)ndp * 8); ctx->n_draw_points = ndp;
    ctx->draw_indices = (uint32_t *)dup(di, (size_t)nds * 8); ctx->n_draw_segments = nds;
    ctx->clip_points = (float *)dup(cp, (size_t)ncp * 8); ctx->n_clip_points = ncp;
    ctx->clip_indices = (uint32_t *)dup(ci, (size_t)ncs * 8); ctx->n_clip_segments = ncs;
}

void orc_end_frame(orc_ctx *ctx) {
    for (int i = 0; i < MAX_BATCHES; i++)
        if (ctx->batches[i].used) free_batch(&ctx->batches[i]);
}

/* ---------------------------------------------------------------- dice */

/* dice.cs.glsl:76-90 emitMicroline. */
static void emit_microline(orc_batch *b, float x0, float y0, float x1, float y1, uint32_t path,
                           uint32_t index) {
    if (index >= b->cap_microlines) {
        uint32_t cap = b->cap_microlines ? b->cap_microlines : 1024;
        while (cap <= index) cap *= 2;
        b->microlines = (uint32_t *)realloc(b->microlines, (size_t)cap * 16);
        b->cap_microlines = cap;
    }
    float v[4] = {x0, y0, x1, y1};
    int32_t sub[4], px[4], fr[4];
    for (int i = 0; i < 4; i++) {
        sub[i] = (int32_t)rintf(clampf_glsl(v[i], -32768.0f, 32767.0f) * 256.0f);
        px[i] = (int32_t)floorf((float)sub[i] / 256.0f);
        fr[i] = sub[i] - px[i] * 256;
    }
    uint32_t *o = b->microlines + (size_t)index * 4;
    o[0] = ((uint32_t)px[0] & 0xffffu) | ((uint32_t)px[1] << 16);
    o[1] = ((uint32_t)px[2] & 0xffffu) | ((uint32_t)px[3] << 16);
    o[2] = (uint32_t)fr[0] | ((uint32_t)fr[1] << 8) | ((uint32_t)fr[2] << 16) | ((uint32_t)fr[3] << 24);
    o[3] = path;
}

/* dice.cs.glsl:126-128 getPoint: uTransform * p + uTranslation (mat2 column-major). */
static inline v2 get_point(const float *pts, uint32_t i, const float *T) {
    float px = pts[2 * (size_t)i], py = pts[2 * (size_t)i + 1];
    v2 r;
    r.x = (T[0] * px + T[1] * py) + T[4];
    r.y = (T[2] * px + T[3] * py) + T[5];
    return r;
}

/* dice.cs.glsl:117-124 sampleCurve. */
static inline v2 sample_curve(v2 p0, v2 p1, v2 p2, v2 p3, float t, int plant_swap) {
    v2 a = {mixf(p0.x, p1.x, t), mixf(p0.y, p1.y, t)};
    v2 b = {mixf(p1.x, p2.x, t), mixf(p1.y, p2.y, t)};
    v2 c = {mixf(p2.x, p3.x, t), mixf(p2.y, p3.y, t)};
    v2 ab = {mixf(a.x, b.x, t), mixf(a.y, b.y, t)};
    v2 bc = {mixf(b.x, c.x, t), mixf(b.y, c.y, t)};
    v2 r = {mixf(ab.x, bc.x, t), mixf(ab.y, bc.y, t)};
    if (plant_swap) {
        r.x = mixf(bc.x, ab.x, t);
        r.y = mixf(bc.y, ab.y, t);
    }
    return r;
}

/* dice.cs.glsl:130-217, invocations in batch-segment order. */
static int stage_dice(orc_ctx *ctx, orc_batch *b) {
    const float *pts = b->path_source ? ctx->clip_points : ctx->draw_points;
    const uint32_t *idx = b->path_source ? ctx->clip_indices : ctx->draw_indices;
    uint32_t n_idx = b->path_source ? ctx->n_clip_segments : ctx->n_draw_segments;
    uint32_t n_pts = b->path_source ? ctx->n_clip_points : ctx->n_draw_points;
    const float *T = b->xform;
    uint32_t counter = 0;
    for (uint32_t seg = 0; seg < b->segment_count; seg++) {
        uint32_t lo = 0, hi = b->path_count;
        int it = 0;
        while (it < 1024 && lo + 1 < hi) {
            uint32_t mid = lo + (hi - lo) / 2;
            uint32_t mid_seg = b->dice_md[mid * 4 + 2];
            if (seg < mid_seg) {
                hi = mid;
            } else {
                lo = mid;
                if (seg == mid_seg) break;
            }
            it++;
        }
        uint32_t path = lo;
        uint32_t gseg = seg - b->dice_md[path * 4 + 2] + b->dice_md[path * 4 + 1];
        if (gseg >= n_idx) return -1;
        uint32_t from = idx[2 * (size_t)gseg], flags = idx[2 * (size_t)gseg + 1];
        uint32_t to = from;
        if (flags & FLAG_CUBIC) to += 3;
        else if (flags & FLAG_QUAD) to += 2;
        else to += 1;
        if (to >= n_pts) return -1;
        v2 p0 = get_point(pts, from, T), p3 = get_point(pts, to, T);
        v2 c0 = {0, 0}, c1 = {0, 0};
        float countf;
        int curve = (flags & (FLAG_CUBIC | FLAG_QUAD)) != 0;
        if (curve) {
            v2 k = get_point(pts, from + 1, T);
            if (flags & FLAG_QUAD) {
                const float third = 1.0f / 3.0f;
                float kx2 = k.x * 2.0f, ky2 = k.y * 2.0f;
                c0.x = (p0.x + kx2) * third; c0.y = (p0.y + ky2) * third;
                c1.x = (p3.x + kx2) * third; c1.y = (p3.y + ky2) * third;
            } else {
                c0 = k;
                c1 = get_point(pts, from + 2, T);
            }
            /* bound = 6 * max(|ctrl.zw - 2 ctrl.xy + baseline.xy|, |baseline.zw - 2 ctrl.zw + ctrl.xy|) */
            float ax = fabsf((c1.x - 2.0f * c0.x) + p0.x), ay = fabsf((c1.y - 2.0f * c0.y) + p0.y);
            float bx = fabsf((p3.x - 2.0f * c1.x) + c0.x), by = fabsf((p3.y - 2.0f * c1.y) + c0.y);
            float mx = 6.0f * maxf_glsl(ax, bx), my = 6.0f * maxf_glsl(ay, by);
            float len = sqrtf(mx * mx + my * my);
            countf = sqrtf(len / (8.0f * TOLERANCE));
        } else {
            float dx = p3.x - p0.x, dy = p3.y - p0.y;
            countf = sqrtf(dx * dx + dy * dy) / MICROLINE_LENGTH;
        }
        int32_t n = f2i_sat(ceilf(countf));
        if (n < 1) n = 1;
        if (ctx->opt_plant_dice == 1) n *= 2;
        if (ctx->opt_plant_dice == 2) n = (n + 1) / 2;
        if (n > (1 << 24)) return -2;
        uint32_t first = counter;
        counter += (uint32_t)n;
        v2 prev = p0;
        for (int32_t i = 0; i < n; i++) {
            float t = (float)(i + 1) / (float)n;
            v2 next;
            if (curve) {
                next = sample_curve(p0, c0, c1, p3, t, ctx->opt_plant_dice == 3);
            } else {
                next.x = mixf(p0.x, p3.x, t);
                next.y = mixf(p0.y, p3.y, t);
            }
            emit_microline(b, prev.x, prev.y, next.x, next.y, path, first + (uint32_t)i);
            prev = next;
        }
    }
    b->n_microlines = counter;
    return 0;
}

/* ---------------------------------------------------------------- bound */

/* bound.cs.glsl:49-84. Independent per tile. */
static void stage_bound(orc_batch *b) {
    int64_t T = b->tile_count;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t t = 0; t < T; t++) {
        uint32_t lo = 0, hi = b->path_count;
        int it = 0;
        while (it < 1024 && lo + 1 < hi) {
            uint32_t mid = lo + (hi - lo) / 2;
            uint32_t mid_tile = b->tpi[mid * 4 + 2];
            if ((uint32_t)t < mid_tile) {
                hi = mid;
            } else {
                lo = mid;
                if ((uint32_t)t == mid_tile) break;
            }
            it++;
        }
        uint32_t *tile = b->tiles + (size_t)t * 4;
        tile[0] = 0xffffffffu;
        tile[1] = 0xffffffffu;
        tile[2] = 0x00ffffffu;
        tile[3] = b->tpi[lo * 4 + 3];
    }
}

/* ---------------------------------------------------------------- bin */

typedef struct { int32_t x0, y0, x1, y1; } irect;

static inline int outcode_any(int32_t tx, int32_t ty, irect r) {
    return tx < r.x0 || ty < r.y0 || tx >= r.x1 || ty >= r.y1;
}

/* bin.cs.glsl:98-121 addFill. */
static void add_fill(orc_batch *b, float x0, float y0, float x1, float y1, int32_t tx, int32_t ty,
                     irect r, uint32_t off) {
    if (outcode_any(tx, ty, r)) return;
    uint32_t t = off + (uint32_t)(tx - r.x0) + (uint32_t)(ty - r.y0) * (uint32_t)(r.x1 - r.x0);
    float ox = (float)(tx * 16), oy = (float)(ty * 16);
    uint32_t sx0 = f2u_sat((x0 - ox) * 256.0f), sy0 = f2u_sat((y0 - oy) * 256.0f);
    uint32_t sx1 = f2u_sat((x1 - ox) * 256.0f), sy1 = f2u_sat((y1 - oy) * 256.0f);
    if (sx0 == sx1) return;
    uint32_t f = b->n_fills++;
    uint32_t link = b->tiles[(size_t)t * 4 + 1];
    b->tiles[(size_t)t * 4 + 1] = f;
    if (f >= b->cap_fills) {
        uint32_t cap = b->cap_fills ? b->cap_fills * 2 : 4096;
        b->fills = (uint32_t *)realloc(b->fills, (size_t)cap * 12);
        b->cap_fills = cap;
    }
    b->fills[(size_t)f * 3 + 0] = sx0 | (sy0 << 16);
    b->fills[(size_t)f * 3 + 1] = sx1 | (sy1 << 16);
    b->fills[(size_t)f * 3 + 2] = link;
}

/* bin.cs.glsl:123-143 adjustBackdrop. */
static void adjust_backdrop(orc_batch *b, int32_t delta, int32_t tx, int32_t ty, irect r,
                            uint32_t off, uint32_t boff) {
    int ox = tx < r.x0, oy = ty < r.y0, oz = tx >= r.x1, ow = ty >= r.y1;
    if (ox || oy || oz || ow) {
        if (!ox && oy && !oz) {
            uint32_t bi = boff + (uint32_t)(tx - r.x0);
            b->backdrops[(size_t)bi * 3] += delta;
        }
    } else {
        uint32_t t = off + (uint32_t)(tx - r.x0) + (uint32_t)(ty - r.y0) * (uint32_t)(r.x1 - r.x0);
        b->tiles[(size_t)t * 4 + 2] += (uint32_t)delta << 24;
    }
}

/* bin.cs.glsl:154-256, invocations in microline order. */
static void stage_bin(orc_batch *b) {
    b->n_fills = 0;
    for (uint32_t m = 0; m < b->n_microlines; m++) {
        const uint32_t *ml = b->microlines + (size_t)m * 4;
        uint32_t path = ml[3];
        int32_t w0 = (int32_t)ml[0], w1 = (int32_t)ml[1], w2 = (int32_t)ml[2];
        float lx0 = (float)((int32_t)((uint32_t)w0 << 16) >> 16) + (float)(w2 & 0xff) / 256.0f;
        float ly0 = (float)(w0 >> 16) + (float)((w2 >> 8) & 0xff) / 256.0f;
        float lx1 = (float)((int32_t)((uint32_t)w1 << 16) >> 16) + (float)((w2 >> 16) & 0xff) / 256.0f;
        float ly1 = (float)(w1 >> 16) + (float)((w2 >> 24) & 0xff) / 256.0f;
        const int32_t *md = b->prop_md + (size_t)path * 12;
        irect r = {md[0], md[1], md[2], md[3]};
        uint32_t off = (uint32_t)md[4], boff = (uint32_t)md[8];

        int32_t ftx = (int32_t)floorf(lx0 / 16.0f), fty = (int32_t)floorf(ly0 / 16.0f);
        int32_t ttx = (int32_t)floorf(lx1 / 16.0f), tty = (int32_t)floorf(ly1 / 16.0f);
        float vx = lx1 - lx0, vy = ly1 - ly0;
        int32_t sx = vx < 0.0f ? -1 : 1, sy = vy < 0.0f ? -1 : 1;
        float fcx = (float)((ftx + (vx >= 0.0f ? 1 : 0)) * 16);
        float fcy = (float)((fty + (vy >= 0.0f ? 1 : 0)) * 16);
        float tmx = (fcx - lx0) / vx, tmy = (fcy - ly0) / vy;
        float tdx = fabsf(16.0f / vx), tdy = fabsf(16.0f / vy);
        float cx = lx0, cy = ly0;
        int32_t tx = ftx, ty = fty;
        int last = 0; /* 0 none, 1 x, 2 y */
        for (uint32_t it = 0; it < MAX_ITERATIONS; it++) {
            int next;
            if (tmx < tmy) next = 1;
            else if (tmx > tmy) next = 2;
            else if (sx > 0) next = 1;
            else next = 2;
            float nt = minf_glsl(next == 1 ? tmx : tmy, 1.0f);
            if (tx == ttx && ty == tty) next = 0;
            float nx = mixf(lx0, lx1, nt), ny = mixf(ly0, ly1, nt);
            add_fill(b, cx, cy, nx, ny, tx, ty, r, off);
            if (sy < 0 && next == 2) {
                add_fill(b, nx, ny, (float)(tx * 16), (float)(ty * 16), tx, ty, r, off);
            } else if (sy > 0 && last == 2) {
                add_fill(b, (float)(tx * 16), (float)(ty * 16), cx, cy, tx, ty, r, off);
            }
            if (sx < 0 && last == 1) {
                adjust_backdrop(b, 1, tx, ty, r, off, boff);
            } else if (sx > 0 && next == 1) {
                adjust_backdrop(b, -1, tx, ty, r, off, boff);
            }
            if (next == 1) {
                tmx += tdx;
                tx += sx;
            } else if (next == 2) {
                tmy += tdy;
                ty += sy;
            } else {
                break;
            }
            cx = nx; cy = ny;
            last = next;
        }
    }
}

/* ---------------------------------------------------------------- propagate */

/* propagate.cs.glsl:103-218, invocations in column order. */
static int stage_propagate(orc_ctx *ctx, orc_batch *b, orc_batch *clip) {
    int32_t fbw = ctx->fbw, fbh = ctx->fbh;
    int64_t fbn = (int64_t)fbw * fbh;
    b->n_alpha = 0;
    b->alpha_start = ctx->alpha_tile_count;
    for (uint32_t col = 0; col < b->column_count; col++) {
        int32_t backdrop = b->backdrops[(size_t)col * 3 + 0];
        int32_t tile_x = b->backdrops[(size_t)col * 3 + 1];
        uint32_t dpath = (uint32_t)b->backdrops[(size_t)col * 3 + 2];
        if (dpath >= b->path_count) return -1;
        const int32_t *md = b->prop_md + (size_t)dpath * 12;
        uint32_t rx0 = (uint32_t)md[0], ry0 = (uint32_t)md[1], rx1 = (uint32_t)md[2], ry1 = (uint32_t)md[3];
        uint32_t height = ry1 - ry0, width = rx1 - rx0;
        uint32_t dofs = (uint32_t)md[4];
        int zwrite = md[6] != 0;
        int32_t cpi = md[7];
        uint32_t crx0 = 0, cry0 = 0, crx1 = 0, cry1 = 0, cofs = 0;
        if (cpi >= 0) {
            if (!clip) return -2;
            size_t stride = ctx->opt_clip_stride_literal ? 8 : 12;
            size_t base = (size_t)cpi * stride;
            if (base + 5 > (size_t)clip->path_count * 12) return -3;
            crx0 = (uint32_t)clip->prop_md[base + 0]; cry0 = (uint32_t)clip->prop_md[base + 1];
            crx1 = (uint32_t)clip->prop_md[base + 2]; cry1 = (uint32_t)clip->prop_md[base + 3];
            cofs = (uint32_t)clip->prop_md[base + 4];
        }
        for (uint32_t y = 0; y < height; y++) {
            uint32_t t = dofs + y * width + (uint32_t)tile_x;
            uint32_t *tile = b->tiles + (size_t)t * 4;
            int32_t draw_alpha = -1, clip_alpha = -1;
            int32_t first_fill = (int32_t)tile[1];
            int32_t delta = (int32_t)tile[2] >> 24;
            uint32_t word = tile[3] & 0x00ffffffu;
            int32_t tile_backdrop = backdrop;
            int have_mask = first_fill >= 0;
            int need_alpha = have_mask;
            if (cpi >= 0) {
                uint32_t ux = (uint32_t)tile_x + rx0, uy = y + ry0; /* unsigned, as in the shader */
                if (ux >= crx0 && uy >= cry0 && ux < crx1 && uy < cry1) {
                    uint32_t ct = cofs + (uy - cry0) * (crx1 - crx0) + (ux - crx0);
                    if ((size_t)ct >= clip->tile_count) return -4;
                    const uint32_t *ctile = clip->tiles + (size_t)ct * 4;
                    int32_t this_clip_alpha = (int32_t)(ctile[2] << 8) >> 8;
                    int32_t clip_backdrop = (int32_t)ctile[3] >> 24;
                    if (this_clip_alpha >= 0) {
                        if (have_mask) {
                            clip_alpha = this_clip_alpha;
                            need_alpha = 1;
                        } else if (tile_backdrop != 0) {
                            draw_alpha = this_clip_alpha;
                            clip_alpha = -1;
                            need_alpha = 0;
                        } else {
                            draw_alpha = -1;
                            clip_alpha = -1;
                            need_alpha = 0;
                        }
                    } else if (clip_backdrop == 0) {
                        tile_backdrop = 0;
                        need_alpha = 0;
                    }
                } else {
                    tile_backdrop = 0;
                    need_alpha = 0;
                }
            }
            if (need_alpha) {
                uint32_t ai = b->n_alpha++;
                b->alpha[(size_t)ai * 2 + 0] = t;
                b->alpha[(size_t)ai * 2 + 1] = (uint32_t)clip_alpha;
                draw_alpha = (int32_t)(ai + b->alpha_start);
            }
            tile[2] = ((uint32_t)draw_alpha & 0x00ffffffu) | ((uint32_t)delta << 24);
            tile[3] = word | ((uint32_t)tile_backdrop << 24);
            int32_t gx = tile_x + (int32_t)rx0, gy = (int32_t)y + (int32_t)ry0;
            if (gx >= 0 && gy >= 0 && gx < fbw && gy < fbh) {
                int64_t pos = (int64_t)gy * fbw + gx;
                if (pos < fbn) {
                    if (zwrite && tile_backdrop != 0 && draw_alpha < 0) {
                        if ((int32_t)t > b->zbuffer[8 + pos]) b->zbuffer[8 + pos] = (int32_t)t;
                    }
                    if (tile_backdrop != 0 || draw_alpha >= 0) {
                        int32_t nxt = b->ftm[pos];
                        b->ftm[pos] = (int32_t)t;
                        tile[0] = (uint32_t)nxt;
                    }
                }
            }
            backdrop += delta;
        }
    }
    b->zbuffer[4] += (int32_t)b->n_alpha;
    ctx->alpha_tile_count += b->n_alpha;
    return 0;
}

/* ---------------------------------------------------------------- sort */

/* sort.cs.glsl:60-95; fb tiles are independent. */
static void stage_sort(orc_ctx *ctx, orc_batch *b) {
    int64_t fbn = (int64_t)ctx->fbw * ctx->fbh;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 16)
    for (int64_t g = 0; g < fbn; g++) {
        int32_t z = b->zbuffer[8 + g];
        int32_t unsorted = b->ftm[g], sorted = -1;
        while (unsorted >= 0) {
            int32_t cur = unsorted;
            unsorted = (int32_t)b->tiles[(size_t)cur * 4];
            if (cur >= z) {
                int32_t prev = -1, trial = sorted;
                for (;;) {
                    if (trial < 0 || cur < trial) {
                        if (prev < 0) {
                            b->tiles[(size_t)cur * 4] = (uint32_t)sorted;
                            sorted = cur;
                        } else {
                            b->tiles[(size_t)cur * 4] = (uint32_t)trial;
                            b->tiles[(size_t)prev * 4] = (uint32_t)cur;
                        }
                        break;
                    }
                    prev = trial;
                    trial = (int32_t)b->tiles[(size_t)trial * 4];
                }
            }
        }
        b->ftm[g] = sorted;
    }
}

/* ---------------------------------------------------------------- fill */

/* texture(uAreaLUT, uv) with GL_LINEAR + CLAMP_TO_EDGE, fp32 weights; each
 * linear interpolation is llvmpipe's v0 + w * (v1 - v0) as one fused
 * multiply-add (lp_build_lerp), along s first, then along t. */
static inline void sample_lut(const orc_ctx *ctx, float u, float v, float out[4]) {
    float s = u * 256.0f - 0.5f, t = v * 256.0f - 0.5f;
    float fs = floorf(s), ft = floorf(t);
    float a = s - fs, bb = t - ft;
    int32_t i0 = f2i_sat(fs), j0 = f2i_sat(ft);
    int32_t i1 = i0 < 2147483647 ? i0 + 1 : i0, j1 = j0 < 2147483647 ? j0 + 1 : j0;
    i0 = i0 < 0 ? 0 : (i0 > 255 ? 255 : i0);
    i1 = i1 < 0 ? 0 : (i1 > 255 ? 255 : i1);
    j0 = j0 < 0 ? 0 : (j0 > 255 ? 255 : j0);
    j1 = j1 < 0 ? 0 : (j1 > 255 ? 255 : j1);
    const float *t00 = ctx->lut + ((size_t)j0 * 256 + i0) * 4;
    const float *t10 = ctx->lut + ((size_t)j0 * 256 + i1) * 4;
    const float *t01 = ctx->lut + ((size_t)j1 * 256 + i0) * 4;
    const float *t11 = ctx->lut + ((size_t)j1 * 256 + i1) * 4;
    for (int c = 0; c < 4; c++) {
        float r0 = fmaf(a, t10[c] - t00[c], t00[c]);
        float r1 = fmaf(a, t11[c] - t01[c], t01[c]);
        out[c] = fmaf(bb, r1 - r0, r0);
    }
}

/* fill_area.inc.glsl:11-27 computeCoverage. */
static inline void compute_coverage(const orc_ctx *ctx, float fx, float fy, float tx, float ty,
                                    float out[4]) {
    int lt = fx < tx;
    float lx = lt ? fx : tx, ly = lt ? fy : ty, rx = lt ? tx : fx, ry = lt ? ty : fy;
    float wx = clampf_glsl(fx, -0.5f, 0.5f), wy = clampf_glsl(tx, -0.5f, 0.5f);
    float offset = mixf(wx, wy, 0.5f) - lx;
    float t = offset / (rx - lx);
    float y = mixf(ly, ry, t);
    float d = (ry - ly) / (rx - lx);
    float dX = wx - wy;
    float s[4];
    sample_lut(ctx, (y + 8.0f) / 16.0f, fabsf(d * dX) / 16.0f, s);
    for (int c = 0; c < 4; c++) out[c] = s[c] * dX;
}

void orc_compute_coverage(const orc_ctx *ctx, float fx, float fy, float tx, float ty, float out[4]) {
    compute_coverage(ctx, fx, fy, tx, ty, out);
}

static void ensure_mask(orc_ctx *ctx, uint32_t tiles) {
    if (tiles <= ctx->mask_cap_tiles) return;
    uint32_t cap = ctx->mask_cap_tiles ? ctx->mask_cap_tiles : 65536;
    while (cap < tiles) cap *= 2;
    ctx->mask = (uint32_t *)realloc(ctx->mask, (size_t)cap * 64 * 4);
    memset(ctx->mask + (size_t)ctx->mask_cap_tiles * 64, 0, (size_t)(cap - ctx->mask_cap_tiles) * 256);
    ctx->mask_cap_tiles = cap;
}

/* fill.cs.glsl:61-88 with fill_compute.inc.glsl:11-25. Alpha tiles are independent. */
static void stage_fill(orc_ctx *ctx, orc_batch *b) {
    ensure_mask(ctx, b->alpha_start + b->n_alpha);
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 8)
    for (int64_t ba = 0; ba < (int64_t)b->n_alpha; ba++) {
        uint32_t alpha_index = (uint32_t)ba + b->alpha_start;
        uint32_t t = b->alpha[(size_t)ba * 2];
        const uint32_t *tile = b->tiles + (size_t)t * 4;
        if (((int32_t)(tile[2] << 8) >> 8) < 0) continue;
        int32_t first_fill = (int32_t)tile[1];
        int32_t backdrop = (int32_t)tile[3] >> 24;
        int32_t clip_alpha = (int32_t)b->alpha[(size_t)ba * 2 + 1];
        for (int ly = 0; ly < 4; ly++) {
            for (int lx = 0; lx < 16; lx++) {
                float fragx = (float)lx + 0.5f, fragy = (float)(ly * 4) + 0.5f;
                float acc[4] = {0, 0, 0, 0};
                int32_t f = first_fill;
                int it = 0;
                do {
                    uint32_t from = b->fills[(size_t)f * 3], to = b->fills[(size_t)f * 3 + 1];
                    float sx = (float)(from & 0xffffu) / 256.0f - fragx;
                    float sy = (float)(from >> 16) / 256.0f - fragy;
                    float ex = (float)(to & 0xffffu) / 256.0f - fragx;
                    float ey = (float)(to >> 16) / 256.0f - fragy;
                    float cov[4];
                    compute_coverage(ctx, sx, sy, ex, ey, cov);
                    for (int c = 0; c < 4; c++) acc[c] += cov[c];
                    f = (int32_t)b->fills[(size_t)f * 3 + 2];
                    it++;
                } while (f >= 0 && it < 1024);
                float c4[4];
                for (int c = 0; c < 4; c++) {
                    float v = (float)backdrop + acc[c];
                    c4[c] = clampf_glsl(fabsf(v), 0.0f, 1.0f);
                }
                if (clip_alpha >= 0) {
                    uint32_t ct = ctx->mask[(size_t)(uint32_t)clip_alpha * 64 + ly * 16 + lx];
                    for (int c = 0; c < 4; c++)
                        c4[c] = minf_glsl(c4[c], (float)((ct >> (8 * c)) & 0xffu) / 255.0f);
                }
                uint32_t out = 0;
                for (int c = 0; c < 4; c++) out |= unorm8(c4[c]) << (8 * c);
                ctx->mask[(size_t)alpha_index * 64 + ly * 16 + lx] = out;
            }
        }
    }
}

/* ---------------------------------------------------------------- tile */

/* tile.cs.glsl:71-163 with calculateColor (tile_fragment.inc.glsl:577-614),
 * sampleMask (:539-556) and the base colour fetch of computeTileVaryings
 * (tile_vertex.inc.glsl:15-46). Solid colours, Normal composite. */
static void stage_tile(orc_ctx *ctx, orc_batch *b) {
    int load = !ctx->has_background || ctx->dest_dirty;
    int32_t fbw = ctx->fbw, fbh = ctx->fbh;
    int64_t fbn = (int64_t)fbw * fbh;
    int64_t pitch = (int64_t)fbw * 16;
    orc_tex tex = {NULL, b->color_flags};
    if (b->color_page >= 0 && (uint32_t)b->color_page < ctx->page_count &&
        ctx->pages[b->color_page].rgba)
        tex.page = &ctx->pages[b->color_page];
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 8)
    for (int64_t g = 0; g < fbn; g++) {
        int32_t first = b->ftm[g];
        if (first < 0 && load) continue;
        int32_t gx = (int32_t)(g % fbw), gy = (int32_t)(g / fbw);
        for (int py = 0; py < 16; py++) {
            int ly = py >> 2, sub = py & 3;
            for (int lx = 0; lx < 16; lx++) {
                uint32_t *px = ctx->fb + (size_t)(gy * 16 + py) * pitch + gx * 16 + lx;
                float d[4];
                if (!load) {
                    for (int c = 0; c < 4; c++) d[c] = ctx->background[c];
                } else {
                    for (int c = 0; c < 4; c++) d[c] = (float)((*px >> (8 * c)) & 0xffu) / 255.0f;
                }
                int32_t t = first;
                while (t >= 0) {
                    const uint32_t *tile = b->tiles + (size_t)t * 4;
                    int32_t alpha = (int32_t)(tile[2] << 8) >> 8;
                    uint32_t word = tile[3];
                    uint32_t color = word & 0xffffu;
                    int32_t ctrl = (int32_t)((word >> 16) & 0xffu);
                    float backdrop;
                    if (alpha >= 0) {
                        backdrop = 0.0f;
                    } else {
                        backdrop = (float)((int32_t)word >> 24);
                        ctrl &= ~0x3;
                    }
                    float mask_alpha = 1.0f;
                    int mask_ctrl = ctrl & 0x3;
                    if (mask_ctrl != 0) {
                        uint32_t texel = ctx->mask[(size_t)(uint32_t)alpha * 64 + ly * 16 + lx];
                        float cov = (float)((texel >> (8 * sub)) & 0xffu) / 255.0f + backdrop;
                        if (mask_ctrl & 0x1) {
                            cov = fabsf(cov);
                        } else {
                            float m = cov - 2.0f * floorf(cov / 2.0f);
                            cov = 1.0f - fabsf(1.0f - m);
                        }
                        mask_alpha = minf_glsl(mask_alpha, cov);
                    }
                    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                    if (color < ctx->metadata_count)
                        calculate_color(ctx, &tex, ctx->metadata + (size_t)color * 40,
                                        (float)(gx * 16 + lx) + 0.5f,
                                        (float)(gy * 16 + py) + 0.5f, mask_alpha, s);
                    float ia = 1.0f - s[3];
                    for (int c = 0; c < 4; c++) d[c] = d[c] * ia + s[c];
                    t = (int32_t)tile[0];
                }
                uint32_t out = 0;
                for (int c = 0; c < 4; c++) out |= unorm8(d[c]) << (8 * c);
                *px = out;
            }
        }
    }
    ctx->dest_dirty = 1;
}

/* ---------------------------------------------------------------- driver */

int orc_prepare_batch(orc_ctx *ctx, uint32_t batch_id, uint32_t path_source, uint32_t path_count,
                      uint32_t tile_count, uint32_t segment_count, const float xform[6],
                      const uint32_t *dice_md, const uint32_t *tpi, const int32_t *prop_md,
                      const int32_t *backdrops, uint32_t column_count, int32_t clip_batch_id) {
    if (find_batch(ctx, batch_id)) return -10;
    orc_batch *b = NULL;
    for (int i = 0; i < MAX_BATCHES; i++)
        if (!ctx->batches[i].used) { b = &ctx->batches[i]; break; }
    if (!b) return -11;
    orc_batch *clip = NULL;
    if (clip_batch_id >= 0) {
        clip = find_batch(ctx, (uint32_t)clip_batch_id);
        if (!clip) return -12;
    }
    memset(b, 0, sizeof(*b));
    b->used = 1;
    b->color_page = -1;
    b->batch_id = batch_id; b->path_source = path_source; b->path_count = path_count;
    b->tile_count = tile_count; b->segment_count = segment_count; b->column_count = column_count;
    b->clip_batch_id = clip_batch_id;
    memcpy(b->xform, xform, sizeof(b->xform));
    b->dice_md = (uint32_t *)dup(dice_md, (size_t)path_count * 16);
    b->tpi = (uint32_t *)dup(tpi, (size_t)path_count * 16);
    b->prop_md = (int32_t *)dup(prop_md, (size_t)path_count * 48);
    b->backdrops = (int32_t *)dup(backdrops, (size_t)column_count * 12);
    b->tiles = (uint32_t *)malloc((size_t)(tile_count ? tile_count : 1) * 16);
    b->alpha = (uint32_t *)malloc((size_t)(tile_count ? tile_count : 1) * 8);
    int64_t fbn = (int64_t)ctx->fbw * ctx->fbh;
    b->zbuffer = (int32_t *)calloc((size_t)fbn + 8, 4);
    b->ftm = (int32_t *)malloc((size_t)(fbn ? fbn : 1) * 4);
    for (int64_t i = 0; i < fbn; i++) b->ftm[i] = -1;

    int rc;
    TIMED(ORC_STAGE_DICE, rc = stage_dice(ctx, b));
    if (rc) return rc;
    TIMED(ORC_STAGE_BOUND, stage_bound(b));
    /* bin_segments uploads [6, 0, 0, 0, 0, microline_count, 0, 0] to the header
     * (d3d11/renderer.rs:174-180); bin bumps [1]. */
    int32_t hdr[8] = {6, 0, 0, 0, 0, (int32_t)b->n_microlines, 0, 0};
    TIMED(ORC_STAGE_BIN, stage_bin(b));
    hdr[1] = (int32_t)b->n_fills;
    /* propagate_tiles uploads tile_area zero words at offset 0 (:562-566). */
    for (int i = 0; i < 8; i++) b->zbuffer[i] = (i < fbn) ? 0 : hdr[i];
    TIMED(ORC_STAGE_PROPAGATE, rc = stage_propagate(ctx, b, clip));
    if (rc) return rc;
    TIMED(ORC_STAGE_FILL, stage_fill(ctx, b));
    TIMED(ORC_STAGE_SORT, stage_sort(ctx, b));
    return 0;
}

int orc_draw_batch(orc_ctx *ctx, uint32_t batch_id) {
    orc_batch *b = find_batch(ctx, batch_id);
    if (!b) return -1;
    TIMED(ORC_STAGE_TILE, stage_tile(ctx, b));
    return 0;
}

int64_t orc_count(orc_ctx *ctx, uint32_t batch_id, int what) {
    orc_batch *b = find_batch(ctx, batch_id);
    if (!b) return -1;
    int64_t fbn = (int64_t)ctx->fbw * ctx->fbh;
    switch (what) {
    case ORC_MICROLINES: return b->n_microlines;
    case ORC_FILLS: return b->n_fills;
    case ORC_TILES: return b->tile_count;
    case ORC_ZBUFFER: return fbn + 8;
    case ORC_FIRST_TILE_MAP: return fbn;
    case ORC_ALPHA_TILES: return b->n_alpha;
    case ORC_BACKDROPS: return b->column_count;
    case ORC_ALPHA_RANGE: return 2;
    }
    return -1;
}

int64_t orc_read(orc_ctx *ctx, uint32_t batch_id, int what, void *dst, size_t cap) {
    orc_batch *b = find_batch(ctx, batch_id);
    if (!b) return -1;
    int64_t n = orc_count(ctx, batch_id, what);
    const void *src = NULL;
    size_t esz = 0;
    uint32_t range[2] = {b->alpha_start, b->alpha_start + b->n_alpha};
    switch (what) {
    case ORC_MICROLINES: src = b->microlines; esz = 16; break;
    case ORC_FILLS: src = b->fills; esz = 12; break;
    case ORC_TILES: src = b->tiles; esz = 16; break;
    case ORC_ZBUFFER: src = b->zbuffer; esz = 4; break;
    case ORC_FIRST_TILE_MAP: src = b->ftm; esz = 4; break;
    case ORC_ALPHA_TILES: src = b->alpha; esz = 8; break;
    case ORC_BACKDROPS: src = b->backdrops; esz = 12; break;
    case ORC_ALPHA_RANGE: src = range; esz = 4; break;
    default: return -1;
    }
    size_t bytes = (size_t)n * esz;
    if (bytes > cap) return -2;
    if (bytes) memcpy(dst, src, bytes);
    return (int64_t)bytes;
}

uint32_t orc_alpha_tile_count(orc_ctx *ctx) { return ctx->alpha_tile_count; }

int64_t orc_read_mask(orc_ctx *ctx, uint32_t *dst, size_t cap_words) {
    size_t n = (size_t)ctx->alpha_tile_count * 64;
    if (n > cap_words) return -2;
    if (n) memcpy(dst, ctx->mask, n * 4);
    return (int64_t)n;
}

int64_t orc_read_framebuffer(orc_ctx *ctx, uint32_t *dst, size_t cap_words) {
    size_t n = (size_t)ctx->width * ctx->height;
    if (n > cap_words) return -2;
    size_t pitch = (size_t)ctx->fbw * 16;
    for (int32_t y = 0; y < ctx->height; y++)
        memcpy(dst + (size_t)y * ctx->width, ctx->fb + (size_t)y * pitch, (size_t)ctx->width * 4);
    return (int64_t)n;
}
