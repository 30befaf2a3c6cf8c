/*
 * pf_d3d9.c - the D3D9-level pipeline of the reference restated on the CPU:
 * the tiler (renderer/src/tiler.rs) with the object packing of
 * renderer/src/builder.rs:415-616, 998-1041, and the raster of
 * renderer/src/gpu/d3d9/renderer.rs:150-472 with the shaders in shaders/d3d9.
 *
 * TEST INFRASTRUCTURE ONLY (a second opinion on whole frames, never the
 * product). The D3D9 level differs from the D3D11 compute path by design:
 * curves are split at t = 0.5 until flat (no 1/256 microline quantization),
 * lines are clipped to the view box columns, fill coordinates are rounded to
 * nearest (not truncated), masks accumulate in binary16 with additive
 * blending, the fill rule is applied at composite, and every tile is blended
 * into the RGBA8 target on its own. The two levels therefore agree only to a
 * tolerance, which the tests state.
 *
 * Choices where the reference leaves the result to the GPU or to thread
 * timing:
 *   - alpha tile ids are numbered path by path in build order (the reference
 *     takes them from an atomic counter across threads); ids do not change
 *     pixels;
 *   - a binary16 blend is the fp32 sum rounded to nearest even;
 *   - fragment varyings are evaluated exactly at the pixel centre;
 *   - the RGBA8 blend result is stored as floor(clamp(x) * 255 + 0.5);
 *   - a solid tile samples the mask outside the allocated pages and reads 0.
 * Segments are built as the D3D11 packing builds them (builder.rs:770-840:
 * each contour closed back to its first point), which is what the tiler's
 * contour iterator yields for contours that end on an on-curve point.
 */
#include "pf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define D9_TILE 16
#define D9_FLAG_QUAD 0x80000000u
#define D9_FLAG_CUBIC 0x40000000u
#define D9_TOLERANCE 0.25f
#define D9_NO_TILE (-1)
#define D9_MAX_DEPTH 24
#define D9_MAX_STEPS (1u << 20)

typedef struct { float x, y; } d9v;

typedef struct {
    uint16_t x0, y0, x1, y1;
    int32_t local; /* local alpha tile number within the path */
} d9_fill;

typedef struct {
    int32_t dest, dest_backdrop, src, src_backdrop; /* global alpha ids */
} d9_clip_job;

typedef struct {
    int32_t rx0, ry0, rx1, ry1; /* tile rect */
    int32_t w, h;
    int32_t *alpha;             /* per tile: local alpha number, then global id, or -1 */
    int8_t *backdrop;           /* per tile */
    int32_t *col_backdrop;      /* per column (tiles above the rect) */
    d9_fill *fills;
    uint32_t n_fills, cap_fills;
    int32_t n_alpha, alpha_base;
    d9_clip_job *jobs;
    uint32_t n_jobs;
    int fill_rule, clip, occludes;
    uint32_t paint;
} d9_path;

static void d9_free_path(d9_path *p) {
    free(p->alpha); free(p->backdrop); free(p->col_backdrop); free(p->fills); free(p->jobs);
    memset(p, 0, sizeof(*p));
}

static inline d9v d9_xform(const float *m, float x, float y) {
    d9v r = {(m[0] * x + m[1] * y) + m[4], (m[2] * x + m[3] * y) + m[5]};
    return r;
}

/* builder.rs:570-600 tile_coords_to_local_index / add_fill. */
static void d9_add_fill(d9_path *p, d9v from, d9v to, int32_t tx, int32_t ty) {
    if (tx < p->rx0 || ty < p->ry0 || tx >= p->rx1 || ty >= p->ry1) return;
    float ox = (float)(tx * D9_TILE), oy = (float)(ty * D9_TILE);
    float v[4] = {(from.x - ox) * 256.0f, (from.y - oy) * 256.0f, (to.x - ox) * 256.0f,
                  (to.y - oy) * 256.0f};
    int32_t q[4];
    for (int i = 0; i < 4; i++) {
        /* F32x4::clamp: max then min, as SSE orders them (NaN -> the bound) */
        float c = v[i] > 0.0f ? v[i] : 0.0f;
        c = c < (float)(D9_TILE * 256 - 1) ? c : (float)(D9_TILE * 256 - 1);
        q[i] = (int32_t)rintf(c); /* _mm_cvtps_epi32: round to nearest even */
    }
    if (q[0] == q[2]) return;
    int32_t local = (tx - p->rx0) + (ty - p->ry0) * p->w;
    if (p->alpha[local] < 0) p->alpha[local] = p->n_alpha++;
    if (p->n_fills == p->cap_fills) {
        p->cap_fills = p->cap_fills ? p->cap_fills * 2 : 256;
        p->fills = (d9_fill *)realloc(p->fills, (size_t)p->cap_fills * sizeof(d9_fill));
    }
    d9_fill f = {(uint16_t)q[0], (uint16_t)q[1], (uint16_t)q[2], (uint16_t)q[3], p->alpha[local]};
    p->fills[p->n_fills++] = f;
}

/* builder.rs:602-616 adjust_alpha_tile_backdrop. */
static void d9_adjust_backdrop(d9_path *p, int32_t tx, int32_t ty, int delta) {
    int32_t ox = tx - p->rx0, oy = ty - p->ry0;
    if (ox < 0 || ox >= p->w || oy >= p->h) return;
    if (oy < 0) {
        p->col_backdrop[ox] += delta;
        return;
    }
    int32_t i = ox + oy * p->w;
    p->backdrop[i] = (int8_t)(p->backdrop[i] + delta);
}

/* clip.rs:494-547 clip_line_segment_to_rect (Cohen-Sutherland). */
static int d9_outcode(d9v pt, const float *r) {
    int o = 0;
    if (pt.x < r[0]) o |= 1;
    if (pt.y < r[1]) o |= 4;
    if (pt.x > r[2]) o |= 2;
    if (pt.y > r[3]) o |= 8;
    return o;
}

static inline float d9_lerp(float a, float b, float t) { return a + (b - a) * t; }

static int d9_clip_line(d9v *from, d9v *to, const float *r) {
    int of = d9_outcode(*from, r), ot = d9_outcode(*to, r);
    for (int guard = 0; guard < 64; guard++) {
        if (!of && !ot) return 1;
        if (of & ot) return 0;
        int clip_from = of > ot;
        d9v pt = clip_from ? *from : *to;
        int o = clip_from ? of : ot;
        if (o & 1)
            pt = (d9v){r[0], d9_lerp(from->y, to->y, (r[0] - from->x) / (to->x - from->x))};
        else if (o & 2)
            pt = (d9v){r[2], d9_lerp(from->y, to->y, (r[2] - from->x) / (to->x - from->x))};
        else if (o & 4)
            pt = (d9v){d9_lerp(from->x, to->x, (r[1] - from->y) / (to->y - from->y)), r[1]};
        else if (o & 8)
            pt = (d9v){d9_lerp(from->x, to->x, (r[3] - from->y) / (to->y - from->y)), r[3]};
        if (clip_from) {
            *from = pt;
            of = d9_outcode(pt, r);
        } else {
            *to = pt;
            ot = d9_outcode(pt, r);
        }
    }
    return 0;
}

/* tiler.rs:189-301 process_line_segment. */
static void d9_line(d9_path *p, d9v from, d9v to, const float *clip_box) {
    if (!d9_clip_line(&from, &to, clip_box)) return;
    const float inv = 1.0f / (float)D9_TILE;
    int32_t ftx = (int32_t)floorf(from.x * inv), fty = (int32_t)floorf(from.y * inv);
    int32_t ttx = (int32_t)floorf(to.x * inv), tty = (int32_t)floorf(to.y * inv);
    float vx = to.x - from.x, vy = to.y - from.y;
    int32_t sx = vx < 0.0f ? -1 : 1, sy = vy < 0.0f ? -1 : 1;
    float fcx = (float)((ftx + (vx < 0.0f ? 0 : 1)) * D9_TILE);
    float fcy = (float)((fty + (vy < 0.0f ? 0 : 1)) * D9_TILE);
    float tmx = (fcx - from.x) / vx, tmy = (fcy - from.y) / vy;
    float tdx = fabsf((float)D9_TILE / vx), tdy = fabsf((float)D9_TILE / vy);
    d9v cur = from;
    int32_t tx = ftx, ty = fty;
    int last = 0; /* 0 none, 1 x, 2 y */
    for (uint32_t it = 0; it < D9_MAX_STEPS; it++) {
        int next;
        if (tmx < tmy) next = 1;
        else if (tmx > tmy) next = 2;
        else next = sx > 0 ? 1 : 2;
        float nt = next == 1 ? tmx : tmy;
        nt = nt < 1.0f ? nt : 1.0f;
        if (tx == ttx && ty == tty) next = 0;
        d9v pos = {from.x + vx * nt, from.y + vy * nt};
        d9_add_fill(p, cur, pos, tx, ty);
        d9v corner = {(float)(tx * D9_TILE), (float)(ty * D9_TILE)};
        if (sy < 0 && next == 2) d9_add_fill(p, pos, corner, tx, ty);
        else if (sy > 0 && last == 2) d9_add_fill(p, corner, cur, tx, ty);
        if (sx < 0 && last == 1) d9_adjust_backdrop(p, tx, ty, 1);
        else if (sx > 0 && next == 1) d9_adjust_backdrop(p, tx, ty, -1);
        if (next == 0) break;
        if (next == 1) {
            if (tx == ttx) break;
            tmx += tdx;
            tx += sx;
        } else {
            if (ty == tty) break;
            tmy += tdy;
            ty += sy;
        }
        cur = pos;
        last = next;
    }
}

/* segment.rs:292-300 CubicSegment::is_flat. */
static int d9_is_flat(const d9v c[4]) {
    float ux0 = ((3.0f * c[1].x - c[0].x) - c[0].x) - c[3].x;
    float uy0 = ((3.0f * c[1].y - c[0].y) - c[0].y) - c[3].y;
    float ux1 = ((3.0f * c[2].x - c[3].x) - c[3].x) - c[0].x;
    float uy1 = ((3.0f * c[2].y - c[3].y) - c[3].y) - c[0].y;
    ux0 *= ux0; uy0 *= uy0; ux1 *= ux1; uy1 *= uy1;
    float mx = ux0 > ux1 ? ux0 : ux1, my = uy0 > uy1 ? uy0 : uy1;
    return mx + my <= 16.0f * D9_TOLERANCE * D9_TOLERANCE;
}

/* tiler.rs:166-184 process_segment for a cubic: split at t = 0.5 (de
 * Casteljau, segment.rs:321-345) until flat, then the baseline. */
static void d9_cubic(d9_path *p, const d9v c[4], const float *clip_box, int depth) {
    if (d9_is_flat(c) || depth >= D9_MAX_DEPTH) {
        d9_line(p, c[0], c[3], clip_box);
        return;
    }
    const float t = 0.5f;
    d9v p01 = {c[0].x + t * (c[1].x - c[0].x), c[0].y + t * (c[1].y - c[0].y)};
    d9v p12 = {c[1].x + t * (c[2].x - c[1].x), c[1].y + t * (c[2].y - c[1].y)};
    d9v p23 = {c[2].x + t * (c[3].x - c[2].x), c[2].y + t * (c[3].y - c[2].y)};
    d9v p012 = {p01.x + t * (p12.x - p01.x), p01.y + t * (p12.y - p01.y)};
    d9v p123 = {p12.x + t * (p23.x - p12.x), p12.y + t * (p23.y - p12.y)};
    d9v m = {p012.x + t * (p123.x - p012.x), p012.y + t * (p123.y - p012.y)};
    d9v a[4] = {c[0], p01, p012, m}, b[4] = {m, p123, p23, c[3]};
    d9_cubic(p, a, clip_box, depth + 1);
    d9_cubic(p, b, clip_box, depth + 1);
}

/* Tiler::new + generate_fills (tiler.rs:40-92): one path. */
static void d9_tile_path(d9_path *p, const orc_d3d9_paths *src, uint32_t i, const float *xf,
                         const float *vb) {
    uint32_t s0 = i ? src->path_seg_end[i - 1] : 0, s1 = src->path_seg_end[i];
    /* transformed outline bounds (every point, control points included) */
    float bx0 = INFINITY, by0 = INFINITY, bx1 = -INFINITY, by1 = -INFINITY;
    for (uint32_t s = s0; s < s1; s++) {
        uint32_t first = src->indices[2 * s], flags = src->indices[2 * s + 1];
        uint32_t n = (flags & D9_FLAG_CUBIC) ? 4 : (flags & D9_FLAG_QUAD) ? 3 : 2;
        for (uint32_t k = 0; k < n; k++) {
            d9v q = d9_xform(xf, src->points[2 * (first + k)], src->points[2 * (first + k) + 1]);
            bx0 = q.x < bx0 ? q.x : bx0; by0 = q.y < by0 ? q.y : by0;
            bx1 = q.x > bx1 ? q.x : bx1; by1 = q.y > by1 ? q.y : by1;
        }
    }
    /* RectF::intersection with the view box (geometry/src/rect.rs:121-137:
     * strict overlap); none -> RectF::default(). An empty outline's bounds
     * are RectF::default() too. */
    if (s0 == s1) bx0 = by0 = bx1 = by1 = 0.0f;
    float ix0 = bx0 > vb[0] ? bx0 : vb[0], iy0 = by0 > vb[1] ? by0 : vb[1];
    float ix1 = bx1 < vb[2] ? bx1 : vb[2], iy1 = by1 < vb[3] ? by1 : vb[3];
    if (!(bx0 < vb[2] && by0 < vb[3] && vb[0] < bx1 && vb[1] < by1))
        ix0 = iy0 = ix1 = iy1 = 0.0f;
    /* tiles.rs:64-66 round_rect_out_to_tile_bounds */
    p->rx0 = (int32_t)floorf(ix0 / (float)D9_TILE);
    p->ry0 = (int32_t)floorf(iy0 / (float)D9_TILE);
    p->rx1 = (int32_t)ceilf(ix1 / (float)D9_TILE);
    p->ry1 = (int32_t)ceilf(iy1 / (float)D9_TILE);
    p->w = p->rx1 - p->rx0;
    p->h = p->ry1 - p->ry0;
    if (p->w < 0) p->w = 0;
    if (p->h < 0) p->h = 0;
    size_t nt = (size_t)p->w * p->h;
    p->alpha = (int32_t *)malloc((nt ? nt : 1) * sizeof(int32_t));
    p->backdrop = (int8_t *)calloc(nt ? nt : 1, 1);
    p->col_backdrop = (int32_t *)calloc(p->w ? p->w : 1, sizeof(int32_t));
    for (size_t k = 0; k < nt; k++) p->alpha[k] = D9_NO_TILE;
    if (!nt) return;
    /* tiler.rs:194-200: clip to the view box columns and its bottom */
    float clip_box[4] = {vb[0], -INFINITY, vb[2], vb[3]};
    for (uint32_t s = s0; s < s1; s++) {
        uint32_t first = src->indices[2 * s], flags = src->indices[2 * s + 1];
        d9v c[4];
        uint32_t n = (flags & D9_FLAG_CUBIC) ? 4 : (flags & D9_FLAG_QUAD) ? 3 : 2;
        for (uint32_t k = 0; k < n; k++)
            c[k] = d9_xform(xf, src->points[2 * (first + k)], src->points[2 * (first + k) + 1]);
        if (n == 2) {
            d9_line(p, c[0], c[1], clip_box);
        } else if (n == 3) {
            /* segment.rs:171-182 to_cubic */
            d9v p12 = {c[1].x + c[1].x, c[1].y + c[1].y};
            const float third = 1.0f / 3.0f;
            d9v q[4] = {c[0],
                        {(c[0].x + p12.x) * third, (c[0].y + p12.y) * third},
                        {(p12.x + c[2].x) * third, (p12.y + c[2].y) * third},
                        c[2]};
            d9_cubic(p, q, clip_box, 0);
        } else {
            d9_cubic(p, c, clip_box, 0);
        }
    }
}

/* Tiler::prepare_tiles (tiler.rs:94-163): backdrop propagation and clip
 * rules against the clip path's prepared tiles. */
static void d9_prepare(d9_path *p, const d9_path *clip) {
    for (int32_t i = 0; i < p->w * p->h; i++) {
        int32_t col = i % p->w;
        int32_t tx = p->rx0 + col, ty = p->ry0 + i / p->w;
        int delta = p->backdrop[i];
        int32_t draw_alpha = p->alpha[i];
        int8_t draw_backdrop = (int8_t)p->col_backdrop[col];
        if (clip) {
            int32_t cx = tx - clip->rx0, cy = ty - clip->ry0;
            if (cx >= 0 && cy >= 0 && cx < clip->w && cy < clip->h) {
                int32_t ci = cx + cy * clip->w;
                int32_t clip_alpha = clip->alpha[ci];
                int8_t clip_backdrop = clip->backdrop[ci];
                if (clip_alpha != D9_NO_TILE && draw_alpha != D9_NO_TILE) {
                    d9_clip_job j = {draw_alpha, draw_backdrop, clip_alpha, clip_backdrop};
                    p->jobs[p->n_jobs++] = j;
                    draw_backdrop = 0;
                } else if (clip_alpha != D9_NO_TILE && draw_alpha == D9_NO_TILE &&
                           draw_backdrop != 0) {
                    draw_alpha = clip_alpha;
                    draw_backdrop = clip_backdrop;
                } else if (clip_alpha == D9_NO_TILE && clip_backdrop == 0) {
                    draw_alpha = D9_NO_TILE;
                    draw_backdrop = 0;
                }
            } else {
                draw_alpha = D9_NO_TILE;
                draw_backdrop = 0;
            }
        }
        p->alpha[i] = draw_alpha;
        p->backdrop[i] = draw_backdrop;
        p->col_backdrop[col] += delta;
    }
}

static void d9_globalize(d9_path *p, int32_t base) {
    p->alpha_base = base;
    for (int32_t i = 0; i < p->w * p->h; i++)
        if (p->alpha[i] != D9_NO_TILE) p->alpha[i] += base;
}

static inline float d9_h2f(uint16_t h) { return orc_f16_to_f32(h); }
static inline uint16_t d9_f2h(float f) { return orc_f32_to_f16(f); }

/* d3d9/fill.{vs,fs}.glsl: every fill of a path as one instanced quad over
 * pixel columns floor(min x) .. ceil(max x) and mask rows floor(min y) / 4 ..
 * 3, each fragment adding computeCoverage to the binary16 texel. */
static void d9_raster_path(const orc_ctx *ctx, const d9_path *p, uint16_t *mask) {
    for (uint32_t k = 0; k < p->n_fills; k++) {
        const d9_fill *f = &p->fills[k];
        float fx = (float)f->x0 / 256.0f, fy = (float)f->y0 / 256.0f;
        float tx = (float)f->x1 / 256.0f, ty = (float)f->y1 / 256.0f;
        int32_t x0 = (int32_t)floorf(fx < tx ? fx : tx), x1 = (int32_t)ceilf(fx > tx ? fx : tx);
        int32_t r0 = (int32_t)floorf(floorf(fy < ty ? fy : ty) * 0.25f);
        uint16_t *tile = mask + (size_t)(p->alpha_base + f->local) * 64 * 4;
        for (int32_t r = r0; r < 4; r++) {
            for (int32_t x = x0; x < x1; x++) {
                if (x < 0 || x >= D9_TILE || r < 0) continue;
                float cov[4];
                float ox = (float)x + 0.5f, oy = (float)(4 * r) + 0.5f;
                orc_compute_coverage(ctx, fx - ox, fy - oy, tx - ox, ty - oy, cov);
                uint16_t *texel = tile + (size_t)(r * 16 + x) * 4;
                for (int c = 0; c < 4; c++) texel[c] = d9_f2h(d9_h2f(texel[c]) + cov[c]);
            }
        }
    }
}

/* tile_clip_copy + tile_clip_combine: dest = min(|dest + db|, |src + sb|)
 * from copies of both tiles taken before any of the batch's combines. */
static void d9_combine(const d9_path *p, uint16_t *mask, const uint16_t *copy_src) {
    for (uint32_t j = 0; j < p->n_jobs; j++) {
        const d9_clip_job *job = &p->jobs[j];
        uint16_t *dest = mask + (size_t)job->dest * 256;
        const uint16_t *src = copy_src + (size_t)job->src * 256;
        for (int k = 0; k < 256; k++) {
            float d = fabsf(d9_h2f(dest[k]) + (float)job->dest_backdrop);
            float s = fabsf(d9_h2f(src[k]) + (float)job->src_backdrop);
            dest[k] = d9_f2h(s < d ? s : d);
        }
    }
}

static inline uint32_t d9_unorm8(float x) {
    float c = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
    return (uint32_t)floorf(c * 255.0f + 0.5f);
}

/* Tiles every clip path, then every draw path, numbers alpha tiles path by
 * path and prepares the tiles (backdrops, clip jobs). */
static int d9_build_paths(const orc_d3d9_scene *s, int nthreads, d9_path **clips_out,
                          d9_path **draws_out, int32_t *n_alpha_out) {
    const float *xf = s->transform, *vb = s->view_box;
    uint32_t nc = s->clip.n_paths, nd = s->draw.n_paths;
    d9_path *clips = (d9_path *)calloc(nc ? nc : 1, sizeof(d9_path));
    d9_path *draws = (d9_path *)calloc(nd ? nd : 1, sizeof(d9_path));
    if (!clips || !draws) return -1;

    /* builder.rs build_clip_paths then build_paths: tile every path */
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int64_t i = 0; i < (int64_t)nc; i++) {
        d9_tile_path(&clips[i], &s->clip, (uint32_t)i, xf, vb);
        clips[i].fill_rule = s->clip.fill_rule[i];
        clips[i].clip = s->clip.clip[i];
    }
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int64_t i = 0; i < (int64_t)nd; i++) {
        d9_tile_path(&draws[i], &s->draw, (uint32_t)i, xf, vb);
        draws[i].fill_rule = s->draw.fill_rule[i];
        draws[i].clip = s->draw.clip[i];
        draws[i].paint = s->draw.paint[i];
    }
    /* alpha ids, path by path in build order */
    int32_t next = 0;
    for (uint32_t i = 0; i < nc; i++) { d9_globalize(&clips[i], next); next += clips[i].n_alpha; }
    for (uint32_t i = 0; i < nd; i++) { d9_globalize(&draws[i], next); next += draws[i].n_alpha; }
    int32_t n_alpha = next;
    /* prepare: clip paths without a clip of their own (build_clip_path_on_cpu,
     * builder.rs:261-278, hands the tiler no built clip paths, so a clip path
     * with a clip is not supported at this level: the reference would index
     * an empty slice), then draw paths against their clip path */
    for (uint32_t i = 0; i < nc; i++) {
        d9_path *p = &clips[i];
        p->jobs = (d9_clip_job *)malloc(sizeof(d9_clip_job));
        d9_prepare(p, NULL);
    }
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int64_t i = 0; i < (int64_t)nd; i++) {
        d9_path *p = &draws[i];
        p->jobs = (d9_clip_job *)malloc(((size_t)p->w * p->h + 1) * sizeof(d9_clip_job));
        const d9_path *c = (p->clip >= 0 && (uint32_t)p->clip < nc) ? &clips[p->clip] : NULL;
        d9_prepare(p, c);
    }
    *clips_out = clips;
    *draws_out = draws;
    *n_alpha_out = n_alpha;
    return 0;
}

static void d9_free_paths(d9_path *clips, uint32_t nc, d9_path *draws, uint32_t nd) {
    for (uint32_t i = 0; i < nc; i++) d9_free_path(&clips[i]);
    for (uint32_t i = 0; i < nd; i++) d9_free_path(&draws[i]);
    free(clips);
    free(draws);
}

int orc_d3d9_build(const orc_d3d9_scene *s, int threads, orc_d3d9_built *out) {
    memset(out, 0, sizeof(*out));
    uint32_t nc = s->clip.n_paths, nd = s->draw.n_paths;
    d9_path *clips, *draws;
    int32_t n_alpha;
    if (d9_build_paths(s, threads > 0 ? threads : 1, &clips, &draws, &n_alpha) != 0) return -1;
    /* fills: AddFillsD3D9 of every clip path, then every draw path */
    uint64_t nf = 0, ntiles = 0, njobs = 0;
    for (uint32_t i = 0; i < nc + nd; i++) nf += (i < nc ? &clips[i] : &draws[i - nc])->n_fills;
    for (uint32_t i = 0; i < nd; i++) {
        for (int32_t t = 0; t < draws[i].w * draws[i].h; t++)
            if (draws[i].alpha[t] != D9_NO_TILE || draws[i].backdrop[t] != 0) ntiles++;
        njobs += draws[i].n_jobs;
    }
    out->fills = (uint32_t *)malloc((nf ? nf : 1) * 12);
    out->tiles = (uint32_t *)malloc((ntiles ? ntiles : 1) * 16);
    out->clips = (int32_t *)malloc((njobs ? njobs : 1) * 16);
    /* z-buffer: DenseTileMap over the view box's tile rect (builder.rs:976-987) */
    int32_t zr[4] = {(int32_t)floorf(s->view_box[0] / (float)D9_TILE),
                     (int32_t)floorf(s->view_box[1] / (float)D9_TILE),
                     (int32_t)ceilf(s->view_box[2] / (float)D9_TILE),
                     (int32_t)ceilf(s->view_box[3] / (float)D9_TILE)};
    int64_t zw = zr[2] - zr[0] > 0 ? zr[2] - zr[0] : 0, zh = zr[3] - zr[1] > 0 ? zr[3] - zr[1] : 0;
    out->z = (int32_t *)calloc((zw * zh) > 0 ? zw * zh : 1, sizeof(int32_t));
    memcpy(out->z_rect, zr, sizeof(zr));
    if (!out->fills || !out->tiles || !out->clips || !out->z) return -1;
    for (uint32_t i = 0; i < nc + nd; i++) {
        const d9_path *p = i < nc ? &clips[i] : &draws[i - nc];
        for (uint32_t k = 0; k < p->n_fills; k++) {
            const d9_fill *f = &p->fills[k];
            uint32_t *w = out->fills + (size_t)out->n_fills++ * 3;
            w[0] = (uint32_t)f->x0 | ((uint32_t)f->y0 << 16);
            w[1] = (uint32_t)f->x1 | ((uint32_t)f->y1 << 16);
            w[2] = (uint32_t)(p->alpha_base + f->local);
        }
    }
    /* DrawTileBatchD3D9 (builder.rs:998-1041): every tile with an alpha tile
     * or a backdrop, path by path; z = the last occluding path over a solid
     * tile; clip tiles where both tiles exist */
    for (uint32_t i = 0; i < nd; i++) {
        const d9_path *p = &draws[i];
        uint32_t pc = p->paint < s->n_paints ? p->paint : 0;
        int occludes = s->n_paints && s->paints_rgba[pc * 4 + 3] == 255;
        for (int32_t t = 0; t < p->w * p->h; t++) {
            if (p->alpha[t] == D9_NO_TILE && p->backdrop[t] == 0) continue;
            int32_t tx = p->rx0 + t % p->w, ty = p->ry0 + t / p->w;
            uint32_t *w = out->tiles + (size_t)out->n_tiles++ * 4;
            w[0] = (uint32_t)(uint16_t)tx | ((uint32_t)(uint16_t)ty << 16);
            w[1] = p->alpha[t] == D9_NO_TILE ? 0xffffffffu : (uint32_t)p->alpha[t];
            w[2] = i;
            w[3] = (p->paint & 0xffffu) | ((uint32_t)(p->fill_rule ? 2 : 1) << 16) |
                   ((uint32_t)(uint8_t)p->backdrop[t] << 24);
            if (occludes && p->alpha[t] == D9_NO_TILE) {
                int64_t zx = tx - zr[0], zy = ty - zr[1];
                if (zx >= 0 && zy >= 0 && zx < zw && zy < zh) {
                    int32_t *zv = &out->z[zy * zw + zx];
                    if ((int32_t)i > *zv) *zv = (int32_t)i;
                }
            }
        }
        for (uint32_t j = 0; j < p->n_jobs; j++) {
            int32_t *w = out->clips + (size_t)out->n_clips++ * 4;
            w[0] = p->jobs[j].dest;
            w[1] = p->jobs[j].dest_backdrop;
            w[2] = p->jobs[j].src;
            w[3] = p->jobs[j].src_backdrop;
        }
    }
    out->n_alpha = (uint32_t)n_alpha;
    d9_free_paths(clips, nc, draws, nd);
    return 0;
}

void orc_d3d9_built_free(orc_d3d9_built *b) {
    free(b->fills); free(b->tiles); free(b->clips); free(b->z);
    memset(b, 0, sizeof(*b));
}

int orc_d3d9_render(const orc_ctx *ctx, const orc_d3d9_scene *s, uint32_t *fb, int threads) {
    int nthreads = threads > 0 ? threads : 1;
    uint32_t nc = s->clip.n_paths, nd = s->draw.n_paths;
    d9_path *clips, *draws;
    int32_t n_alpha;
    double t_start = orc_now();
    if (d9_build_paths(s, nthreads, &clips, &draws, &n_alpha) != 0) return -1;
    double t_tiled = orc_now();
    orc_stage_seconds[ORC_STAGE_D3D9_TILER] += t_tiled - t_start;
    /* masks: binary16 RGBA, 64 texels per alpha tile, cleared */
    uint16_t *mask = (uint16_t *)calloc(((size_t)n_alpha + 1) * 256, sizeof(uint16_t));
    if (!mask) return -1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int64_t i = 0; i < (int64_t)(nc + nd); i++)
        d9_raster_path(ctx, i < (int64_t)nc ? &clips[i] : &draws[i - nc], mask);
    uint16_t *copy = (uint16_t *)malloc(((size_t)n_alpha + 1) * 256 * sizeof(uint16_t));
    if (!copy) return -1;
    memcpy(copy, mask, ((size_t)n_alpha + 1) * 256 * sizeof(uint16_t));
    for (uint32_t i = 0; i < nd; i++) d9_combine(&draws[i], mask, copy);
    free(copy);
    double t_masked = orc_now();
    orc_stage_seconds[ORC_STAGE_D3D9_RASTER] += t_masked - t_tiled;

    /* builder.rs:1013-1028: z-buffer of occluding solid tiles; per-framebuffer
     * tile lists in path order (tile.vs.glsl culls a tile whose path index is
     * below the z value) */
    int32_t fbw = (s->width + D9_TILE - 1) / D9_TILE, fbh = (s->height + D9_TILE - 1) / D9_TILE;
    int64_t fbn = (int64_t)fbw * fbh;
    int32_t *z = (int32_t *)calloc(fbn ? fbn : 1, sizeof(int32_t));
    uint32_t *counts = (uint32_t *)calloc(fbn + 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < nd; i++) {
        d9_path *p = &draws[i];
        uint32_t pc = p->paint < s->n_paints ? p->paint : 0;
        p->occludes = s->n_paints && s->paints_rgba[pc * 4 + 3] == 255;
        for (int32_t t = 0; t < p->w * p->h; t++) {
            if (p->alpha[t] == D9_NO_TILE && p->backdrop[t] == 0) continue;
            int32_t tx = p->rx0 + t % p->w, ty = p->ry0 + t / p->w;
            if (tx < 0 || ty < 0 || tx >= fbw || ty >= fbh) continue;
            counts[(int64_t)ty * fbw + tx]++;
            if (p->occludes && p->alpha[t] == D9_NO_TILE) {
                int32_t *zv = &z[(int64_t)ty * fbw + tx];
                if ((int32_t)i > *zv) *zv = (int32_t)i;
            }
        }
    }
    uint32_t *start = (uint32_t *)malloc((fbn + 1) * sizeof(uint32_t));
    uint32_t acc = 0;
    for (int64_t g = 0; g < fbn; g++) { start[g] = acc; acc += counts[g]; counts[g] = 0; }
    start[fbn] = acc;
    uint64_t *lists = (uint64_t *)malloc((acc ? acc : 1) * sizeof(uint64_t));
    for (uint32_t i = 0; i < nd; i++) {
        d9_path *p = &draws[i];
        for (int32_t t = 0; t < p->w * p->h; t++) {
            if (p->alpha[t] == D9_NO_TILE && p->backdrop[t] == 0) continue;
            int32_t tx = p->rx0 + t % p->w, ty = p->ry0 + t / p->w;
            if (tx < 0 || ty < 0 || tx >= fbw || ty >= fbh) continue;
            int64_t g = (int64_t)ty * fbw + tx;
            lists[start[g] + counts[g]++] = ((uint64_t)i << 32) | (uint32_t)t;
        }
    }
    /* paints: ColorU::to_f32 then binary16 (gpu/renderer.rs:724-727) */
    float *base = (float *)malloc(((size_t)s->n_paints + 1) * 4 * sizeof(float));
    const float inv255 = 1.0f / 255.0f;
    for (uint32_t k = 0; k < 4 * s->n_paints; k++)
        base[k] = orc_f16_to_f32(orc_f32_to_f16((float)s->paints_rgba[k] * inv255));

    /* d3d9/tile.{vs,fs}.glsl + calculateColor (tile_fragment.inc.glsl:539-614),
     * blended src-over into the RGBA8 target one tile at a time */
    int64_t pitch = (int64_t)fbw * D9_TILE;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 8)
    for (int64_t g = 0; g < fbn; g++) {
        int32_t gx = (int32_t)(g % fbw), gy = (int32_t)(g / fbw);
        for (int py = 0; py < D9_TILE; py++) {
            for (int px = 0; px < D9_TILE; px++) {
                uint32_t *out = fb + (size_t)(gy * D9_TILE + py) * pitch + gx * D9_TILE + px;
                uint32_t d8 = 0;
                if (s->has_background)
                    for (int c = 0; c < 4; c++) d8 |= d9_unorm8(s->background[c]) << (8 * c);
                for (uint32_t e = start[g]; e < start[g + 1]; e++) {
                    uint32_t pi = (uint32_t)(lists[e] >> 32), t = (uint32_t)lists[e];
                    if ((int32_t)pi < z[g]) continue;
                    const d9_path *p = &draws[pi];
                    float texel = 0.0f;
                    if (p->alpha[t] != D9_NO_TILE)
                        texel = d9_h2f(mask[(size_t)p->alpha[t] * 256 +
                                            (size_t)((py >> 2) * 16 + px) * 4 + (py & 3)]);
                    float cov = texel + (float)p->backdrop[t];
                    if (p->fill_rule == 0) {
                        cov = fabsf(cov);
                    } else {
                        float m = cov - 2.0f * floorf(cov / 2.0f);
                        cov = 1.0f - fabsf(1.0f - m);
                    }
                    float mask_alpha = cov < 1.0f ? cov : 1.0f;
                    uint32_t pc = p->paint < s->n_paints ? p->paint : 0;
                    const float *b = base + (size_t)pc * 4;
                    float a = b[3] * mask_alpha;
                    uint32_t o = 0;
                    for (int c = 0; c < 4; c++) {
                        float dc = (float)((d8 >> (8 * c)) & 0xffu) / 255.0f;
                        float sc = c < 3 ? b[c] * a : a;
                        o |= d9_unorm8(sc + dc * (1.0f - a)) << (8 * c);
                    }
                    d8 = o;
                }
                *out = d8;
            }
        }
    }
    orc_stage_seconds[ORC_STAGE_D3D9_COMPOSITE] += orc_now() - t_masked;
    free(base); free(lists); free(start); free(counts); free(z); free(mask);
    d9_free_paths(clips, nc, draws, nd);
    return 0;
}
