// pfh_d3d9.cpp - the D3D9 level's CPU tiler (see pfh_d3d9.h for the
// reference functions each part follows).
#include "pfh_d3d9.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

namespace pfh9 {
namespace {

constexpr int32_t kTile = 16;                 // TILE_WIDTH, TILE_HEIGHT
constexpr uint32_t kFlagQuad = 0x80000000u;   // SegmentIndicesD3D11 flags
constexpr uint32_t kFlagCubic = 0x40000000u;
constexpr float kTolerance = 0.25f;           // FLATTENING_TOLERANCE
constexpr uint32_t kNoTile = 0xffffffffu;     // AlphaTileId(!0)
// process_segment recurses until a piece is flat, which a finite curve is
// after a few dozen halvings; the cap only stops non-finite input
constexpr int kMaxDepth = 24;
constexpr int kMaxClipRounds = 64;            // Cohen-Sutherland ends within 4

struct V {
    float x, y;
};

// One path's BuiltPathBinCPUData: the tile rect, per tile the alpha tile
// (local number until numbered globally) and backdrop, the column
// backdrops above the rect, its fills and (draw paths) clip jobs.
struct PathTiles {
    int32_t x0 = 0, y0 = 0, w = 0, h = 0;
    std::vector<uint32_t> alpha;
    std::vector<int8_t> backdrop;
    std::vector<int32_t> column;
    std::vector<pfh_fill_d3d9> fills;
    uint32_t n_alpha = 0;
    std::vector<pfh_clip_d3d9> jobs;
};

inline V xform(const pfh_transform2f &t, const pfh_vector2f &p) {
    return V{(t.m00 * p.x + t.m01 * p.y) + t.tx, (t.m10 * p.x + t.m11 * p.y) + t.ty};
}

inline float lerp(float a, float b, float t) { return a + (b - a) * t; }

// ObjectBuilder::add_fill
void add_fill(PathTiles &p, V from, V to, int32_t tx, int32_t ty) {
    if (tx < p.x0 || ty < p.y0 || tx >= p.x0 + p.w || ty >= p.y0 + p.h) return;
    float ox = (float)tx * (float)kTile, oy = (float)ty * (float)kTile;
    float v[4] = {(from.x - ox) * 256.0f, (from.y - oy) * 256.0f, (to.x - ox) * 256.0f,
                  (to.y - oy) * 256.0f};
    int32_t q[4];
    for (int i = 0; i < 4; i++) {
        // F32x4::clamp = max then min (the SSE forms: a NaN lane takes the bound)
        float c = v[i] > 0.0f ? v[i] : 0.0f;
        c = c < (float)(kTile * 256 - 1) ? c : (float)(kTile * 256 - 1);
        q[i] = (int32_t)std::nearbyint(c);  // to_i32x4: round to nearest even
    }
    if (q[0] == q[2]) return;
    uint32_t local = (uint32_t)((tx - p.x0) + (ty - p.y0) * p.w);
    if (p.alpha[local] == kNoTile) p.alpha[local] = p.n_alpha++;
    p.fills.push_back(pfh_fill_d3d9{(uint16_t)q[0], (uint16_t)q[1], (uint16_t)q[2],
                                    (uint16_t)q[3], p.alpha[local]});
}

// ObjectBuilder::adjust_alpha_tile_backdrop
void adjust_backdrop(PathTiles &p, int32_t tx, int32_t ty, int8_t delta) {
    int32_t ox = tx - p.x0, oy = ty - p.y0;
    if (ox < 0 || ox >= p.w || oy >= p.h) return;
    if (oy < 0) {
        p.column[ox] += delta;
        return;
    }
    int8_t &b = p.backdrop[(size_t)ox + (size_t)oy * p.w];
    b = (int8_t)(b + delta);
}

// clip_line_segment_to_rect (Cohen-Sutherland)
enum : int { kLeft = 1, kRight = 2, kTop = 4, kBottom = 8 };

int outcode(V pt, const float r[4]) {
    int o = 0;
    if (pt.x < r[0]) o |= kLeft;
    if (pt.y < r[1]) o |= kTop;
    if (pt.x > r[2]) o |= kRight;
    if (pt.y > r[3]) o |= kBottom;
    return o;
}

bool clip_line(V &from, V &to, const float r[4]) {
    int of = outcode(from, r), ot = outcode(to, r);
    for (int round = 0; round < kMaxClipRounds; round++) {
        if (!of && !ot) return true;
        if (of & ot) return false;
        bool clip_from = of > ot;
        int o = clip_from ? of : ot;
        V pt = clip_from ? from : to;
        if (o & kLeft)
            pt = V{r[0], lerp(from.y, to.y, (r[0] - from.x) / (to.x - from.x))};
        else if (o & kRight)
            pt = V{r[2], lerp(from.y, to.y, (r[2] - from.x) / (to.x - from.x))};
        else if (o & kTop)
            pt = V{lerp(from.x, to.x, (r[1] - from.y) / (to.y - from.y)), r[1]};
        else if (o & kBottom)
            pt = V{lerp(from.x, to.x, (r[3] - from.y) / (to.y - from.y)), r[3]};
        if (clip_from) {
            from = pt;
            of = outcode(pt, r);
        } else {
            to = pt;
            ot = outcode(pt, r);
        }
    }
    return false;
}

// process_line_segment: the line clipped to the view box's columns and
// bottom, then walked tile by tile (Amanatides-Woo) with the fills of each
// tile it crosses, the auxiliary fills along tile tops, and the backdrop
// steps where it crosses a tile's right edge.
void process_line(PathTiles &p, V from, V to, const float clip_box[4]) {
    if (!clip_line(from, to, clip_box)) return;
    const float inv = 1.0f / (float)kTile;
    int32_t ftx = (int32_t)std::floor(from.x * inv), fty = (int32_t)std::floor(from.y * inv);
    int32_t ttx = (int32_t)std::floor(to.x * inv), tty = (int32_t)std::floor(to.y * inv);
    float vx = to.x - from.x, vy = to.y - from.y;
    bool neg_x = vx < 0.0f, neg_y = vy < 0.0f;
    int32_t sx = neg_x ? -1 : 1, sy = neg_y ? -1 : 1;
    float cross_x = (float)((ftx + (neg_x ? 0 : 1)) * kTile);
    float cross_y = (float)((fty + (neg_y ? 0 : 1)) * kTile);
    float t_max_x = (cross_x - from.x) / vx, t_max_y = (cross_y - from.y) / vy;
    float t_delta_x = std::fabs((float)kTile / vx), t_delta_y = std::fabs((float)kTile / vy);
    V cur = from;
    int32_t tx = ftx, ty = fty;
    enum { kNone, kX, kY } last = kNone;
    for (;;) {
        int next;
        if (t_max_x < t_max_y) next = kX;
        else if (t_max_x > t_max_y) next = kY;
        else next = sx > 0 ? kX : kY;
        float t = next == kX ? t_max_x : t_max_y;
        t = std::fmin(t, 1.0f);
        if (tx == ttx && ty == tty) next = kNone;
        V pos{from.x + vx * t, from.y + vy * t};
        add_fill(p, cur, pos, tx, ty);
        V corner{(float)tx * (float)kTile, (float)ty * (float)kTile};
        if (sy < 0 && next == kY) add_fill(p, pos, corner, tx, ty);
        else if (sy > 0 && last == kY) add_fill(p, corner, cur, tx, ty);
        if (sx < 0 && last == kX) adjust_backdrop(p, tx, ty, 1);
        else if (sx > 0 && next == kX) adjust_backdrop(p, tx, ty, -1);
        if (next == kNone) break;
        if (next == kX) {
            if (tx == ttx) break;
            t_max_x += t_delta_x;
            tx += sx;
        } else {
            if (ty == tty) break;
            t_max_y += t_delta_y;
            ty += sy;
        }
        cur = pos;
        last = next == kX ? kX : kY;
    }
}

// CubicSegment::is_flat: max over both control points of
// (3 ctrl - 2 end - other end)^2 per axis, summed, against 16 tol^2.
bool is_flat(const V c[4]) {
    float ux0 = ((3.0f * c[1].x - c[0].x) - c[0].x) - c[3].x;
    float uy0 = ((3.0f * c[1].y - c[0].y) - c[0].y) - c[3].y;
    float ux1 = ((3.0f * c[2].x - c[3].x) - c[3].x) - c[0].x;
    float uy1 = ((3.0f * c[2].y - c[3].y) - c[3].y) - c[0].y;
    ux0 *= ux0;
    uy0 *= uy0;
    ux1 *= ux1;
    uy1 *= uy1;
    float mx = ux0 > ux1 ? ux0 : ux1, my = uy0 > uy1 ? uy0 : uy1;  // _mm_max_ps
    return mx + my <= 16.0f * kTolerance * kTolerance;
}

// process_segment for a cubic: the baseline once flat, else both halves of
// the de Casteljau split at t = 1/2.
void process_cubic(PathTiles &p, const V c[4], const float clip_box[4], int depth) {
    if (is_flat(c) || depth >= kMaxDepth) {
        process_line(p, c[0], c[3], clip_box);
        return;
    }
    const float t = 0.5f;
    V p01{c[0].x + t * (c[1].x - c[0].x), c[0].y + t * (c[1].y - c[0].y)};
    V p12{c[1].x + t * (c[2].x - c[1].x), c[1].y + t * (c[2].y - c[1].y)};
    V p23{c[2].x + t * (c[3].x - c[2].x), c[2].y + t * (c[3].y - c[2].y)};
    V p012{p01.x + t * (p12.x - p01.x), p01.y + t * (p12.y - p01.y)};
    V p123{p12.x + t * (p23.x - p12.x), p12.y + t * (p23.y - p12.y)};
    V m{p012.x + t * (p123.x - p012.x), p012.y + t * (p123.y - p012.y)};
    const V a[4] = {c[0], p01, p012, m}, b[4] = {m, p123, p23, c[3]};
    process_cubic(p, a, clip_box, depth + 1);
    process_cubic(p, b, clip_box, depth + 1);
}

// Tiler::new (the tile rect of the transformed outline's bounds within the
// view box) and generate_fills.
void tile_path(PathTiles &p, const pfh_vector2f *pts, const pfh_segment_indices *idx,
               const PathInput &in, const pfh_transform2f &xf, const float vb[4]) {
    float bx0 = INFINITY, by0 = INFINITY, bx1 = -INFINITY, by1 = -INFINITY;
    auto npoints = [](uint32_t flags) {
        return (flags & kFlagCubic) ? 4u : (flags & kFlagQuad) ? 3u : 2u;
    };
    for (uint32_t s = in.seg_begin; s < in.seg_end; s++) {
        for (uint32_t k = 0; k < npoints(idx[s].flags); k++) {
            V q = xform(xf, pts[idx[s].first_point_index + k]);
            bx0 = std::min(bx0, q.x);
            by0 = std::min(by0, q.y);
            bx1 = std::max(bx1, q.x);
            by1 = std::max(by1, q.y);
        }
    }
    if (in.seg_begin == in.seg_end) bx0 = by0 = bx1 = by1 = 0.0f;  // Outline::bounds of nothing
    // RectF::intersection (strict overlap), RectF::default() when there is none
    float ix0 = 0, iy0 = 0, ix1 = 0, iy1 = 0;
    if (bx0 < vb[2] && by0 < vb[3] && vb[0] < bx1 && vb[1] < by1) {
        ix0 = std::max(bx0, vb[0]);
        iy0 = std::max(by0, vb[1]);
        ix1 = std::min(bx1, vb[2]);
        iy1 = std::min(by1, vb[3]);
    }
    // round_rect_out_to_tile_bounds
    const float s16 = 1.0f / (float)kTile;
    p.x0 = (int32_t)std::floor(ix0 * s16);
    p.y0 = (int32_t)std::floor(iy0 * s16);
    p.w = std::max((int32_t)std::ceil(ix1 * s16) - p.x0, 0);
    p.h = std::max((int32_t)std::ceil(iy1 * s16) - p.y0, 0);
    size_t n = (size_t)p.w * p.h;
    p.alpha.assign(n, kNoTile);
    p.backdrop.assign(n, 0);
    p.column.assign((size_t)p.w, 0);
    if (!n) return;
    const float clip_box[4] = {vb[0], -INFINITY, vb[2], vb[3]};
    for (uint32_t s = in.seg_begin; s < in.seg_end; s++) {
        uint32_t first = idx[s].first_point_index, flags = idx[s].flags;
        V c[4];
        uint32_t np = npoints(flags);
        for (uint32_t k = 0; k < np; k++) c[k] = xform(xf, pts[first + k]);
        if (np == 2) {
            process_line(p, c[0], c[1], clip_box);
        } else if (np == 3) {
            // Segment::to_cubic: (p0 + 2 c) / 3 and (2 c + p3) / 3 as products by 1/3
            V c2{c[1].x + c[1].x, c[1].y + c[1].y};
            const float third = 1.0f / 3.0f;
            const V q[4] = {c[0],
                            {(c[0].x + c2.x) * third, (c[0].y + c2.y) * third},
                            {(c2.x + c[2].x) * third, (c2.y + c[2].y) * third},
                            c[2]};
            process_cubic(p, q, clip_box, 0);
        } else {
            process_cubic(p, c, clip_box, 0);
        }
    }
}

// Tiler::prepare_tiles: backdrops propagated down each column, and the clip
// path's tiles applied to a draw path's.
void prepare(PathTiles &p, const PathTiles *clip) {
    for (int32_t i = 0; i < p.w * p.h; i++) {
        int32_t col = i % p.w;
        int32_t tx = p.x0 + col, ty = p.y0 + i / p.w;
        int32_t delta = p.backdrop[i];
        uint32_t draw_alpha = p.alpha[i];
        int8_t draw_backdrop = (int8_t)p.column[col];
        if (clip) {
            int32_t cx = tx - clip->x0, cy = ty - clip->y0;
            if (cx >= 0 && cy >= 0 && cx < clip->w && cy < clip->h) {
                size_t ci = (size_t)cx + (size_t)cy * clip->w;
                uint32_t clip_alpha = clip->alpha[ci];
                int8_t clip_backdrop = clip->backdrop[ci];
                if (clip_alpha != kNoTile && draw_alpha != kNoTile) {
                    p.jobs.push_back(pfh_clip_d3d9{draw_alpha, draw_backdrop, clip_alpha,
                                                   clip_backdrop});
                    draw_backdrop = 0;
                } else if (clip_alpha != kNoTile && draw_alpha == kNoTile && draw_backdrop != 0) {
                    draw_alpha = clip_alpha;
                    draw_backdrop = clip_backdrop;
                } else if (clip_alpha == kNoTile && clip_backdrop == 0) {
                    draw_alpha = kNoTile;
                    draw_backdrop = 0;
                }
            } else {
                draw_alpha = kNoTile;
                draw_backdrop = 0;
            }
        }
        p.alpha[i] = draw_alpha;
        p.backdrop[i] = draw_backdrop;
        p.column[col] += delta;
    }
}

template <typename F>
void parallel_for(size_t n, int threads, F fn) {
    int t = (int)std::min<size_t>((size_t)std::max(threads, 1), (n + 7) / 8);
    if (t <= 1) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < t; k++) pool.emplace_back(work);
    work();
    for (std::thread &th : pool) th.join();
}

}  // namespace

void build(const pfh_vector2f *clip_points, const pfh_segment_indices *clip_indices,
           const std::vector<PathInput> &clip_paths, const pfh_vector2f *draw_points,
           const pfh_segment_indices *draw_indices, const std::vector<PathInput> &draw_paths,
           const pfh_transform2f &transform, const float view_box[4], int threads, Built *out) {
    *out = Built();
    size_t nc = clip_paths.size(), nd = draw_paths.size();
    std::vector<PathTiles> clips(nc), draws(nd);
    // build_paths_on_cpu: clip paths, then draw paths
    parallel_for(nc + nd, threads, [&](size_t i) {
        if (i < nc)
            tile_path(clips[i], clip_points, clip_indices, clip_paths[i], transform, view_box);
        else
            tile_path(draws[i - nc], draw_points, draw_indices, draw_paths[i - nc], transform,
                      view_box);
    });
    // alpha tiles numbered path by path in build order
    uint32_t base = 0;
    out->fills.resize(nc + nd);
    for (size_t i = 0; i < nc + nd; i++) {
        PathTiles &p = i < nc ? clips[i] : draws[i - nc];
        for (uint32_t &a : p.alpha)
            if (a != kNoTile) a += base;
        std::vector<pfh_fill_d3d9> &f = out->fills[i];
        f = std::move(p.fills);
        for (pfh_fill_d3d9 &x : f) x.alpha_tile_index += base;
        base += p.n_alpha;
    }
    out->alpha_tile_count = base;
    parallel_for(nc, threads, [&](size_t i) { prepare(clips[i], nullptr); });
    parallel_for(nd, threads, [&](size_t i) {
        int32_t c = draw_paths[i].clip_path;
        prepare(draws[i], c >= 0 && (size_t)c < nc ? &clips[c] : nullptr);
    });
    // the draw batch: every tile with an alpha tile or a backdrop, in draw
    // order; z = the last occluding path over a solid tile; clip tiles
    const float s16 = 1.0f / (float)kTile;
    pfh_rect_i zr{(int32_t)std::floor(view_box[0] * s16), (int32_t)std::floor(view_box[1] * s16),
                  (int32_t)std::ceil(view_box[2] * s16), (int32_t)std::ceil(view_box[3] * s16)};
    int64_t zw = std::max(zr.max_x - zr.min_x, 0), zh = std::max(zr.max_y - zr.min_y, 0);
    out->z_rect = zr;
    out->z.assign((size_t)(zw * zh), 0);
    for (size_t i = 0; i < nd; i++) {
        const PathTiles &p = draws[i];
        const PathInput &in = draw_paths[i];
        uint8_t ctrl = in.fill_rule == PFH_FILL_RULE_EVEN_ODD ? 2 : 1;
        for (int32_t t = 0; t < p.w * p.h; t++) {
            if (p.alpha[t] == kNoTile && p.backdrop[t] == 0) continue;
            int32_t tx = p.x0 + t % p.w, ty = p.y0 + t / p.w;
            out->tiles.push_back(pfh_tile_object_d3d9{(int16_t)tx, (int16_t)ty, p.alpha[t],
                                                      (uint32_t)i, (uint16_t)in.paint, ctrl,
                                                      p.backdrop[t]});
            if (in.occludes && p.alpha[t] == kNoTile) {
                int64_t zx = tx - zr.min_x, zy = ty - zr.min_y;
                if (zx >= 0 && zy >= 0 && zx < zw && zy < zh) {
                    int32_t &zv = out->z[(size_t)(zy * zw + zx)];
                    zv = std::max(zv, (int32_t)i);
                }
            }
        }
        out->clips.insert(out->clips.end(), p.jobs.begin(), p.jobs.end());
    }
}

}  // namespace pfh9
