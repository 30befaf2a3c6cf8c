// pfh_palette.cpp - gradient and pattern paints and the palette build.
//
// renderer/src/paint.rs: build_paint_info :398-437, assign_paint_locations
// :453-583, calculate_texture_transforms :585-639, create_texture_metadata
// :641-659, allocate_textures :661-673, create_render_commands :675-692,
// GradientTileBuilder :813-880, PaintMetadata::filter :747-769.
// content/src/gradient.rs: add :159-164, sample :188-213, is_opaque :215.
// Transforms follow geometry/src/transform2d.rs (Matrix2x2F stored column by
// column, products as separate multiplies and adds, inverse through the
// adjugate times 1/det), so every float matches the reference's.
#include "pfh_palette.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

namespace pfh {
namespace {

constexpr int32_t kGradientTileLength = 256;  // GRADIENT_TILE_LENGTH, paint.rs:35

// Transform2F as the reference stores it: matrix lanes (m00, m10, m01, m11)
// (column-major), then the translation.
struct Xform {
    float s[4];
    float tx, ty;
};

Xform from_c(const pfh_transform2f &t) { return Xform{{t.m00, t.m10, t.m01, t.m11}, t.tx, t.ty}; }

pfh_transform2f to_c(const Xform &x) {
    return pfh_transform2f{x.s[0], x.s[2], x.s[1], x.s[3], x.tx, x.ty};
}

// Matrix2x2F * Matrix2x2F (transform2d.rs:115-121):
// self.xyxy * other.xxzz + self.zwzw * other.yyww
void mat_mul(const float a[4], const float b[4], float out[4]) {
    float r0 = a[0] * b[0] + a[2] * b[1];
    float r1 = a[1] * b[0] + a[3] * b[1];
    float r2 = a[0] * b[2] + a[2] * b[3];
    float r3 = a[1] * b[2] + a[3] * b[3];
    out[0] = r0; out[1] = r1; out[2] = r2; out[3] = r3;
}

// Matrix2x2F * Vector2F (:123-130): halves = m * (x, x, y, y); halves.xy + halves.zw
void mat_vec(const float a[4], float x, float y, float *ox, float *oy) {
    float h0 = a[0] * x, h1 = a[1] * x, h2 = a[2] * y, h3 = a[3] * y;
    *ox = h0 + h2;
    *oy = h1 + h3;
}

// Transform2F * Transform2F (:301-310)
Xform xmul(const Xform &a, const Xform &b) {
    Xform r;
    mat_mul(a.s, b.s, r.s);
    float vx, vy;
    mat_vec(a.s, b.tx, b.ty, &vx, &vy);
    r.tx = vx + a.tx;
    r.ty = vy + a.ty;
    return r;
}

// Transform2F::inverse (:294-298) with Matrix2x2F::{det, adjugate, inverse} (:60-72)
Xform xinverse(const Xform &a) {
    float det = a.s[0] * a.s[3] - a.s[2] * a.s[1];
    float r = 1.0f / det;
    Xform o;
    o.s[0] = r * a.s[3];
    o.s[1] = r * -a.s[1];
    o.s[2] = r * -a.s[2];
    o.s[3] = r * a.s[0];
    float vx, vy;
    mat_vec(o.s, a.tx, a.ty, &vx, &vy);
    o.tx = -vx;
    o.ty = -vy;
    return o;
}

Xform identity() { return Xform{{1, 0, 0, 1}, 0, 0}; }

// Transform2F::from_scale(scale).translate(origin) = from_translation(origin) * from_scale(scale)
Xform scale_then_translate(float sx, float sy, float ox, float oy) {
    Xform s{{sx, 0, 0, sy}, 0, 0};
    Xform t{{1, 0, 0, 1}, ox, oy};
    return xmul(t, s);
}

template <typename T>
void put(std::string &k, const T &v) {
    k.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

bool filter_valid_for_pattern(const pfh_filter &f) {
    return f.kind == PFH_FILTER_NONE || f.kind == PFH_FILTER_BLUR ||
           f.kind == PFH_FILTER_COLOR_MATRIX;
}

}  // namespace

bool paint_from_gradient(const pfh_color_u &base, const pfh_gradient &g, int32_t composite_op,
                         PaintDesc *out) {
    if (g.geometry != PFH_GRADIENT_LINEAR && g.geometry != PFH_GRADIENT_RADIAL) return false;
    if (g.wrap != PFH_GRADIENT_WRAP_CLAMP && g.wrap != PFH_GRADIENT_WRAP_REPEAT) return false;
    if (g.stop_count && !g.stops) return false;
    if (composite_op != PFH_PAINT_COMPOSITE_SRC_IN && composite_op != PFH_PAINT_COMPOSITE_DEST_IN)
        return false;
    PaintDesc p;
    p.base = base;
    p.overlay = OVERLAY_GRADIENT;
    p.composite_op = composite_op;
    p.gradient.geometry = g.geometry;
    p.gradient.from = g.from;
    p.gradient.to = g.to;
    p.gradient.wrap = g.wrap;
    if (g.geometry == PFH_GRADIENT_RADIAL) {
        p.gradient.radius0 = g.radius0;
        p.gradient.radius1 = g.radius1;
        p.gradient.transform = g.transform;
    }
    // Gradient::add inserts after every stop whose offset is <= the new one:
    // a stable sort by offset
    p.gradient.stops.assign(g.stops, g.stops + g.stop_count);
    std::stable_sort(p.gradient.stops.begin(), p.gradient.stops.end(),
                     [](const pfh_color_stop &a, const pfh_color_stop &b) {
                         return a.offset < b.offset;
                     });
    *out = std::move(p);
    return true;
}

bool paint_from_pattern(const pfh_color_u &base, const pfh_pattern &pt, int32_t composite_op,
                        PaintDesc *out) {
    if (pt.width <= 0 || pt.height <= 0 || pt.width > 16382 || pt.height > 16382 || !pt.pixels)
        return false;
    if (!filter_valid_for_pattern(pt.filter)) return false;
    if (composite_op != PFH_PAINT_COMPOSITE_SRC_IN && composite_op != PFH_PAINT_COMPOSITE_DEST_IN)
        return false;
    PaintDesc p;
    p.base = base;
    p.overlay = OVERLAY_PATTERN;
    p.composite_op = composite_op;
    p.pattern.width = pt.width;
    p.pattern.height = pt.height;
    p.pattern.pixels.assign(pt.pixels, pt.pixels + (size_t)pt.width * pt.height);
    p.pattern.transform = pt.transform;
    p.pattern.flags = pt.flags & (PFH_PATTERN_REPEAT_X | PFH_PATTERN_REPEAT_Y |
                                  PFH_PATTERN_NO_SMOOTHING);
    p.pattern.filter.kind = pt.filter.kind;
    int used = pt.filter.kind == PFH_FILTER_BLUR ? 2 : pt.filter.kind == PFH_FILTER_COLOR_MATRIX
                                                          ? 20 : 0;
    for (int i = 0; i < 20; i++) p.pattern.filter.params[i] = i < used ? pt.filter.params[i] : 0;
    *out = std::move(p);
    return true;
}

std::string paint_key(const PaintDesc &p) {
    std::string k;
    put(k, p.overlay);
    put(k, p.base);
    if (p.overlay == OVERLAY_NONE) return k;
    put(k, p.composite_op);
    if (p.overlay == OVERLAY_GRADIENT) {
        const GradientDesc &g = p.gradient;
        put(k, g.geometry);
        put(k, g.from);
        put(k, g.to);
        if (g.geometry == PFH_GRADIENT_RADIAL) {
            put(k, g.radius0);
            put(k, g.radius1);
            put(k, g.transform);
        }
        put(k, g.wrap);
        for (const pfh_color_stop &s : g.stops) {
            put(k, s.offset);
            put(k, s.color);
        }
    } else {
        const PatternDesc &q = p.pattern;
        put(k, q.width);
        put(k, q.height);
        put(k, q.transform);
        put(k, q.flags);
        put(k, q.filter);
        k.append(reinterpret_cast<const char *>(q.pixels.data()), q.pixels.size() * 4);
    }
    return k;
}

bool paint_is_opaque(const PaintDesc &p) {
    if (p.base.a != 255) return false;
    if (p.overlay == OVERLAY_GRADIENT) {
        for (const pfh_color_stop &s : p.gradient.stops)
            if (s.color.a != 255) return false;
        return true;
    }
    if (p.overlay == OVERLAY_PATTERN) {
        for (const pfh_color_u &c : p.pattern.pixels)
            if (c.a != 255) return false;
    }
    return true;
}

pfh_color_u gradient_sample(const std::vector<pfh_color_stop> &stops, float t) {
    if (stops.empty()) return pfh_color_u{0, 0, 0, 0};
    t = std::fmin(1.0f, std::fmax(0.0f, t));  // geometry util::clamp
    // binary_search_by with Less for (offset < t || offset == 0): the first
    // stop past that prefix
    size_t lo = 0, hi = stops.size();
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (stops[mid].offset < t || stops[mid].offset == 0.0f) lo = mid + 1;
        else hi = mid;
    }
    size_t last = stops.size() - 1;
    size_t upper = std::min(lo, last);
    size_t lower = upper > 0 ? upper - 1 : upper;
    const pfh_color_stop &a = stops[lower], &b = stops[upper];
    float denom = b.offset - a.offset;
    if (denom == 0.0f) return a.color;
    float ratio = std::fmin((t - a.offset) / denom, 1.0f);
    // ColorU::to_f32 (x * 1/255), ColorF::lerp (a + (b - a) t), ColorF::to_u8
    // (round half to even of x * 255, low byte)
    const float inv = 1.0f / 255.0f;
    const uint8_t ca[4] = {a.color.r, a.color.g, a.color.b, a.color.a};
    const uint8_t cb[4] = {b.color.r, b.color.g, b.color.b, b.color.a};
    uint8_t o[4];
    for (int c = 0; c < 4; c++) {
        float fa = (float)ca[c] * inv, fb = (float)cb[c] * inv;
        float v = fa + (fb - fa) * ratio;
        o[c] = (uint8_t)(int32_t)std::nearbyint(v * 255.0f);
    }
    return pfh_color_u{o[0], o[1], o[2], o[3]};
}

void build_palette(const std::vector<PaintDesc> &paints, const pfh_transform2f &build_transform,
                   PaletteBuild *out) {
    *out = PaletteBuild();
    // the render transform is the inverse of the build's 2D transform
    // (builder.rs:168-171)
    Xform render = xinverse(from_c(build_transform));

    struct GradientTile {
        uint32_t page;
        std::vector<pfh_color_u> texels;
        int32_t next_index;
    };
    std::vector<GradientTile> tiles;
    std::vector<TexelUpload> images;
    std::map<std::string, uint32_t> image_pages;  // image and border -> page

    auto new_page = [&](int32_t w, int32_t h) {
        out->pages.push_back(PageSize{w, h});
        return (uint32_t)(out->pages.size() - 1);
    };

    for (const PaintDesc &p : paints) {
        pfh_texture_metadata_entry e;
        std::memset(&e, 0, sizeof(e));
        e.color_0_transform = pfh_transform2f{1, 0, 0, 1, 0, 0};
        e.base_color = p.base;
        e.blend_mode = PFH_BLEND_MODE_SRC_OVER;  // PaintMetadata.blend_mode (paint.rs:573)
        TileBatchTexture tex;
        if (p.overlay == OVERLAY_GRADIENT) {
            const GradientDesc &g = p.gradient;
            // GradientTileBuilder::allocate: a row of the last 256x256 page
            if (tiles.empty() || tiles.back().next_index == kGradientTileLength) {
                uint32_t page = new_page(kGradientTileLength, kGradientTileLength);
                tiles.push_back(GradientTile{
                    page,
                    std::vector<pfh_color_u>((size_t)kGradientTileLength * kGradientTileLength,
                                             pfh_color_u{0, 0, 0, 255}),
                    0});
            }
            GradientTile &tile = tiles.back();
            int32_t row = tile.next_index++;
            for (int32_t x = 0; x < kGradientTileLength; x++) {
                float t = ((float)x + 0.5f) / (float)kGradientTileLength;
                tile.texels[(size_t)row * kGradientTileLength + x] = gradient_sample(g.stops, t);
            }
            const float scale = 1.0f / (float)kGradientTileLength;  // page_scale
            Xform xf;
            if (g.geometry == PFH_GRADIENT_LINEAR) {
                // project the gradient line onto (0..1, v0), v0 the row centre
                float v0 = ((float)row + 1.0f * 0.5f) * scale;
                float dx = g.to.x - g.from.x, dy = g.to.y - g.from.y;
                float len2 = dx * dx + dy * dy;
                float m0x = dx / len2, m0y = dy / len2;
                float m13x = m0x * -g.from.x, m13y = m0y * -g.from.y;
                // Transform2F::row_major(m0.x, m0.y, m13.x + m13.y, 0, 0, v0)
                xf = Xform{{m0x, 0.0f, m0y, 0.0f}, m13x + m13y, v0};
            } else {
                xf = xinverse(from_c(g.transform));
                // Filter::RadialGradient: line, radii, uv origin = the row's
                // uv rect contracted by half a texel vertically
                e.filter.kind = PFH_FILTER_RADIAL_GRADIENT;
                e.filter.params[0] = g.from.x;
                e.filter.params[1] = g.from.y;
                e.filter.params[2] = g.to.x;
                e.filter.params[3] = g.to.y;
                e.filter.params[4] = g.radius0;
                e.filter.params[5] = g.radius1;
                e.filter.params[6] = 0.0f * scale + 0.0f;
                e.filter.params[7] = (float)row * scale + scale * 0.5f;
            }
            e.color_0_transform = to_c(xmul(xf, render));
            e.color_0_combine_mode = PFH_COMBINE_SRC_IN;
            tex.page = (int32_t)tile.page;
            tex.sampling_flags = g.wrap == PFH_GRADIENT_WRAP_REPEAT ? PFH_SAMPLING_REPEAT_U : 0u;
            tex.composite_op = p.composite_op;
        } else if (p.overlay == OVERLAY_PATTERN) {
            const PatternDesc &q = p.pattern;
            int32_t bx = (q.flags & PFH_PATTERN_REPEAT_X) ? 0 : 1;
            int32_t by = (q.flags & PFH_PATTERN_REPEAT_Y) ? 0 : 1;
            int32_t pw = q.width + 2 * bx, ph = q.height + 2 * by;
            // one page per image (AllocationMode::OwnPage), shared by the
            // paints that show the same image with the same border
            std::string key;
            put(key, q.width);
            put(key, q.height);
            put(key, bx);
            put(key, by);
            key.append(reinterpret_cast<const char *>(q.pixels.data()), q.pixels.size() * 4);
            auto it = image_pages.find(key);
            uint32_t page;
            if (it == image_pages.end()) {
                page = new_page(pw, ph);
                image_pages.emplace(std::move(key), page);
            } else {
                page = it->second;
            }
            images.push_back(TexelUpload{page, pfh_rect_i{bx, by, bx + q.width, by + q.height},
                                         q.pixels});
            // texture_scale * (uv origin of the page rect, (0, 0)) after the
            // pattern's inverse transform
            float sx = 1.0f / (float)pw, sy = 1.0f / (float)ph;
            Xform xf = xmul(scale_then_translate(sx, sy, 0.0f * sx, 0.0f * sy),
                            xinverse(from_c(q.transform)));
            e.color_0_transform = to_c(xmul(xf, render));
            e.color_0_combine_mode = PFH_COMBINE_SRC_IN;
            e.filter = q.filter;
            tex.page = (int32_t)page;
            tex.sampling_flags = ((q.flags & PFH_PATTERN_REPEAT_X) ? PFH_SAMPLING_REPEAT_U : 0u) |
                                 ((q.flags & PFH_PATTERN_REPEAT_Y) ? PFH_SAMPLING_REPEAT_V : 0u) |
                                 ((q.flags & PFH_PATTERN_NO_SMOOTHING)
                                      ? (PFH_SAMPLING_NEAREST_MIN | PFH_SAMPLING_NEAREST_MAG)
                                      : 0u);
            tex.composite_op = p.composite_op;
        }
        out->metadata.push_back(e);
        out->opaque.push_back(paint_is_opaque(p) ? 1 : 0);
        out->textures.push_back(tex);
    }
    // create_render_commands: gradient tiles (whole pages), then images
    for (GradientTile &t : tiles)
        out->uploads.push_back(TexelUpload{
            t.page, pfh_rect_i{0, 0, kGradientTileLength, kGradientTileLength}, std::move(t.texels)});
    for (TexelUpload &u : images) out->uploads.push_back(std::move(u));
}

}  // namespace pfh
