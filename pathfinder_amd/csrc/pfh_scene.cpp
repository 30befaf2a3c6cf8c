// pfh_scene.cpp - Scene / SceneBuilder front-end for the D3D11-level path.
//
// Mirrors, for 2D transforms:
//   Scene::{push_paint, push_draw_path, push_clip_path, set_view_box}
//       renderer/src/scene.rs:77-226, palette dedup renderer/src/paint.rs:381-390
//   Palette::build_paint_info (gradients, patterns)   paint.rs:398-437 (pfh_palette.cpp)
//   SceneBuilder::build                    renderer/src/builder.rs:148-222
//   BuiltSegments::from_scene / SegmentsD3D11::add_path   builder.rs:777-840
//   build_tile_batches_for_draw_path_display_item         builder.rs:885-1056
//   prepare_draw_path_for_gpu_binning                     builder.rs:1058-1095
//   add_clip_path_to_batch / prepare_clip_path_...        builder.rs:1124-1220
//   TileBatchDataD3D11::push, init_backdrops              builder.rs:653-768
//   BuiltPath::new tile bounds                            builder.rs:415-485
//   round_rect_out_to_tile_bounds                         renderer/src/tiles.rs:64-66
// and emits the same RenderCommand stream (gpu_data.rs:37-105).
#include "pfh_internal.h"
#include "pfh_palette.h"
#include "pfh_d3d9.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace {

struct RectF {
    float x0 = 0, y0 = 0, x1 = 0, y1 = 0;
};

struct PathRecord {
    uint32_t point_begin = 0, point_end = 0;      // into the source's point arrays
    uint32_t contour_begin = 0, contour_end = 0;  // into the source's contour_ends
    RectF bounds;                                  // Outline::bounds
    uint32_t paint = 0;
    int32_t fill_rule = PFH_FILL_RULE_WINDING;
    int32_t clip_path = -1;
    int32_t blend_mode = PFH_BLEND_MODE_SRC_OVER;
};

struct PathSet {
    std::vector<pfh_vector2f> points;
    std::vector<uint8_t> flags;
    std::vector<uint32_t> contour_ends;  // absolute indices into points
    std::vector<PathRecord> paths;
};

// union_rect over every point of every non-empty contour (outline.rs:180-192, 560-568, 1086-1091).
RectF outline_bounds(const pfh_vector2f *pts, const uint32_t *ends, uint32_t begin_point,
                     uint32_t contour_begin, uint32_t contour_end) {
    RectF r;
    bool first_contour = true;
    uint32_t start = begin_point;
    for (uint32_t c = contour_begin; c < contour_end; c++) {
        uint32_t end = ends[c];
        if (end > start) {
            RectF cb{pts[start].x, pts[start].y, pts[start].x, pts[start].y};
            for (uint32_t i = start + 1; i < end; i++) {
                cb.x0 = std::min(cb.x0, pts[i].x);
                cb.y0 = std::min(cb.y0, pts[i].y);
                cb.x1 = std::max(cb.x1, pts[i].x);
                cb.y1 = std::max(cb.y1, pts[i].y);
            }
            if (first_contour) {
                r = cb;
                first_contour = false;
            } else {
                r.x0 = std::min(r.x0, cb.x0);
                r.y0 = std::min(r.y0, cb.y0);
                r.x1 = std::max(r.x1, cb.x1);
                r.y1 = std::max(r.y1, cb.y1);
            }
        }
        start = end;
    }
    return r;
}

// Transform2F * Vector2F (geometry/src/transform2d.rs:123-129, 312-318).
inline void xform_point(const pfh_transform2f &t, float x, float y, float *ox, float *oy) {
    *ox = (t.m00 * x + t.m01 * y) + t.tx;
    *oy = (t.m10 * x + t.m11 * y) + t.ty;
}

// Transform2F * RectF (transform2d.rs:328-338).
RectF xform_rect(const pfh_transform2f &t, const RectF &r) {
    float ux0, uy0, ux1, uy1, lx0, ly0, lx1, ly1;
    xform_point(t, r.x0, r.y0, &ux0, &uy0);  // upper left
    xform_point(t, r.x1, r.y0, &ux1, &uy1);  // upper right
    xform_point(t, r.x0, r.y1, &lx0, &ly0);  // lower left
    xform_point(t, r.x1, r.y1, &lx1, &ly1);  // lower right
    RectF o;
    o.x0 = std::min(std::min(std::min(ux0, ux1), lx0), lx1);
    o.y0 = std::min(std::min(std::min(uy0, uy1), ly0), ly1);
    o.x1 = std::max(std::max(std::max(ux0, ux1), lx0), lx1);
    o.y1 = std::max(std::max(std::max(uy0, uy1), ly0), ly1);
    return o;
}

// RectF::intersection (geometry/src/rect.rs:122-137).
bool intersect(const RectF &a, const RectF &b, RectF *out) {
    if (!(a.x0 < b.x1 && a.y0 < b.y1 && b.x0 < a.x1 && b.y0 < a.y1)) return false;
    out->x0 = std::max(a.x0, b.x0);
    out->y0 = std::max(a.y0, b.y0);
    out->x1 = std::min(a.x1, b.x1);
    out->y1 = std::min(a.y1, b.y1);
    return true;
}

struct RectI {
    int32_t x0, y0, x1, y1;
    int32_t width() const { return x1 - x0; }
    int32_t height() const { return y1 - y0; }
};

// BlendMode::is_destructive (effects.rs:234-262).
bool blend_is_destructive(int32_t mode) {
    return mode == PFH_BLEND_MODE_CLEAR || mode == PFH_BLEND_MODE_COPY ||
           mode == PFH_BLEND_MODE_SRC_IN || mode == PFH_BLEND_MODE_DEST_IN ||
           mode == PFH_BLEND_MODE_SRC_OUT || mode == PFH_BLEND_MODE_DEST_ATOP;
}

// BlendMode::occludes_backdrop (effects.rs:202-233).
bool blend_occludes_backdrop(int32_t mode) {
    return mode == PFH_BLEND_MODE_SRC_OVER || mode == PFH_BLEND_MODE_CLEAR;
}

// round_rect_out_to_tile_bounds (tiles.rs:64-66): (rect * 1/16).round_out().to_i32().
RectI tile_bounds(const RectF &r) {
    const float s = 1.0f / 16.0f;
    RectI o;
    o.x0 = (int32_t)std::floor(r.x0 * s);
    o.y0 = (int32_t)std::floor(r.y0 * s);
    o.x1 = (int32_t)std::ceil(r.x1 * s);
    o.y1 = (int32_t)std::ceil(r.y1 * s);
    return o;
}

}  // namespace

struct pfh_scene {
    RectF view_box;
    std::vector<pfh_color_u> paints;         // base colours, by PaintId
    std::vector<pfh::PaintDesc> paint_descs;  // the paints, by PaintId
    std::map<std::string, uint32_t> paint_cache;  // pfh::paint_key -> PaintId
    PathSet draw, clip;
    // pfh_scene_path_arrays: the per-path columns of the last export
    std::vector<uint32_t> ex_path_contour_ends, ex_paints;
    std::vector<int32_t> ex_fill_rules, ex_clip_paths, ex_blend_modes;
};

struct pfh_command {
    int32_t kind;
    uint32_t batch_slot;  // index into batches, pages or uploads
};

struct pfh_command_list {
    std::vector<pfh_command> commands;
    std::vector<pfh_color_u> paints;
    std::vector<pfh_texture_metadata_entry> metadata;
    std::vector<pfh::PageSize> pages;
    std::vector<pfh::TexelUpload> uploads;
    std::vector<pfh_vector2f> draw_points, clip_points;
    std::vector<pfh_segment_indices> draw_indices, clip_indices;
    std::vector<pfh::BatchData> batches;
    uint32_t path_count = 0;
    // the D3D9 level: fills per path (ADD_FILLS_D3D9 slots) and the batch
    std::vector<std::vector<pfh_fill_d3d9>> fills_d3d9;
    pfh9::Built batch_d3d9;
};

namespace {

uint32_t push_paths(PathSet &set, const pfh_vector2f *points, const uint8_t *flags,
                    uint32_t point_count, const uint32_t *contour_ends, uint32_t contour_count,
                    const uint32_t *path_contour_ends, uint32_t path_count, const uint32_t *paints,
                    const int32_t *fill_rules, const int32_t *clip_paths,
                    const int32_t *blend_modes) {
    uint32_t first_id = (uint32_t)set.paths.size();
    uint32_t base_point = (uint32_t)set.points.size();
    uint32_t base_contour = (uint32_t)set.contour_ends.size();
    set.points.insert(set.points.end(), points, points + point_count);
    if (flags) {
        set.flags.insert(set.flags.end(), flags, flags + point_count);
    } else {
        set.flags.insert(set.flags.end(), point_count, 0);
    }
    for (uint32_t c = 0; c < contour_count; c++)
        set.contour_ends.push_back(base_point + std::min(contour_ends[c], point_count));
    uint32_t prev_contour = 0;
    for (uint32_t i = 0; i < path_count; i++) {
        uint32_t ce = std::min(path_contour_ends[i], contour_count);
        if (ce < prev_contour) ce = prev_contour;
        PathRecord rec;
        rec.contour_begin = base_contour + prev_contour;
        rec.contour_end = base_contour + ce;
        rec.point_begin = prev_contour == 0 ? base_point : set.contour_ends[rec.contour_begin - 1];
        rec.point_end = ce == 0 ? base_point : set.contour_ends[rec.contour_end - 1];
        if (rec.point_end < rec.point_begin) rec.point_end = rec.point_begin;
        rec.bounds = outline_bounds(set.points.data(), set.contour_ends.data(), rec.point_begin,
                                    rec.contour_begin, rec.contour_end);
        rec.paint = paints ? paints[i] : 0;
        rec.fill_rule = fill_rules ? fill_rules[i] : PFH_FILL_RULE_WINDING;
        rec.clip_path = clip_paths ? clip_paths[i] : -1;
        rec.blend_mode = blend_modes ? blend_modes[i] : PFH_BLEND_MODE_SRC_OVER;
        set.paths.push_back(rec);
        prev_contour = ce;
    }
    return first_id;
}

// SegmentsD3D11::add_path (builder.rs:804-840): every on-curve point starts a
// segment whose kind comes from the flags of the next two points, and each
// contour gets its first point appended so it closes.
void add_segments(const PathSet &set, const PathRecord &rec, std::vector<pfh_vector2f> &points,
                  std::vector<pfh_segment_indices> &indices, uint32_t *range_begin,
                  uint32_t *range_end) {
    *range_begin = (uint32_t)indices.size();
    uint32_t start = rec.point_begin;
    for (uint32_t c = rec.contour_begin; c < rec.contour_end; c++) {
        uint32_t end = set.contour_ends[c];
        uint32_t n = end > start ? end - start : 0;
        if (n > 0) {
            for (uint32_t i = 0; i < n; i++) {
                uint8_t f = set.flags[start + i];
                if (!(f & (PFH_POINT_CONTROL_0 | PFH_POINT_CONTROL_1))) {
                    uint32_t flags = 0;
                    if (i + 1 < n && (set.flags[start + i + 1] & PFH_POINT_CONTROL_0)) {
                        if (i + 2 < n && (set.flags[start + i + 2] & PFH_POINT_CONTROL_1))
                            flags = PFH_CURVE_IS_CUBIC;
                        else
                            flags = PFH_CURVE_IS_QUADRATIC;
                    }
                    indices.push_back(pfh_segment_indices{(uint32_t)points.size(), flags});
                }
                points.push_back(set.points[start + i]);
            }
            points.push_back(set.points[start]);
        }
        start = std::max(start, end);
    }
    *range_end = (uint32_t)indices.size();
}

struct BuiltPathGPU {
    RectI tile_bounds;
    uint32_t paint;
    uint8_t ctrl;
};

// TileBatchDataD3D11::new (builder.rs:620-651).
pfh::BatchData new_batch(uint32_t batch_id, uint32_t path_source, const pfh_transform2f &t) {
    pfh::BatchData b;
    b.batch_id = batch_id;
    b.path_source = path_source;
    b.transform = t;
    return b;
}

// TileBatchDataD3D11::push (builder.rs:653-760) + init_backdrops (:762-768).
uint32_t batch_push(pfh::BatchData &b, const BuiltPathGPU &path, uint32_t global_path_id,
                    uint32_t segment_begin, uint32_t segment_end, bool has_clip,
                    uint32_t clip_batch_id, uint32_t clip_path_index, bool z_write) {
    uint32_t path_index = b.path_count++;
    pfh_propagate_metadata pm;
    std::memset(&pm, 0, sizeof(pm));
    pm.tile_rect[0] = path.tile_bounds.x0;
    pm.tile_rect[1] = path.tile_bounds.y0;
    pm.tile_rect[2] = path.tile_bounds.x1;
    pm.tile_rect[3] = path.tile_bounds.y1;
    pm.tile_offset = b.tile_count;
    pm.path_index = path_index;
    pm.z_write = z_write ? 1u : 0u;
    pm.clip_path_index = has_clip ? clip_path_index : 0xffffffffu;
    pm.backdrop_offset = (uint32_t)b.backdrops.size();
    b.propagate_metadata.push_back(pm);
    for (int32_t x = 0; x < path.tile_bounds.width(); x++)
        b.backdrops.push_back(pfh_backdrop_info{0, x, path_index});
    b.dice_metadata.push_back(
        pfh_dice_metadata{global_path_id, segment_begin, b.segment_count, 0});
    pfh_tile_path_info tpi;
    tpi.tile_min_x = (int16_t)path.tile_bounds.x0;
    tpi.tile_min_y = (int16_t)path.tile_bounds.y0;
    tpi.tile_max_x = (int16_t)path.tile_bounds.x1;
    tpi.tile_max_y = (int16_t)path.tile_bounds.y1;
    tpi.first_tile_index = b.tile_count;
    tpi.color = (uint16_t)path.paint;
    tpi.ctrl = path.ctrl;
    tpi.backdrop = 0;
    b.tile_path_info.push_back(tpi);
    b.tile_count += (uint32_t)(path.tile_bounds.width() * path.tile_bounds.height());
    b.segment_count += segment_end - segment_begin;
    if (has_clip && b.clip_batch_id < 0) b.clip_batch_id = (int32_t)clip_batch_id;
    return path_index;
}

struct ClipBatches {
    std::vector<pfh::BatchData> prepare_batches;  // submitted in reverse (LIFO) order
    std::map<uint32_t, uint32_t> clip_id_to_path_batch_index;
};

struct BuildState {
    const pfh_scene *scene;
    pfh_transform2f transform;
    RectF view_box;
    std::vector<uint32_t> clip_seg_begin, clip_seg_end;
    ClipBatches clips;
};

// add_clip_path_to_batch (builder.rs:1124-1178) + prepare_clip_path_for_gpu_binning (:1181-1215).
bool add_clip_path_to_batch(BuildState &st, int32_t clip_path_id, uint32_t clip_level,
                            uint32_t *out_batch_id, uint32_t *out_path_index) {
    if (clip_path_id < 0 || (size_t)clip_path_id >= st.scene->clip.paths.size()) return false;
    // Clip nesting deeper than the clip batch id space (MAX_CLIP_BATCHES = 32,
    // builder.rs:50) cannot be expressed; a clip cycle would recurse forever.
    if (clip_level >= 32) return false;
    auto found = st.clips.clip_id_to_path_batch_index.find((uint32_t)clip_path_id);
    if (found != st.clips.clip_id_to_path_batch_index.end()) {
        *out_batch_id = clip_level;
        *out_path_index = found->second;
        return true;
    }
    const PathRecord &cp = st.scene->clip.paths[(size_t)clip_path_id];
    uint32_t sub_batch = 0, sub_index = 0;
    bool has_sub = add_clip_path_to_batch(st, cp.clip_path, clip_level + 1, &sub_batch, &sub_index);
    // The clip path's bounds are not intersected with the view box (builder.rs:1207).
    RectF bounds = xform_rect(st.transform, cp.bounds);
    BuiltPathGPU built{tile_bounds(bounds), 0, 0};
    while (clip_level >= st.clips.prepare_batches.size()) {
        uint32_t id = (uint32_t)st.clips.prepare_batches.size();
        st.clips.prepare_batches.push_back(new_batch(id, PFH_PATH_SOURCE_CLIP, st.transform));
    }
    uint32_t index = batch_push(st.clips.prepare_batches[clip_level], built, (uint32_t)clip_path_id,
                                st.clip_seg_begin[(size_t)clip_path_id],
                                st.clip_seg_end[(size_t)clip_path_id], has_sub, sub_batch,
                                sub_index, true);
    st.clips.clip_id_to_path_batch_index[(uint32_t)clip_path_id] = index;
    *out_batch_id = clip_level;
    *out_path_index = index;
    return true;
}

}  // namespace

extern "C" {

pfh_scene *pfh_scene_create(void) { return new (std::nothrow) pfh_scene(); }

void pfh_scene_destroy(pfh_scene *scene) { delete scene; }

void pfh_scene_set_view_box(pfh_scene *scene, const pfh_rect_f *vb) {
    if (!scene || !vb) return;
    scene->view_box = RectF{vb->origin.x, vb->origin.y, vb->lower_right.x, vb->lower_right.y};
}

namespace {
// Palette::push_paint (paint.rs:381-390): the id of an equal paint, else the next id.
uint32_t push_paint_desc(pfh_scene *scene, pfh::PaintDesc &&desc) {
    std::string key = pfh::paint_key(desc);
    auto it = scene->paint_cache.find(key);
    if (it != scene->paint_cache.end()) return it->second;
    uint32_t id = (uint32_t)scene->paints.size();
    scene->paints.push_back(desc.base);
    scene->paint_descs.push_back(std::move(desc));
    scene->paint_cache.emplace(std::move(key), id);
    return id;
}
}  // namespace

uint32_t pfh_scene_push_paint(pfh_scene *scene, const pfh_color_u *color) {
    if (!scene || !color) return 0;
    pfh::PaintDesc d;
    d.base = *color;
    return push_paint_desc(scene, std::move(d));
}

pfh_status pfh_scene_set_draw_points(pfh_scene *scene, const pfh_vector2f *points,
                                     uint32_t count) {
    if (!scene || (count && !points)) return PFH_ERROR_INVALID_ARGUMENT;
    PathSet &set = scene->draw;
    if (count != set.points.size()) return PFH_ERROR_INVALID_ARGUMENT;
    std::copy(points, points + count, set.points.begin());
    for (PathRecord &r : set.paths)
        r.bounds = outline_bounds(set.points.data(), set.contour_ends.data(), r.point_begin,
                                  r.contour_begin, r.contour_end);
    return PFH_OK;
}

pfh_status pfh_scene_set_paint_colors(pfh_scene *scene, const pfh_color_u *colors,
                                      uint32_t count) {
    if (!scene || (count && !colors)) return PFH_ERROR_INVALID_ARGUMENT;
    if (count > 65536) return PFH_ERROR_LIMIT;
    scene->paints.assign(colors, colors + count);
    scene->paint_descs.assign(count, pfh::PaintDesc());
    scene->paint_cache.clear();
    for (uint32_t i = 0; i < count; i++) {
        scene->paint_descs[i].base = colors[i];
        scene->paint_cache.emplace(pfh::paint_key(scene->paint_descs[i]), i);  // first id wins
    }
    return PFH_OK;
}

uint32_t pfh_scene_push_gradient_paint(pfh_scene *scene, const pfh_color_u *base_color,
                                       const pfh_gradient *gradient, int32_t composite_op) {
    pfh::PaintDesc d;
    if (!scene || !base_color || !gradient ||
        !pfh::paint_from_gradient(*base_color, *gradient, composite_op, &d) ||
        scene->paints.size() >= 65536)
        return 0xffffffffu;
    return push_paint_desc(scene, std::move(d));
}

uint32_t pfh_scene_push_pattern_paint(pfh_scene *scene, const pfh_color_u *base_color,
                                      const pfh_pattern *pattern, int32_t composite_op) {
    pfh::PaintDesc d;
    if (!scene || !base_color || !pattern ||
        !pfh::paint_from_pattern(*base_color, *pattern, composite_op, &d) ||
        scene->paints.size() >= 65536)
        return 0xffffffffu;
    return push_paint_desc(scene, std::move(d));
}

uint32_t pfh_scene_push_clip_path(pfh_scene *scene, const pfh_outline *o, int32_t fill_rule,
                                  int32_t clip_path) {
    if (!scene || !o) return 0xffffffffu;
    uint32_t ends = o->contour_count;
    return push_paths(scene->clip, o->points, o->flags, o->point_count, o->contour_ends,
                      o->contour_count, &ends, 1, nullptr, &fill_rule, &clip_path, nullptr);
}

uint32_t pfh_scene_push_draw_path(pfh_scene *scene, const pfh_outline *o, uint32_t paint,
                                  int32_t fill_rule, int32_t clip_path, int32_t blend_mode) {
    if (!scene || !o) return 0xffffffffu;
    uint32_t ends = o->contour_count;
    return push_paths(scene->draw, o->points, o->flags, o->point_count, o->contour_ends,
                      o->contour_count, &ends, 1, &paint, &fill_rule, &clip_path, &blend_mode);
}

uint32_t pfh_scene_push_draw_paths(pfh_scene *scene, const pfh_vector2f *points,
                                   const uint8_t *flags, uint32_t point_count,
                                   const uint32_t *contour_ends, uint32_t contour_count,
                                   const uint32_t *path_contour_ends, uint32_t path_count,
                                   const uint32_t *paints, const int32_t *fill_rules,
                                   const int32_t *clip_paths, const int32_t *blend_modes) {
    if (!scene || (point_count && !points) || (contour_count && !contour_ends) ||
        (path_count && !path_contour_ends))
        return 0xffffffffu;
    return push_paths(scene->draw, points, flags, point_count, contour_ends, contour_count,
                      path_contour_ends, path_count, paints, fill_rules, clip_paths, blend_modes);
}

uint32_t pfh_scene_push_clip_paths(pfh_scene *scene, const pfh_vector2f *points,
                                   const uint8_t *flags, uint32_t point_count,
                                   const uint32_t *contour_ends, uint32_t contour_count,
                                   const uint32_t *path_contour_ends, uint32_t path_count,
                                   const int32_t *fill_rules, const int32_t *clip_paths) {
    if (!scene || (point_count && !points) || (contour_count && !contour_ends) ||
        (path_count && !path_contour_ends))
        return 0xffffffffu;
    return push_paths(scene->clip, points, flags, point_count, contour_ends, contour_count,
                      path_contour_ends, path_count, nullptr, fill_rules, clip_paths, nullptr);
}

pfh_status pfh_scene_path_arrays(pfh_scene *scene, uint32_t path_source, pfh_path_arrays *out) {
    if (!scene || !out || path_source > PFH_PATH_SOURCE_CLIP) return PFH_ERROR_INVALID_ARGUMENT;
    const PathSet &set = path_source == PFH_PATH_SOURCE_CLIP ? scene->clip : scene->draw;
    size_t n = set.paths.size();
    scene->ex_path_contour_ends.resize(n);
    scene->ex_paints.resize(n);
    scene->ex_fill_rules.resize(n);
    scene->ex_clip_paths.resize(n);
    scene->ex_blend_modes.resize(n);
    for (size_t i = 0; i < n; i++) {
        const PathRecord &r = set.paths[i];
        scene->ex_path_contour_ends[i] = r.contour_end;
        scene->ex_paints[i] = r.paint;
        scene->ex_fill_rules[i] = r.fill_rule;
        scene->ex_clip_paths[i] = r.clip_path;
        scene->ex_blend_modes[i] = r.blend_mode;
    }
    out->points = set.points.data();
    out->flags = set.flags.data();
    out->point_count = (uint32_t)set.points.size();
    out->contour_ends = set.contour_ends.data();
    out->contour_count = (uint32_t)set.contour_ends.size();
    out->path_contour_ends = scene->ex_path_contour_ends.data();
    out->path_count = (uint32_t)n;
    out->paints = scene->ex_paints.data();
    out->fill_rules = scene->ex_fill_rules.data();
    out->clip_paths = scene->ex_clip_paths.data();
    out->blend_modes = scene->ex_blend_modes.data();
    return PFH_OK;
}

pfh_status pfh_scene_paints(const pfh_scene *scene, const pfh_color_u **colors, uint32_t *count) {
    if (!scene || !colors || !count) return PFH_ERROR_INVALID_ARGUMENT;
    *colors = scene->paints.data();
    *count = (uint32_t)scene->paints.size();
    return PFH_OK;
}

pfh_status pfh_scene_get_view_box(const pfh_scene *scene, pfh_rect_f *out) {
    if (!scene || !out) return PFH_ERROR_INVALID_ARGUMENT;
    *out = pfh_rect_f{{scene->view_box.x0, scene->view_box.y0},
                      {scene->view_box.x1, scene->view_box.y1}};
    return PFH_OK;
}

uint32_t pfh_scene_draw_path_count(const pfh_scene *scene) {
    return scene ? (uint32_t)scene->draw.paths.size() : 0;
}

pfh_status pfh_scene_build(pfh_scene *scene, const pfh_build_options *options,
                           pfh_command_list **out) {
    if (!scene || !out) return PFH_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    std::unique_ptr<pfh_command_list> list(new (std::nothrow) pfh_command_list());
    if (!list) return PFH_ERROR_OUT_OF_MEMORY;
    pfh_transform2f identity{1, 0, 0, 1, 0, 0};
    BuildState st;
    st.scene = scene;
    st.transform = options ? options->transform : identity;
    st.view_box = scene->view_box;  // effective_view_box without subpixel AA (scene.rs:276-282)

    list->path_count = (uint32_t)(scene->clip.paths.size() + scene->draw.paths.size());
    list->commands.push_back({PFH_COMMAND_START, 0});
    // Palette::build_paint_info (paint.rs:398-437): metadata, then a page
    // for every new texture page, then the texel uploads
    pfh::PaletteBuild palette;
    pfh::build_palette(scene->paint_descs, st.transform, &palette);
    list->paints = scene->paints;
    list->metadata = palette.metadata;
    list->commands.push_back({PFH_COMMAND_UPLOAD_TEXTURE_METADATA, 0});
    list->pages = palette.pages;
    for (size_t i = 0; i < list->pages.size(); i++)
        list->commands.push_back({PFH_COMMAND_ALLOCATE_TEXTURE_PAGE, (uint32_t)i});
    list->uploads = std::move(palette.uploads);
    for (size_t i = 0; i < list->uploads.size(); i++)
        list->commands.push_back({PFH_COMMAND_UPLOAD_TEXEL_DATA, (uint32_t)i});

    // BuiltSegments::from_scene: clip paths first, then draw paths (builder.rs:782-797).
    st.clip_seg_begin.resize(scene->clip.paths.size());
    st.clip_seg_end.resize(scene->clip.paths.size());
    for (size_t i = 0; i < scene->clip.paths.size(); i++)
        add_segments(scene->clip, scene->clip.paths[i], list->clip_points, list->clip_indices,
                     &st.clip_seg_begin[i], &st.clip_seg_end[i]);
    std::vector<uint32_t> draw_seg_begin(scene->draw.paths.size()),
        draw_seg_end(scene->draw.paths.size());
    for (size_t i = 0; i < scene->draw.paths.size(); i++)
        add_segments(scene->draw, scene->draw.paths[i], list->draw_points, list->draw_indices,
                     &draw_seg_begin[i], &draw_seg_end[i]);
    list->commands.push_back({PFH_COMMAND_UPLOAD_SCENE_D3D11, 0});

    // One DrawPaths display item holds every draw path (scene.rs:84-94). A
    // path whose paint has a colour texture other than the batch's starts a
    // new batch (fixup_batch_for_new_path_if_possible, builder.rs:914-942,
    // 1227-1243); a batch without one adopts the first path's.
    const uint32_t MAX_CLIP_BATCHES = 32;
    uint32_t next_batch_id = MAX_CLIP_BATCHES;
    bool have_draw_batch = false;
    pfh::BatchData draw_batch;
    pfh::TileBatchTexture batch_texture;
    std::vector<pfh::BatchData> draw_batches;
    for (size_t i = 0; i < scene->draw.paths.size(); i++) {
        const PathRecord &dp = scene->draw.paths[i];
        RectF bounds = xform_rect(st.transform, dp.bounds);
        RectF clipped;
        if (!intersect(bounds, st.view_box, &clipped)) continue;
        // BuiltPath::new (builder.rs:430-436): a destructive blend mode tiles
        // the whole view box, anything else its clipped bounds
        uint8_t ctrl = dp.fill_rule == PFH_FILL_RULE_EVEN_ODD ? 0x2 : 0x1;
        BuiltPathGPU built{tile_bounds(blend_is_destructive(dp.blend_mode) ? st.view_box : clipped),
                           dp.paint, ctrl};
        pfh::TileBatchTexture path_texture;
        if (dp.paint < palette.textures.size()) path_texture = palette.textures[dp.paint];
        if (have_draw_batch && path_texture.page >= 0) {
            if (batch_texture.page < 0) {
                batch_texture = path_texture;
                draw_batch.color_texture_page = path_texture.page;
                draw_batch.color_texture_sampling_flags = path_texture.sampling_flags;
            } else if (path_texture != batch_texture) {
                draw_batches.push_back(std::move(draw_batch));
                have_draw_batch = false;
            }
        }
        if (!have_draw_batch) {
            draw_batch = new_batch(next_batch_id++, PFH_PATH_SOURCE_DRAW, st.transform);
            batch_texture = path_texture;
            draw_batch.color_texture_page = path_texture.page;
            draw_batch.color_texture_sampling_flags = path_texture.sampling_flags;
            have_draw_batch = true;
        }
        uint32_t cb = 0, ci = 0;
        bool has_clip = add_clip_path_to_batch(st, dp.clip_path, 0, &cb, &ci);
        // BuiltDrawPath::new (builder.rs:83): occludes = opaque paint and a
        // mode that occludes the backdrop
        bool opaque = dp.paint < palette.opaque.size() && palette.opaque[dp.paint] &&
                      blend_occludes_backdrop(dp.blend_mode);
        batch_push(draw_batch, built, (uint32_t)i, draw_seg_begin[i], draw_seg_end[i], has_clip,
                   cb, ci, opaque);
    }

    // send_to (builder.rs:1097-1113): clip batches LIFO, then draws.
    for (size_t k = st.clips.prepare_batches.size(); k-- > 0;) {
        if (st.clips.prepare_batches[k].path_count == 0) continue;
        list->batches.push_back(std::move(st.clips.prepare_batches[k]));
        list->commands.push_back(
            {PFH_COMMAND_PREPARE_CLIP_TILES_D3D11, (uint32_t)list->batches.size() - 1});
    }
    if (have_draw_batch) draw_batches.push_back(std::move(draw_batch));
    for (pfh::BatchData &b : draw_batches) {
        list->batches.push_back(std::move(b));
        list->commands.push_back({PFH_COMMAND_DRAW_TILES_D3D11, (uint32_t)list->batches.size() - 1});
    }
    list->commands.push_back({PFH_COMMAND_FINISH, 0});
    *out = list.release();
    return PFH_OK;
}

pfh_status pfh_scene_build_d3d9(pfh_scene *scene, const pfh_build_options *options,
                                pfh_command_list **out) {
    if (!scene || !out) return PFH_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    for (const pfh::PaintDesc &d : scene->paint_descs)
        if (d.overlay != pfh::OVERLAY_NONE) return PFH_ERROR_INVALID_ARGUMENT;
    for (const PathRecord &r : scene->draw.paths)
        if (r.blend_mode != PFH_BLEND_MODE_SRC_OVER) return PFH_ERROR_INVALID_ARGUMENT;
    std::unique_ptr<pfh_command_list> list(new (std::nothrow) pfh_command_list());
    if (!list) return PFH_ERROR_OUT_OF_MEMORY;
    pfh_transform2f transform = options ? options->transform : pfh_transform2f{1, 0, 0, 1, 0, 0};
    list->path_count = (uint32_t)(scene->clip.paths.size() + scene->draw.paths.size());
    list->commands.push_back({PFH_COMMAND_START, 0});
    pfh::PaletteBuild palette;
    pfh::build_palette(scene->paint_descs, transform, &palette);
    list->paints = scene->paints;
    list->metadata = palette.metadata;
    list->commands.push_back({PFH_COMMAND_UPLOAD_TEXTURE_METADATA, 0});
    // the outlines packed as SegmentsD3D11 are the segments the tiler walks
    // (each contour closed back to its first point)
    std::vector<pfh9::PathInput> clips(scene->clip.paths.size()), draws(scene->draw.paths.size());
    for (size_t i = 0; i < clips.size(); i++) {
        add_segments(scene->clip, scene->clip.paths[i], list->clip_points, list->clip_indices,
                     &clips[i].seg_begin, &clips[i].seg_end);
        clips[i].fill_rule = scene->clip.paths[i].fill_rule;
        clips[i].clip_path = -1;  // a clip path's clip is not applied at this level
        clips[i].paint = 0;
        clips[i].occludes = false;
    }
    for (size_t i = 0; i < draws.size(); i++) {
        const PathRecord &r = scene->draw.paths[i];
        add_segments(scene->draw, r, list->draw_points, list->draw_indices, &draws[i].seg_begin,
                     &draws[i].seg_end);
        draws[i].fill_rule = r.fill_rule;
        draws[i].clip_path = r.clip_path;
        draws[i].paint = r.paint;
        draws[i].occludes = r.paint < palette.opaque.size() && palette.opaque[r.paint];
    }
    const float vb[4] = {scene->view_box.x0, scene->view_box.y0, scene->view_box.x1,
                         scene->view_box.y1};
    unsigned hw = std::thread::hardware_concurrency();
    int threads = (int)std::min(16u, std::max(1u, hw));
    pfh9::build(list->clip_points.data(), list->clip_indices.data(), clips,
                list->draw_points.data(), list->draw_indices.data(), draws, transform, vb,
                threads, &list->batch_d3d9);
    list->fills_d3d9 = std::move(list->batch_d3d9.fills);
    for (size_t i = 0; i < list->fills_d3d9.size(); i++)
        if (!list->fills_d3d9[i].empty())
            list->commands.push_back({PFH_COMMAND_ADD_FILLS_D3D9, (uint32_t)i});
    list->commands.push_back({PFH_COMMAND_FLUSH_FILLS_D3D9, 0});
    if (!scene->draw.paths.empty()) list->commands.push_back({PFH_COMMAND_DRAW_TILES_D3D9, 0});
    list->commands.push_back({PFH_COMMAND_FINISH, 0});
    *out = list.release();
    return PFH_OK;
}

pfh_status pfh_command_fills_d3d9(const pfh_command_list *list, uint32_t index,
                                  const pfh_fill_d3d9 **fills, uint32_t *count) {
    if (!list || !fills || !count || index >= list->commands.size() ||
        list->commands[index].kind != PFH_COMMAND_ADD_FILLS_D3D9)
        return PFH_ERROR_INVALID_ARGUMENT;
    const std::vector<pfh_fill_d3d9> &f = list->fills_d3d9[list->commands[index].batch_slot];
    *fills = f.data();
    *count = (uint32_t)f.size();
    return PFH_OK;
}

pfh_status pfh_command_batch_d3d9(const pfh_command_list *list, uint32_t index,
                                  pfh_draw_tile_batch_d3d9 *out) {
    if (!list || !out || index >= list->commands.size() ||
        list->commands[index].kind != PFH_COMMAND_DRAW_TILES_D3D9)
        return PFH_ERROR_INVALID_ARGUMENT;
    const pfh9::Built &b = list->batch_d3d9;
    out->tiles = b.tiles.data();
    out->tile_count = (uint32_t)b.tiles.size();
    out->clips = b.clips.data();
    out->clip_count = (uint32_t)b.clips.size();
    out->z_buffer = b.z.data();
    out->z_buffer_rect = b.z_rect;
    return PFH_OK;
}

void pfh_command_list_destroy(pfh_command_list *list) { delete list; }

uint32_t pfh_command_list_len(const pfh_command_list *list) {
    return list ? (uint32_t)list->commands.size() : 0;
}

int32_t pfh_command_kind_at(const pfh_command_list *list, uint32_t index) {
    if (!list || index >= list->commands.size()) return -1;
    return list->commands[index].kind;
}

pfh_status pfh_command_batch(const pfh_command_list *list, uint32_t index, pfh_tile_batch *out) {
    if (!list || !out || index >= list->commands.size()) return PFH_ERROR_INVALID_ARGUMENT;
    const pfh_command &c = list->commands[index];
    if (c.kind != PFH_COMMAND_PREPARE_CLIP_TILES_D3D11 && c.kind != PFH_COMMAND_DRAW_TILES_D3D11)
        return PFH_ERROR_INVALID_ARGUMENT;
    pfh::batch_view(list->batches[c.batch_slot], out);
    return PFH_OK;
}

pfh_status pfh_command_segments(const pfh_command_list *list, uint32_t index, pfh_segments *draw,
                                pfh_segments *clip) {
    if (!list || index >= list->commands.size() ||
        list->commands[index].kind != PFH_COMMAND_UPLOAD_SCENE_D3D11)
        return PFH_ERROR_INVALID_ARGUMENT;
    if (draw) {
        draw->points = list->draw_points.data();
        draw->point_count = (uint32_t)list->draw_points.size();
        draw->indices = list->draw_indices.data();
        draw->index_count = (uint32_t)list->draw_indices.size();
    }
    if (clip) {
        clip->points = list->clip_points.data();
        clip->point_count = (uint32_t)list->clip_points.size();
        clip->indices = list->clip_indices.data();
        clip->index_count = (uint32_t)list->clip_indices.size();
    }
    return PFH_OK;
}

pfh_status pfh_command_paints(const pfh_command_list *list, uint32_t index,
                              const pfh_color_u **colors, uint32_t *count) {
    if (!list || !colors || !count || index >= list->commands.size() ||
        list->commands[index].kind != PFH_COMMAND_UPLOAD_TEXTURE_METADATA)
        return PFH_ERROR_INVALID_ARGUMENT;
    *colors = list->paints.data();
    *count = (uint32_t)list->paints.size();
    return PFH_OK;
}

pfh_status pfh_command_texture_metadata(const pfh_command_list *list, uint32_t index,
                                        const pfh_texture_metadata_entry **entries,
                                        uint32_t *count) {
    if (!list || !entries || !count || index >= list->commands.size() ||
        list->commands[index].kind != PFH_COMMAND_UPLOAD_TEXTURE_METADATA)
        return PFH_ERROR_INVALID_ARGUMENT;
    *entries = list->metadata.data();
    *count = (uint32_t)list->metadata.size();
    return PFH_OK;
}

pfh_status pfh_command_texture_page(const pfh_command_list *list, uint32_t index,
                                    uint32_t *page_id, int32_t *width, int32_t *height) {
    if (!list || !page_id || !width || !height || index >= list->commands.size() ||
        list->commands[index].kind != PFH_COMMAND_ALLOCATE_TEXTURE_PAGE)
        return PFH_ERROR_INVALID_ARGUMENT;
    uint32_t slot = list->commands[index].batch_slot;
    *page_id = slot;
    *width = list->pages[slot].width;
    *height = list->pages[slot].height;
    return PFH_OK;
}

pfh_status pfh_command_texel_data(const pfh_command_list *list, uint32_t index,
                                  uint32_t *page_id, pfh_rect_i *rect,
                                  const pfh_color_u **texels) {
    if (!list || !page_id || !rect || !texels || index >= list->commands.size() ||
        list->commands[index].kind != PFH_COMMAND_UPLOAD_TEXEL_DATA)
        return PFH_ERROR_INVALID_ARGUMENT;
    const pfh::TexelUpload &u = list->uploads[list->commands[index].batch_slot];
    *page_id = u.page;
    *rect = u.rect;
    *texels = u.texels.data();
    return PFH_OK;
}

pfh_status pfh_render_commands(pfh_context *ctx, const pfh_command_list *list) {
    if (!ctx || !list) return PFH_ERROR_INVALID_ARGUMENT;
    for (size_t i = 0; i < list->commands.size(); i++) {
        const pfh_command &c = list->commands[i];
        pfh_status s = PFH_OK;
        switch (c.kind) {
        case PFH_COMMAND_UPLOAD_TEXTURE_METADATA:
            s = pfh_upload_texture_metadata_entries(ctx, list->metadata.data(),
                                                    (uint32_t)list->metadata.size());
            break;
        case PFH_COMMAND_ALLOCATE_TEXTURE_PAGE: {
            const pfh::PageSize &pg = list->pages[c.batch_slot];
            s = pfh_allocate_texture_page(ctx, c.batch_slot, pg.width, pg.height);
            break;
        }
        case PFH_COMMAND_UPLOAD_TEXEL_DATA: {
            const pfh::TexelUpload &u = list->uploads[c.batch_slot];
            s = pfh_upload_texel_data(ctx, u.page, &u.rect, u.texels.data());
            break;
        }
        case PFH_COMMAND_ADD_FILLS_D3D9: {
            const std::vector<pfh_fill_d3d9> &f = list->fills_d3d9[c.batch_slot];
            s = pfh_d3d9_add_fills(ctx, f.data(), (uint32_t)f.size());
            break;
        }
        case PFH_COMMAND_FLUSH_FILLS_D3D9:
            s = pfh_d3d9_flush_fills(ctx);
            break;
        case PFH_COMMAND_DRAW_TILES_D3D9: {
            pfh_draw_tile_batch_d3d9 b;
            pfh_command_batch_d3d9(list, (uint32_t)i, &b);
            s = pfh_d3d9_draw_tiles(ctx, &b);
            break;
        }
        case PFH_COMMAND_UPLOAD_SCENE_D3D11: {
            pfh_segments d, k;
            pfh_command_segments(list, (uint32_t)i, &d, &k);
            s = pfh_upload_scene(ctx, &d, &k);
            break;
        }
        case PFH_COMMAND_PREPARE_CLIP_TILES_D3D11: {
            pfh_tile_batch b;
            pfh::batch_view(list->batches[c.batch_slot], &b);
            s = pfh_prepare_tiles(ctx, &b);
            break;
        }
        case PFH_COMMAND_DRAW_TILES_D3D11: {
            pfh_tile_batch b;
            pfh::batch_view(list->batches[c.batch_slot], &b);
            s = pfh_prepare_and_draw_tiles(ctx, &b);
            break;
        }
        default:
            break;
        }
        if (s != PFH_OK) return s;
    }
    return PFH_OK;
}

pfh_status pfh_scene_build_and_render(pfh_scene *scene, pfh_context *ctx,
                                      const pfh_frame_options *frame,
                                      const pfh_build_options *options) {
    if (!scene || !ctx || !frame) return PFH_ERROR_INVALID_ARGUMENT;
    pfh_command_list *list = nullptr;
    pfh_status s = pfh_scene_build(scene, options, &list);
    if (s != PFH_OK) return s;
    s = pfh_begin_scene(ctx, frame);
    if (s == PFH_OK) s = pfh_render_commands(ctx, list);
    if (s == PFH_OK) s = pfh_end_scene(ctx);
    pfh_command_list_destroy(list);
    return s;
}

pfh_status pfh_scene_build_and_render_d3d9(pfh_scene *scene, pfh_context *ctx,
                                           const pfh_frame_options *frame,
                                           const pfh_build_options *options) {
    if (!scene || !ctx || !frame) return PFH_ERROR_INVALID_ARGUMENT;
    pfh_command_list *list = nullptr;
    pfh_status s = pfh_scene_build_d3d9(scene, options, &list);
    if (s != PFH_OK) return s;
    s = pfh_begin_scene(ctx, frame);
    if (s == PFH_OK) s = pfh_render_commands(ctx, list);
    if (s == PFH_OK) s = pfh_end_scene(ctx);
    pfh_command_list_destroy(list);
    return s;
}

}  // extern "C"
