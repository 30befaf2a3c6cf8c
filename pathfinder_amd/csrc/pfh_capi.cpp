// pfh_capi.cpp - the reference's C API (c/src/lib.rs) over pfh_scene_* and
// pfh_context_* (include/pathfinder.h).
//
// Canvas: CanvasRenderingContext2D (canvas/src/lib.rs) without text: a state
// stack (transform, fill and stroke colour, line width 1, butt caps, miter
// joins, miter limit 10: canvas/src/lib.rs:655-675), fills with the nonzero
// rule under the current transform, strokes stroked in path space and then
// transformed (as canvas/src/lib.rs does), each pushed as a draw path.
// Path2D: contours of lines, quadratics and cubics; arcs as cubic pieces of
// at most 90 degrees (Contour::push_arc, content/src/outline.rs:615-627: a
// sweep of 2 pi or more is a whole ellipse; otherwise from the start to the
// end angle, increasing for CW (clockwise on the y-down screen), decreasing
// for CCW). arcTo draws the short arc between the two tangent points
// (canvas/src/lib.rs:770-783 always sweeps CW, marked FIXME there).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include <zlib.h>

#include "../../include/pathfinder.h"
#include "../../include/pfhip.h"
#include "pfh_geom.h"

extern "C" const uint8_t pfh_area_lut_rgba[];

using namespace pfh_geom;

struct PFPath {
    std::vector<Subpath> subs;
    bool open_contour = false;  // the last subpath takes further segments
    V2 cur{0, 0};
};

struct PFFillStyle {
    PFColorU color;
};

struct PFResourceLoader {
    std::vector<uint8_t> area_lut;  // 256 x 256 RGBA8
};

struct PFRenderTransform {
    pfh_transform2f t;
};

struct PFBuildOptions {
    pfh_transform2f t{1, 0, 0, 1, 0, 0};
    float dilation[2] = {0, 0};
    bool subpixel_aa = false;
};

struct PFSceneProxy {
    pfh_scene *scene;
    PFRendererLevel level;
};

struct PFSVGScene {
    pfh_scene *scene;
};

struct PFHIPDevice {
    int32_t ordinal;
};

struct PFHIPDestFramebuffer {
    int32_t width, height;
};

struct PFHIPRenderer {
    PFRendererLevel level = PF_RENDERER_LEVEL_D3D11;
    PFHIPDevice *device = nullptr;
    PFHIPDestFramebuffer *dest = nullptr;
    pfh_context *ctx = nullptr;
    pfh_frame_options frame{};
    std::string error;
};

namespace {

struct CanvasState {
    Xf xf;
    PFColorU fill{0, 0, 0, 255}, stroke{0, 0, 0, 255};
    StrokeParams stroke_params;
    CanvasState() { stroke_params.miter_limit = 10.0; }
};

uint32_t push_outline(pfh_scene *scene, const OutlineBuf &o, PFColorU color, int32_t rule) {
    if (o.ends.empty()) return 0xffffffffu;
    pfh_color_u c{color.r, color.g, color.b, color.a};
    uint32_t paint = pfh_scene_push_paint(scene, &c);
    pfh_outline ol{o.pts.data(), o.flags.data(), (uint32_t)o.pts.size(), o.ends.data(),
                   (uint32_t)o.ends.size()};
    return pfh_scene_push_draw_path(scene, &ol, paint, rule, -1, PFH_BLEND_MODE_SRC_OVER);
}

// Cubic pieces (<= 90 degrees each) of the arc of the ellipse
// center + R(rotation) diag(axes) (cos t, sin t) for t from a0 sweeping by `sweep`.
void push_elliptic_arc(PFPath &p, V2 center, V2 axes, double rotation, double a0, double sweep) {
    double cr = std::cos(rotation), sr = std::sin(rotation);
    auto at = [&](double t) {
        double x = axes.x * std::cos(t), y = axes.y * std::sin(t);
        return V2{center.x + cr * x - sr * y, center.y + sr * x + cr * y};
    };
    auto der = [&](double t) {
        double x = -axes.x * std::sin(t), y = axes.y * std::cos(t);
        return V2{cr * x - sr * y, sr * x + cr * y};
    };
    V2 start = at(a0);
    if (!p.open_contour || p.subs.empty()) {
        p.subs.push_back(Subpath());
        p.subs.back().start = start;
        p.open_contour = true;
    } else if (len(start - p.cur) > 0) {
        p.subs.back().segs.push_back({false, {}, {}, start});
    }
    int n = std::max(1, (int)std::ceil(std::fabs(sweep) / (M_PI / 2) - 1e-9));
    double step = sweep / n, k = 4.0 / 3.0 * std::tan(step / 4);
    for (int i = 0; i < n; i++) {
        double ta = a0 + step * i, tb = ta + step;
        V2 a = at(ta), b = at(tb);
        p.subs.back().segs.push_back({true, a + der(ta) * k, b - der(tb) * k, b});
    }
    p.cur = at(a0 + sweep);
}

// Contour::push_arc's sweep (content/src/outline.rs:615-627).
double arc_sweep(double start, double end, bool cw) {
    const double two_pi = 2 * M_PI;
    if (cw) {
        if (end - start >= two_pi) return two_pi;
        double s = std::fmod(end - start, two_pi);
        return s < 0 ? s + two_pi : s;
    }
    if (start - end >= two_pi) return -two_pi;
    double s = std::fmod(start - end, two_pi);
    return -(s < 0 ? s + two_pi : s);
}

void line_to(PFPath &p, V2 to) {
    if (!p.open_contour || p.subs.empty()) {
        p.subs.push_back(Subpath());
        p.subs.back().start = p.cur;
        p.open_contour = true;
    }
    p.subs.back().segs.push_back({false, {}, {}, to});
    p.cur = to;
}

bool decode_png_rgba(const std::vector<uint8_t> &png, uint32_t want_w, uint32_t want_h,
                     std::vector<uint8_t> &out) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (png.size() < 8 || std::memcmp(png.data(), sig, 8) != 0) return false;
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int bpp = 0;
    std::vector<uint8_t> idat;
    auto be32 = [&](size_t o) {
        return ((uint32_t)png[o] << 24) | ((uint32_t)png[o + 1] << 16) | ((uint32_t)png[o + 2] << 8) |
               png[o + 3];
    };
    while (pos + 12 <= png.size()) {
        uint32_t n = be32(pos);
        if (pos + 12 + (size_t)n > png.size()) return false;
        const uint8_t *type = &png[pos + 4];
        const uint8_t *data = &png[pos + 8];
        if (!std::memcmp(type, "IHDR", 4) && n >= 13) {
            w = be32(pos + 8);
            h = be32(pos + 12);
            uint8_t depth = data[8], color = data[9], interlace = data[12];
            if (depth != 8 || interlace != 0) return false;
            bpp = color == 6 ? 4 : color == 2 ? 3 : color == 0 ? 1 : 0;
            if (!bpp) return false;
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + n);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + n;
    }
    if (w != want_w || h != want_h || !bpp) return false;
    size_t stride = (size_t)w * bpp;
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf got = (uLongf)raw.size();
    if (uncompress(raw.data(), &got, idat.data(), (uLong)idat.size()) != Z_OK || got != raw.size())
        return false;
    std::vector<uint8_t> img(stride * h), prev(stride, 0);
    for (uint32_t y = 0; y < h; y++) {
        uint8_t f = raw[y * (stride + 1)];
        const uint8_t *src = &raw[y * (stride + 1) + 1];
        uint8_t *dst = &img[y * stride];
        for (size_t i = 0; i < stride; i++) {
            int a = i >= (size_t)bpp ? dst[i - bpp] : 0, b = prev[i],
                c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            int v = src[i];
            switch (f) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: {
                int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                break;
            }
            default: return false;
            }
            dst[i] = (uint8_t)v;
        }
        std::memcpy(prev.data(), dst, stride);
    }
    out.resize((size_t)w * h * 4);
    for (size_t i = 0; i < (size_t)w * h; i++) {
        for (int c = 0; c < 4; c++)
            out[i * 4 + c] = bpp == 4 ? img[i * 4 + c]
                             : c == 3 ? 255
                             : bpp == 3 ? img[i * 3 + c] : img[i];
    }
    return true;
}

bool read_file(const char *path, std::vector<uint8_t> &out) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return false;
    }
    out.resize((size_t)n);
    size_t got = n ? std::fread(out.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == (size_t)n;
}

}  // namespace

struct PFCanvas {
    pfh_scene *scene = nullptr;
    std::vector<CanvasState> stack{CanvasState()};
    CanvasState &state() { return stack.back(); }
};

extern "C" {

// ------------------------------------------------------------ canvas

PFCanvasRef PFCanvasCreate(const PFVector2F *size) {
    if (!size) return nullptr;
    PFCanvas *c = new (std::nothrow) PFCanvas();
    if (!c) return nullptr;
    c->scene = pfh_scene_create();
    if (!c->scene) {
        delete c;
        return nullptr;
    }
    pfh_rect_f vb{{0, 0}, {size->x, size->y}};
    pfh_scene_set_view_box(c->scene, &vb);
    return c;
}

void PFCanvasDestroy(PFCanvasRef canvas) {
    if (!canvas) return;
    pfh_scene_destroy(canvas->scene);
    delete canvas;
}

PFSceneRef PFCanvasCreateScene(PFCanvasRef canvas) {
    if (!canvas) return nullptr;
    pfh_scene *s = canvas->scene;
    canvas->scene = nullptr;
    delete canvas;
    return s;
}

static void canvas_fill(PFCanvas *c, const PFPath &p) {
    OutlineBuf o;
    fill_outline(p.subs, c->state().xf, o);
    push_outline(c->scene, o, c->state().fill, PFH_FILL_RULE_WINDING);
}

static void canvas_stroke(PFCanvas *c, const PFPath &p) {
    OutlineBuf o;
    stroke_outline(p.subs, c->state().stroke_params, c->state().xf, o);
    push_outline(c->scene, o, c->state().stroke, PFH_FILL_RULE_WINDING);
}

void PFCanvasFillRect(PFCanvasRef canvas, const PFRectF *rect) {
    if (!canvas || !rect) return;
    PFPath p;
    PFPathRect(&p, rect);
    canvas_fill(canvas, p);
}

void PFCanvasStrokeRect(PFCanvasRef canvas, const PFRectF *rect) {
    if (!canvas || !rect) return;
    PFPath p;
    PFPathRect(&p, rect);
    canvas_stroke(canvas, p);
}

void PFCanvasSetLineWidth(PFCanvasRef canvas, float w) {
    if (canvas) canvas->state().stroke_params.stroke_width = w;
}
void PFCanvasSetLineCap(PFCanvasRef canvas, PFLineCap cap) {
    if (canvas) canvas->state().stroke_params.cap = cap <= 2 ? cap : 0;
}
void PFCanvasSetLineJoin(PFCanvasRef canvas, PFLineJoin join) {
    if (canvas) canvas->state().stroke_params.join = join <= 2 ? join : 0;
}
void PFCanvasSetMiterLimit(PFCanvasRef canvas, float limit) {
    if (canvas) canvas->state().stroke_params.miter_limit = limit;
}

void PFCanvasSetTransform(PFCanvasRef canvas, const PFTransform2F *t) {
    if (!canvas || !t) return;
    Xf &x = canvas->state().xf;
    x.a = t->matrix.m00; x.c = t->matrix.m01; x.e = t->vector.x;
    x.b = t->matrix.m10; x.d = t->matrix.m11; x.f = t->vector.y;
}

void PFCanvasResetTransform(PFCanvasRef canvas) {
    if (canvas) canvas->state().xf = Xf();
}

void PFCanvasSave(PFCanvasRef canvas) {
    if (canvas) canvas->stack.push_back(canvas->state());
}

void PFCanvasRestore(PFCanvasRef canvas) {
    if (canvas && canvas->stack.size() > 1) canvas->stack.pop_back();
}

void PFCanvasSetFillStyle(PFCanvasRef canvas, PFFillStyleRef style) {
    if (canvas && style) canvas->state().fill = style->color;
}

void PFCanvasSetStrokeStyle(PFCanvasRef canvas, PFFillStyleRef style) {
    if (canvas && style) canvas->state().stroke = style->color;
}

void PFCanvasFillPath(PFCanvasRef canvas, PFPathRef path) {
    if (canvas && path) canvas_fill(canvas, *path);
    delete path;
}

void PFCanvasStrokePath(PFCanvasRef canvas, PFPathRef path) {
    if (canvas && path) canvas_stroke(canvas, *path);
    delete path;
}

// ------------------------------------------------------------ paths

PFPathRef PFPathCreate(void) { return new (std::nothrow) PFPath(); }
void PFPathDestroy(PFPathRef path) { delete path; }
PFPathRef PFPathClone(PFPathRef path) { return path ? new (std::nothrow) PFPath(*path) : nullptr; }

void PFPathMoveTo(PFPathRef path, const PFVector2F *to) {
    if (!path || !to) return;
    path->subs.push_back(Subpath());
    path->subs.back().start = V2{to->x, to->y};
    path->open_contour = true;
    path->cur = path->subs.back().start;
}

void PFPathLineTo(PFPathRef path, const PFVector2F *to) {
    if (path && to) line_to(*path, V2{to->x, to->y});
}

void PFPathQuadraticCurveTo(PFPathRef path, const PFVector2F *ctrl, const PFVector2F *to) {
    if (!path || !ctrl || !to) return;
    if (!path->open_contour || path->subs.empty()) PFPathMoveTo(path, ctrl);
    V2 q{ctrl->x, ctrl->y}, t{to->x, to->y}, c = path->cur;
    path->subs.back().segs.push_back({true, c + (q - c) * (2.0 / 3.0), t + (q - t) * (2.0 / 3.0), t});
    path->cur = t;
}

void PFPathBezierCurveTo(PFPathRef path, const PFVector2F *c0, const PFVector2F *c1,
                         const PFVector2F *to) {
    if (!path || !c0 || !c1 || !to) return;
    if (!path->open_contour || path->subs.empty()) PFPathMoveTo(path, c0);
    V2 t{to->x, to->y};
    path->subs.back().segs.push_back({true, V2{c0->x, c0->y}, V2{c1->x, c1->y}, t});
    path->cur = t;
}

void PFPathArc(PFPathRef path, const PFVector2F *center, float radius, float start_angle,
               float end_angle, PFArcDirection direction) {
    if (!path || !center) return;
    push_elliptic_arc(*path, V2{center->x, center->y}, V2{radius, radius}, 0.0, start_angle,
                      arc_sweep(start_angle, end_angle, direction == PF_ARC_DIRECTION_CW));
}

void PFPathArcTo(PFPathRef path, const PFVector2F *ctrl, const PFVector2F *to, float radius) {
    if (!path || !ctrl || !to) return;
    V2 c{ctrl->x, ctrl->y}, t{to->x, to->y};
    if (!path->open_contour || path->subs.empty()) PFPathMoveTo(path, ctrl);
    V2 from = path->cur;
    V2 v0 = from - c, v1 = t - c;
    double l0 = len(v0), l1 = len(v1);
    if (radius <= 0 || l0 == 0 || l1 == 0 || std::fabs(cross(v0, v1)) < 1e-12 * l0 * l1) {
        line_to(*path, c);
        return;
    }
    V2 u0 = v0 * (1.0 / l0), u1 = v1 * (1.0 / l1);
    double cos_t = std::max(-1.0, std::min(1.0, dot(u0, u1)));
    double half = std::acos(cos_t) / 2;  // half the angle at the corner
    double d = radius / std::tan(half);  // corner to tangent points
    V2 t0 = c + u0 * d, t1 = c + u1 * d;
    V2 bis = (u0 + u1) * (1.0 / len(u0 + u1));
    V2 center = c + bis * (radius / std::sin(half));
    line_to(*path, t0);
    double a0 = std::atan2(t0.y - center.y, t0.x - center.x);
    double a1 = std::atan2(t1.y - center.y, t1.x - center.x);
    double sweep = a1 - a0;
    while (sweep > M_PI) sweep -= 2 * M_PI;
    while (sweep < -M_PI) sweep += 2 * M_PI;
    push_elliptic_arc(*path, center, V2{radius, radius}, 0.0, a0, sweep);
}

void PFPathRect(PFPathRef path, const PFRectF *rect) {
    if (!path || !rect) return;
    Subpath sp;
    double x0 = rect->origin.x, y0 = rect->origin.y, x1 = rect->lower_right.x,
           y1 = rect->lower_right.y;
    sp.start = {x0, y0};
    sp.segs = {{false, {}, {}, {x1, y0}}, {false, {}, {}, {x1, y1}}, {false, {}, {}, {x0, y1}}};
    sp.closed = true;
    path->subs.push_back(sp);
    path->open_contour = false;
    path->cur = sp.start;
}

void PFPathEllipse(PFPathRef path, const PFVector2F *center, const PFVector2F *axes,
                   float rotation, float start_angle, float end_angle) {
    if (!path || !center || !axes) return;
    push_elliptic_arc(*path, V2{center->x, center->y}, V2{axes->x, axes->y}, rotation,
                      start_angle, arc_sweep(start_angle, end_angle, true));
}

void PFPathClosePath(PFPathRef path) {
    if (!path || path->subs.empty() || !path->open_contour) return;
    path->subs.back().closed = true;
    path->open_contour = false;
    path->cur = path->subs.back().start;
}

// ------------------------------------------------------------ fill styles

PFFillStyleRef PFFillStyleCreateColor(const PFColorU *color) {
    if (!color) return nullptr;
    PFFillStyle *s = new (std::nothrow) PFFillStyle();
    if (s) s->color = *color;
    return s;
}

void PFFillStyleDestroy(PFFillStyleRef fill_style) { delete fill_style; }

// ------------------------------------------------------------ resources

PFResourceLoaderRef PFEmbeddedResourceLoaderCreate(void) {
    PFResourceLoader *l = new (std::nothrow) PFResourceLoader();
    if (l) l->area_lut.assign(pfh_area_lut_rgba, pfh_area_lut_rgba + 256 * 256 * 4);
    return l;
}

PFResourceLoaderRef PFFilesystemResourceLoaderFromPath(const char *path) {
    if (!path) return nullptr;
    std::vector<uint8_t> png;
    if (!read_file((std::string(path) + "/textures/area-lut.png").c_str(), png)) return nullptr;
    PFResourceLoader *l = new (std::nothrow) PFResourceLoader();
    if (!l) return nullptr;
    if (!decode_png_rgba(png, 256, 256, l->area_lut)) {
        delete l;
        return nullptr;
    }
    return l;
}

void PFResourceLoaderDestroy(PFResourceLoaderRef loader) { delete loader; }

// ------------------------------------------------------------ HIP device and renderer

PFHIPDeviceRef PFHIPDeviceCreate(int32_t device_ordinal) {
    PFHIPDevice *d = new (std::nothrow) PFHIPDevice();
    if (d) d->ordinal = device_ordinal;
    return d;
}

void PFHIPDeviceDestroy(PFHIPDeviceRef device) { delete device; }

PFHIPDestFramebufferRef PFHIPDestFramebufferCreateFullWindow(const PFVector2I *size) {
    if (!size || size->x <= 0 || size->y <= 0) return nullptr;
    PFHIPDestFramebuffer *f = new (std::nothrow) PFHIPDestFramebuffer();
    if (f) {
        f->width = size->x;
        f->height = size->y;
    }
    return f;
}

void PFHIPDestFramebufferDestroy(PFHIPDestFramebufferRef dest) { delete dest; }

PFHIPRendererRef PFHIPRendererCreate(PFHIPDeviceRef device, PFResourceLoaderRef resources,
                                     const PFRendererMode *mode,
                                     const PFRendererOptions *options) {
    PFHIPDestFramebuffer *dest =
        options ? static_cast<PFHIPDestFramebuffer *>(options->dest) : nullptr;
    if (!device || !resources || !mode || !options || !dest ||
        (mode->level != PF_RENDERER_LEVEL_D3D11 && mode->level != PF_RENDERER_LEVEL_D3D9) ||
        resources->area_lut.size() != 256 * 256 * 4) {
        delete device;
        delete dest;
        return nullptr;
    }
    PFHIPRenderer *r = new (std::nothrow) PFHIPRenderer();
    if (!r) {
        delete device;
        delete dest;
        return nullptr;
    }
    r->level = mode->level;
    r->device = device;
    r->dest = dest;
    pfh_context_options co;
    std::memset(&co, 0, sizeof(co));
    co.device = device->ordinal;
    co.level = mode->level == PF_RENDERER_LEVEL_D3D9 ? PFH_RENDERER_LEVEL_D3D9
                                                     : PFH_RENDERER_LEVEL_D3D11;
    co.area_lut_rgba = resources->area_lut.data();
    if (pfh_context_create(&co, &r->ctx) != PFH_OK) {
        PFHIPRendererDestroy(r);
        return nullptr;
    }
    r->frame.width = dest->width;
    r->frame.height = dest->height;
    r->frame.has_background =
        (options->flags & PF_RENDERER_OPTIONS_FLAGS_HAS_BACKGROUND_COLOR) != 0;
    r->frame.background_color = {options->background_color.r, options->background_color.g,
                                 options->background_color.b, options->background_color.a};
    return r;
}

void PFHIPRendererDestroy(PFHIPRendererRef renderer) {
    if (!renderer) return;
    if (renderer->ctx) pfh_context_destroy(renderer->ctx);
    delete renderer->device;
    delete renderer->dest;
    delete renderer;
}

PFHIPDeviceRef PFHIPRendererGetDevice(PFHIPRendererRef renderer) {
    return renderer ? renderer->device : nullptr;
}

int32_t PFHIPRendererReadPixels(PFHIPRendererRef renderer, uint8_t *rgba, size_t stride) {
    if (!renderer || !rgba) return PFH_ERROR_INVALID_ARGUMENT;
    pfh_status s = pfh_read_framebuffer(renderer->ctx, rgba, stride);
    if (s != PFH_OK) renderer->error = pfh_last_error(renderer->ctx);
    return s;
}

const char *PFHIPRendererGetLastError(PFHIPRendererRef renderer) {
    return renderer ? renderer->error.c_str() : "";
}

// ------------------------------------------------------------ build options

PFRenderTransformRef PFRenderTransformCreate2D(const PFTransform2F *t) {
    if (!t) return nullptr;
    PFRenderTransform *r = new (std::nothrow) PFRenderTransform();
    if (r)
        r->t = pfh_transform2f{t->matrix.m00, t->matrix.m01, t->matrix.m10, t->matrix.m11,
                               t->vector.x, t->vector.y};
    return r;
}

void PFRenderTransformDestroy(PFRenderTransformRef transform) { delete transform; }

PFBuildOptionsRef PFBuildOptionsCreate(void) { return new (std::nothrow) PFBuildOptions(); }

void PFBuildOptionsDestroy(PFBuildOptionsRef options) { delete options; }

void PFBuildOptionsSetTransform(PFBuildOptionsRef options, PFRenderTransformRef transform) {
    if (options && transform) options->t = transform->t;
    delete transform;
}

void PFBuildOptionsSetDilation(PFBuildOptionsRef options, const PFVector2F *dilation) {
    if (options && dilation) {
        options->dilation[0] = dilation->x;
        options->dilation[1] = dilation->y;
    }
}

void PFBuildOptionsSetSubpixelAAEnabled(PFBuildOptionsRef options, int enabled) {
    if (options) options->subpixel_aa = enabled != 0;
}

// ------------------------------------------------------------ scenes

void PFSceneDestroy(PFSceneRef scene) { pfh_scene_destroy(scene); }

PFSceneProxyRef PFSceneProxyCreateFromSceneAndRayonExecutor(PFSceneRef scene,
                                                            PFRendererLevel level) {
    if (!scene) return nullptr;
    PFSceneProxy *p = new (std::nothrow) PFSceneProxy();
    if (!p) {
        pfh_scene_destroy(scene);
        return nullptr;
    }
    p->scene = scene;
    p->level = level;
    return p;
}

void PFSceneProxyDestroy(PFSceneProxyRef proxy) {
    if (!proxy) return;
    pfh_scene_destroy(proxy->scene);
    delete proxy;
}

void PFSceneProxyBuildAndRenderHIP(PFSceneProxyRef proxy, PFHIPRendererRef renderer,
                                   PFBuildOptionsRef options) {
    if (!proxy || !renderer) return;
    PFBuildOptions defaults;
    const PFBuildOptions &o = options ? *options : defaults;
    if (o.dilation[0] != 0 || o.dilation[1] != 0 || o.subpixel_aa) {
        renderer->error = "dilation and subpixel AA are not supported";
        return;
    }
    if (proxy->level != renderer->level) {
        renderer->error = "the scene proxy's level is not the renderer's";
        return;
    }
    pfh_build_options bo;
    bo.transform = o.t;
    pfh_status s = proxy->level == PF_RENDERER_LEVEL_D3D9
                       ? pfh_scene_build_and_render_d3d9(proxy->scene, renderer->ctx,
                                                         &renderer->frame, &bo)
                       : pfh_scene_build_and_render(proxy->scene, renderer->ctx,
                                                    &renderer->frame, &bo);
    if (s == PFH_OK) renderer->error.clear();
    else if (*pfh_last_error(renderer->ctx)) renderer->error = pfh_last_error(renderer->ctx);
    else renderer->error = "the scene has a paint or blend mode this level does not draw";
}

// ------------------------------------------------------------ SVG

PFSVGSceneRef PFSVGSceneCreateWithMemory(const char *bytes, size_t byte_len) {
    pfh_scene *s = pfh_svg_scene_create(bytes, byte_len, nullptr);
    if (!s) return nullptr;
    PFSVGScene *svg = new (std::nothrow) PFSVGScene();
    if (!svg) {
        pfh_scene_destroy(s);
        return nullptr;
    }
    svg->scene = s;
    return svg;
}

PFSVGSceneRef PFSVGSceneCreateWithPath(const char *path) {
    std::vector<uint8_t> data;
    if (!path || !read_file(path, data)) return nullptr;
    return PFSVGSceneCreateWithMemory(reinterpret_cast<const char *>(data.data()), data.size());
}

PFSceneRef PFSVGSceneIntoScene(PFSVGSceneRef svg) {
    if (!svg) return nullptr;
    pfh_scene *s = svg->scene;
    delete svg;
    return s;
}

}  // extern "C"
