// pfh_svg.cpp - SVG ingest and stroking for the scene front-end.
//
// The reference builds scenes from SVG with its svg crate (svg/src/lib.rs:
// SVGScene::from_tree, :65-290) over usvg, and turns strokes into fills with
// OutlineStrokeToFill (content/src/stroke.rs). This is this library's own
// reader for the subset a filled-and-stroked vector document uses:
//  - elements svg (viewBox, else width/height), g, path, rect (rx/ry),
//    circle, ellipse, line, polyline, polygon; every other element and its
//    children are skipped (and counted);
//  - presentation attributes and inline style: fill, fill-opacity,
//    fill-rule, stroke, stroke-width, stroke-opacity, stroke-linecap,
//    stroke-linejoin, stroke-miterlimit, display, transform (matrix,
//    translate, scale, rotate, skewX, skewY); `opacity` is not applied, as
//    the reference applies none (svg/src/lib.rs:94-123);
//  - colours #rgb, #rrggbb, rgb() (numbers or percentages) and the common
//    named colours; paint alpha = opacity x 255, rounded (svg/src/lib.rs:361);
//  - path data M L H V C S Q T A Z in absolute and relative forms; like
//    usvg, quadratics and arcs become cubics;
//  - a fill is the path's outline under the current transform; a stroke is
//    stroked in user space (width at least 0.0333, svg/src/lib.rs:38, 150)
//    and then transformed, filled with the nonzero rule, after the fill.
//
// Stroking: curves are flattened with Sederberg's bound at a tolerance of
// 0.1 device pixel (the transform's scale), and every segment (a rectangle
// about it), join (bevel triangle, miter quad within the miter limit, or a
// round fan) and cap (square or round) is its own contour, oriented
// counter-clockwise in the scene's y-down coordinates; the nonzero fill of
// their union is the stroke, with no self-intersecting offset curves.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <new>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/pfhip.h"
#include "pfh_geom.h"

struct pfh_outline_buffer {
    std::vector<pfh_vector2f> points;
    std::vector<uint8_t> flags;
    std::vector<uint32_t> contour_ends;
    pfh_outline view;
};

namespace {

using namespace pfh_geom;

// ------------------------------------------------------------ path data

struct PathParser {
    const char *s;
    bool ok = true;

    void skip_ws() {
        while (*s && (std::isspace((unsigned char)*s) || *s == ',')) s++;
    }
    bool number(double *out) {
        skip_ws();
        const char *p = s;
        if (*p == '+' || *p == '-') p++;
        bool digits = false;
        while (std::isdigit((unsigned char)*p)) { p++; digits = true; }
        if (*p == '.') {
            p++;
            while (std::isdigit((unsigned char)*p)) { p++; digits = true; }
        }
        if (!digits) return false;
        if (*p == 'e' || *p == 'E') {
            const char *q = p + 1;
            if (*q == '+' || *q == '-') q++;
            if (std::isdigit((unsigned char)*q)) {
                while (std::isdigit((unsigned char)*q)) q++;
                p = q;
            }
        }
        *out = std::strtod(std::string(s, p).c_str(), nullptr);
        s = p;
        return true;
    }
    bool flag(bool *out) {  // arc flags may be written without separators
        skip_ws();
        if (*s == '0' || *s == '1') {
            *out = *s == '1';
            s++;
            return true;
        }
        return false;
    }
};

bool parse_path_data(const char *d, std::vector<Subpath> &subs) {
    PathParser P{d};
    char cmd = 0;
    V2 cur{0, 0}, start{0, 0}, last_ctrl{0, 0};
    char last_kind = 0;  // 'C' or 'Q' when last_ctrl is a reflectable control point
    Subpath *sp = nullptr;
    auto ensure = [&]() {
        if (!sp) {
            subs.push_back(Subpath());
            sp = &subs.back();
            sp->start = cur;
        }
    };
    while (true) {
        P.skip_ws();
        if (!*P.s) break;
        if (std::isalpha((unsigned char)*P.s)) {
            cmd = *P.s++;
        } else if (!cmd) {
            return false;
        }
        bool rel = std::islower((unsigned char)cmd) != 0;
        char C = (char)std::toupper((unsigned char)cmd);
        V2 base = rel ? cur : V2{0, 0};
        double v[7];
        auto nums = [&](int n) {
            for (int i = 0; i < n; i++)
                if (!P.number(&v[i])) return false;
            return true;
        };
        switch (C) {
        case 'M': {
            if (!nums(2)) return false;
            cur = base + V2{v[0], v[1]};
            start = cur;
            subs.push_back(Subpath());
            sp = &subs.back();
            sp->start = cur;
            cmd = rel ? 'l' : 'L';  // later pairs are implicit lines
            last_kind = 0;
            continue;
        }
        case 'Z':
            if (sp) sp->closed = true;
            cur = start;
            sp = nullptr;
            last_kind = 0;
            cmd = 0;
            continue;
        case 'L':
            if (!nums(2)) return false;
            ensure();
            cur = base + V2{v[0], v[1]};
            sp->segs.push_back({false, {0, 0}, {0, 0}, cur});
            last_kind = 0;
            break;
        case 'H':
            if (!nums(1)) return false;
            ensure();
            cur = V2{(rel ? cur.x : 0) + v[0], cur.y};
            sp->segs.push_back({false, {0, 0}, {0, 0}, cur});
            last_kind = 0;
            break;
        case 'V':
            if (!nums(1)) return false;
            ensure();
            cur = V2{cur.x, (rel ? cur.y : 0) + v[0]};
            sp->segs.push_back({false, {0, 0}, {0, 0}, cur});
            last_kind = 0;
            break;
        case 'C': {
            if (!nums(6)) return false;
            ensure();
            V2 c0 = base + V2{v[0], v[1]}, c1 = base + V2{v[2], v[3]}, to = base + V2{v[4], v[5]};
            sp->segs.push_back({true, c0, c1, to});
            cur = to;
            last_ctrl = c1;
            last_kind = 'C';
            break;
        }
        case 'S': {
            if (!nums(4)) return false;
            ensure();
            V2 c0 = last_kind == 'C' ? cur * 2.0 - last_ctrl : cur;
            V2 c1 = base + V2{v[0], v[1]}, to = base + V2{v[2], v[3]};
            sp->segs.push_back({true, c0, c1, to});
            cur = to;
            last_ctrl = c1;
            last_kind = 'C';
            break;
        }
        case 'Q':
        case 'T': {
            V2 q, to;
            if (C == 'Q') {
                if (!nums(4)) return false;
                q = base + V2{v[0], v[1]};
                to = base + V2{v[2], v[3]};
            } else {
                if (!nums(2)) return false;
                q = last_kind == 'Q' ? cur * 2.0 - last_ctrl : cur;
                to = base + V2{v[0], v[1]};
            }
            ensure();
            // exact degree elevation, as usvg converts quadratics
            sp->segs.push_back({true, cur + (q - cur) * (2.0 / 3.0), to + (q - to) * (2.0 / 3.0),
                                to});
            cur = to;
            last_ctrl = q;
            last_kind = 'Q';
            break;
        }
        case 'A': {
            bool large, sweep;
            if (!P.number(&v[0]) || !P.number(&v[1]) || !P.number(&v[2]) || !P.flag(&large) ||
                !P.flag(&sweep) || !P.number(&v[3]) || !P.number(&v[4]))
                return false;
            ensure();
            V2 to = base + V2{v[3], v[4]};
            arc_to_cubics(cur, v[0], v[1], v[2], large, sweep, to, sp->segs);
            cur = to;
            last_kind = 0;
            break;
        }
        default:
            return false;
        }
    }
    return true;
}

// ------------------------------------------------------------ attributes

struct Style {
    bool fill_none = false, stroke_none = true;
    uint8_t fill[3] = {0, 0, 0}, stroke[3] = {0, 0, 0};
    double fill_opacity = 1, stroke_opacity = 1;
    int fill_rule = PFH_FILL_RULE_WINDING;
    double stroke_width = 1, miter_limit = 4;
    int cap = 0;   // butt, square, round
    int join = 0;  // miter, bevel, round
    bool unsupported_paint = false;
    Xf xf;
};

const std::unordered_map<std::string, uint32_t> &named_colors() {
    static const std::unordered_map<std::string, uint32_t> m = {
        {"black", 0x000000}, {"white", 0xffffff}, {"red", 0xff0000}, {"green", 0x008000},
        {"blue", 0x0000ff}, {"yellow", 0xffff00}, {"cyan", 0x00ffff}, {"aqua", 0x00ffff},
        {"magenta", 0xff00ff}, {"fuchsia", 0xff00ff}, {"gray", 0x808080}, {"grey", 0x808080},
        {"silver", 0xc0c0c0}, {"maroon", 0x800000}, {"olive", 0x808000}, {"lime", 0x00ff00},
        {"teal", 0x008080}, {"navy", 0x000080}, {"purple", 0x800080}, {"orange", 0xffa500},
        {"brown", 0xa52a2a}, {"pink", 0xffc0cb}, {"gold", 0xffd700}, {"indigo", 0x4b0082},
        {"violet", 0xee82ee}, {"darkgray", 0xa9a9a9}, {"darkgrey", 0xa9a9a9},
        {"lightgray", 0xd3d3d3}, {"lightgrey", 0xd3d3d3}, {"darkred", 0x8b0000},
        {"darkgreen", 0x006400}, {"darkblue", 0x00008b}, {"lightblue", 0xadd8e6},
        {"skyblue", 0x87ceeb}, {"steelblue", 0x4682b4}, {"tomato", 0xff6347},
        {"coral", 0xff7f50}, {"salmon", 0xfa8072}, {"khaki", 0xf0e68c}, {"beige", 0xf5f5dc},
        {"tan", 0xd2b48c}, {"chocolate", 0xd2691e}, {"crimson", 0xdc143c},
        {"orchid", 0xda70d6}, {"plum", 0xdda0dd}, {"turquoise", 0x40e0d0},
        {"darkorange", 0xff8c00}, {"forestgreen", 0x228b22}, {"seagreen", 0x2e8b57},
        {"slategray", 0x708090}, {"dimgray", 0x696969}, {"whitesmoke", 0xf5f5f5},
        {"ivory", 0xfffff0}, {"lavender", 0xe6e6fa}, {"wheat", 0xf5deb3}};
    return m;
}

std::string trim(const std::string &v) {
    size_t a = 0, b = v.size();
    while (a < b && std::isspace((unsigned char)v[a])) a++;
    while (b > a && std::isspace((unsigned char)v[b - 1])) b--;
    return v.substr(a, b - a);
}

// 0: a colour, 1: none, 2: unsupported (url(), unknown)
int parse_color(const std::string &raw, uint8_t rgb[3]) {
    std::string v = trim(raw);
    std::string lower;
    for (char ch : v) lower.push_back((char)std::tolower((unsigned char)ch));
    if (lower == "none" || lower == "transparent") return 1;
    if (lower == "currentcolor") {
        rgb[0] = rgb[1] = rgb[2] = 0;
        return 0;
    }
    if (!lower.empty() && lower[0] == '#') {
        std::string h = lower.substr(1);
        auto hv = [](char ch) { return std::isdigit((unsigned char)ch) ? ch - '0' : ch - 'a' + 10; };
        for (char ch : h)
            if (!std::isxdigit((unsigned char)ch)) return 2;
        if (h.size() == 3) {
            for (int i = 0; i < 3; i++) rgb[i] = (uint8_t)(hv(h[i]) * 17);
            return 0;
        }
        if (h.size() == 6) {
            for (int i = 0; i < 3; i++) rgb[i] = (uint8_t)(hv(h[2 * i]) * 16 + hv(h[2 * i + 1]));
            return 0;
        }
        return 2;
    }
    if (lower.compare(0, 4, "rgb(") == 0) {
        const char *p = lower.c_str() + 4;
        for (int i = 0; i < 3; i++) {
            while (*p && (std::isspace((unsigned char)*p) || *p == ',')) p++;
            char *end = nullptr;
            double x = std::strtod(p, &end);
            if (end == p) return 2;
            p = end;
            if (*p == '%') {
                x = x * 255.0 / 100.0;
                p++;
            }
            rgb[i] = (uint8_t)std::lround(std::min(255.0, std::max(0.0, x)));
        }
        return 0;
    }
    auto it = named_colors().find(lower);
    if (it == named_colors().end()) return 2;
    rgb[0] = (uint8_t)(it->second >> 16);
    rgb[1] = (uint8_t)(it->second >> 8);
    rgb[2] = (uint8_t)it->second;
    return 0;
}

bool parse_transform(const std::string &v, Xf *out) {
    Xf acc;
    const char *p = v.c_str();
    while (true) {
        while (*p && (std::isspace((unsigned char)*p) || *p == ',')) p++;
        if (!*p) break;
        const char *n0 = p;
        while (std::isalpha((unsigned char)*p)) p++;
        std::string name(n0, p);
        while (*p && std::isspace((unsigned char)*p)) p++;
        if (*p != '(') return false;
        p++;
        double a[6];
        int n = 0;
        PathParser P{p};
        while (n < 6 && P.number(&a[n])) n++;
        p = P.s;
        while (*p && *p != ')') p++;
        if (*p != ')') return false;
        p++;
        Xf t;
        if (name == "matrix" && n == 6) {
            t.a = a[0]; t.b = a[1]; t.c = a[2]; t.d = a[3]; t.e = a[4]; t.f = a[5];
        } else if (name == "translate" && n >= 1) {
            t.e = a[0];
            t.f = n > 1 ? a[1] : 0;
        } else if (name == "scale" && n >= 1) {
            t.a = a[0];
            t.d = n > 1 ? a[1] : a[0];
        } else if (name == "rotate" && n >= 1) {
            double r = a[0] * M_PI / 180.0, c = std::cos(r), s = std::sin(r);
            Xf rot;
            rot.a = c; rot.b = s; rot.c = -s; rot.d = c;
            if (n >= 3) {
                Xf t0, t1;
                t0.e = a[1]; t0.f = a[2];
                t1.e = -a[1]; t1.f = -a[2];
                t = t0.mul(rot).mul(t1);
            } else {
                t = rot;
            }
        } else if (name == "skewX" && n == 1) {
            t.c = std::tan(a[0] * M_PI / 180.0);
        } else if (name == "skewY" && n == 1) {
            t.b = std::tan(a[0] * M_PI / 180.0);
        } else {
            return false;
        }
        acc = acc.mul(t);
    }
    *out = acc;
    return true;
}

double parse_length(const std::string &v, double fallback) {
    const char *p = v.c_str();
    char *end = nullptr;
    double x = std::strtod(p, &end);
    return end == p ? fallback : x;
}

// ------------------------------------------------------------ XML

struct Element {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    bool self_closing = false, closing = false;
};

std::string decode_entities(const std::string &v) {
    std::string out;
    for (size_t i = 0; i < v.size(); i++) {
        if (v[i] == '&') {
            size_t j = v.find(';', i);
            if (j != std::string::npos) {
                std::string e = v.substr(i + 1, j - i - 1);
                if (e == "amp") { out += '&'; i = j; continue; }
                if (e == "lt") { out += '<'; i = j; continue; }
                if (e == "gt") { out += '>'; i = j; continue; }
                if (e == "quot") { out += '"'; i = j; continue; }
                if (e == "apos") { out += '\''; i = j; continue; }
            }
        }
        out += v[i];
    }
    return out;
}

// Reads the next tag at or after pos; false at the end or on a malformed tag.
bool next_tag(const std::string &x, size_t &pos, Element &el, bool &ok) {
    while (true) {
        size_t lt = x.find('<', pos);
        if (lt == std::string::npos) return false;
        if (x.compare(lt, 4, "<!--") == 0) {
            size_t e = x.find("-->", lt + 4);
            if (e == std::string::npos) { ok = false; return false; }
            pos = e + 3;
            continue;
        }
        if (x.compare(lt, 9, "<![CDATA[") == 0) {
            size_t e = x.find("]]>", lt);
            if (e == std::string::npos) { ok = false; return false; }
            pos = e + 3;
            continue;
        }
        if (x.compare(lt, 2, "<?") == 0 || x.compare(lt, 2, "<!") == 0) {
            size_t e = x.find('>', lt);
            if (e == std::string::npos) { ok = false; return false; }
            pos = e + 1;
            continue;
        }
        size_t p = lt + 1;
        el = Element();
        if (p < x.size() && x[p] == '/') {
            el.closing = true;
            p++;
        }
        size_t n0 = p;
        while (p < x.size() && !std::isspace((unsigned char)x[p]) && x[p] != '>' && x[p] != '/')
            p++;
        el.name = x.substr(n0, p - n0);
        while (true) {
            while (p < x.size() && std::isspace((unsigned char)x[p])) p++;
            if (p >= x.size()) { ok = false; return false; }
            if (x[p] == '>') { pos = p + 1; return true; }
            if (x[p] == '/' && p + 1 < x.size() && x[p + 1] == '>') {
                el.self_closing = true;
                pos = p + 2;
                return true;
            }
            size_t a0 = p;
            while (p < x.size() && x[p] != '=' && !std::isspace((unsigned char)x[p]) && x[p] != '>')
                p++;
            std::string an = x.substr(a0, p - a0);
            while (p < x.size() && std::isspace((unsigned char)x[p])) p++;
            if (p >= x.size() || x[p] != '=') { ok = false; return false; }
            p++;
            while (p < x.size() && std::isspace((unsigned char)x[p])) p++;
            if (p >= x.size() || (x[p] != '"' && x[p] != '\'')) { ok = false; return false; }
            char q = x[p++];
            size_t v0 = p;
            while (p < x.size() && x[p] != q) p++;
            if (p >= x.size()) { ok = false; return false; }
            el.attrs.push_back({an, decode_entities(x.substr(v0, p - v0))});
            p++;
        }
    }
}

void apply_property(Style &st, const std::string &name, const std::string &value, bool *display_none) {
    std::string v = trim(value);
    if (name == "fill") {
        int r = parse_color(v, st.fill);
        st.fill_none = r != 0;
        st.unsupported_paint = st.unsupported_paint || r == 2;
    } else if (name == "stroke") {
        int r = parse_color(v, st.stroke);
        st.stroke_none = r != 0;
        st.unsupported_paint = st.unsupported_paint || r == 2;
    } else if (name == "fill-opacity") {
        st.fill_opacity = std::min(1.0, std::max(0.0, parse_length(v, 1)));
    } else if (name == "stroke-opacity") {
        st.stroke_opacity = std::min(1.0, std::max(0.0, parse_length(v, 1)));
    } else if (name == "fill-rule") {
        st.fill_rule = v == "evenodd" ? PFH_FILL_RULE_EVEN_ODD : PFH_FILL_RULE_WINDING;
    } else if (name == "stroke-width") {
        st.stroke_width = parse_length(v, 1);
    } else if (name == "stroke-miterlimit") {
        st.miter_limit = std::max(1.0, parse_length(v, 4));
    } else if (name == "stroke-linecap") {
        st.cap = v == "square" ? 1 : v == "round" ? 2 : 0;
    } else if (name == "stroke-linejoin") {
        st.join = v == "bevel" ? 1 : v == "round" ? 2 : 0;
    } else if (name == "display") {
        if (v == "none" && display_none) *display_none = true;
    }
}

double attr_num(const Element &el, const char *name, double fallback) {
    for (auto &kv : el.attrs)
        if (kv.first == name) return parse_length(kv.second, fallback);
    return fallback;
}

const std::string *attr(const Element &el, const char *name) {
    for (auto &kv : el.attrs)
        if (kv.first == name) return &kv.second;
    return nullptr;
}

// The subpaths of a basic shape element (SVG 1.1 section 9), or of a path.
bool element_subpaths(const Element &el, std::vector<Subpath> &subs) {
    const std::string &n = el.name;
    if (n == "path") {
        const std::string *d = attr(el, "d");
        return d && parse_path_data(d->c_str(), subs);
    }
    if (n == "rect") {
        double x = attr_num(el, "x", 0), y = attr_num(el, "y", 0);
        double w = attr_num(el, "width", 0), h = attr_num(el, "height", 0);
        if (w <= 0 || h <= 0) return true;
        double rx = attr_num(el, "rx", -1), ry = attr_num(el, "ry", -1);
        if (rx < 0) rx = ry;
        if (ry < 0) ry = rx;
        rx = std::min(std::max(rx, 0.0), w / 2);
        ry = std::min(std::max(ry, 0.0), h / 2);
        Subpath sp;
        sp.closed = true;
        if (rx == 0 || ry == 0) {
            sp.start = {x, y};
            sp.segs = {{false, {}, {}, {x + w, y}}, {false, {}, {}, {x + w, y + h}},
                       {false, {}, {}, {x, y + h}}};
        } else {
            sp.start = {x + rx, y};
            V2 cur = sp.start;
            auto line = [&](V2 to) { sp.segs.push_back({false, {}, {}, to}); cur = to; };
            auto arc = [&](V2 to) { arc_to_cubics(cur, rx, ry, 0, false, true, to, sp.segs); cur = to; };
            line({x + w - rx, y});
            arc({x + w, y + ry});
            line({x + w, y + h - ry});
            arc({x + w - rx, y + h});
            line({x + rx, y + h});
            arc({x, y + h - ry});
            line({x, y + ry});
            arc({x + rx, y});
        }
        subs.push_back(sp);
        return true;
    }
    if (n == "circle" || n == "ellipse") {
        double cx = attr_num(el, "cx", 0), cy = attr_num(el, "cy", 0);
        double rx = n == "circle" ? attr_num(el, "r", 0) : attr_num(el, "rx", 0);
        double ry = n == "circle" ? rx : attr_num(el, "ry", 0);
        if (rx <= 0 || ry <= 0) return true;
        Subpath sp;
        sp.closed = true;
        sp.start = {cx + rx, cy};
        arc_to_cubics(sp.start, rx, ry, 0, false, true, {cx - rx, cy}, sp.segs);
        arc_to_cubics({cx - rx, cy}, rx, ry, 0, false, true, sp.start, sp.segs);
        subs.push_back(sp);
        return true;
    }
    if (n == "line") {
        Subpath sp;
        sp.start = {attr_num(el, "x1", 0), attr_num(el, "y1", 0)};
        sp.segs.push_back({false, {}, {}, {attr_num(el, "x2", 0), attr_num(el, "y2", 0)}});
        subs.push_back(sp);
        return true;
    }
    if (n == "polyline" || n == "polygon") {
        const std::string *pts = attr(el, "points");
        if (!pts) return true;
        PathParser P{pts->c_str()};
        double a, b;
        Subpath sp;
        bool first = true;
        while (P.number(&a) && P.number(&b)) {
            if (first) sp.start = {a, b};
            else sp.segs.push_back({false, {}, {}, {a, b}});
            first = false;
        }
        sp.closed = n == "polygon";
        if (!first) subs.push_back(sp);
        return true;
    }
    return true;
}

StrokeParams stroke_params(const Style &st) {
    StrokeParams p;
    p.stroke_width = st.stroke_width;
    p.miter_limit = st.miter_limit;
    p.cap = st.cap;
    p.join = st.join;
    return p;
}

uint8_t unit_to_u8(double x) { return (uint8_t)std::lround(std::min(1.0, std::max(0.0, x)) * 255.0); }

}  // namespace

extern "C" pfh_scene *pfh_svg_scene_create(const char *bytes, size_t byte_len, pfh_svg_info *info) {
    if (!bytes) return nullptr;
    std::string x(bytes, byte_len);
    pfh_scene *scene = pfh_scene_create();
    if (!scene) return nullptr;
    pfh_svg_info inf;
    std::memset(&inf, 0, sizeof(inf));
    std::vector<Style> stack{Style()};
    size_t pos = 0;
    bool ok = true, have_root = false;
    Element el;
    int depth = 0, skip_depth = -1;
    while (next_tag(x, pos, el, ok)) {
        if (el.closing) {
            depth--;
            if (skip_depth >= 0) {  // inside a skipped subtree: nothing was pushed
                if (depth == skip_depth) skip_depth = -1;
                continue;
            }
            if (stack.size() > 1) stack.pop_back();
            continue;
        }
        bool opens = !el.self_closing;
        if (skip_depth >= 0) {
            if (opens) depth++;
            continue;
        }
        Style st = stack.back();
        bool display_none = false;
        for (auto &kv : el.attrs) {
            if (kv.first == "transform") {
                Xf t;
                if (parse_transform(kv.second, &t)) st.xf = st.xf.mul(t);
                else inf.skipped_elements++;
            } else if (kv.first == "style") {
                std::string s = kv.second;
                size_t a = 0;
                while (a < s.size()) {
                    size_t semi = s.find(';', a);
                    std::string decl = s.substr(a, semi == std::string::npos ? std::string::npos
                                                                             : semi - a);
                    size_t colon = decl.find(':');
                    if (colon != std::string::npos)
                        apply_property(st, trim(decl.substr(0, colon)), decl.substr(colon + 1),
                                       &display_none);
                    if (semi == std::string::npos) break;
                    a = semi + 1;
                }
            } else {
                apply_property(st, kv.first, kv.second, &display_none);
            }
        }
        // style declarations win over presentation attributes
        const std::string &n = el.name;
        bool container = n == "svg" || n == "g";
        bool shape = n == "path" || n == "rect" || n == "circle" || n == "ellipse" ||
                     n == "line" || n == "polyline" || n == "polygon";
        if (n == "svg" && !have_root) {
            have_root = true;
            float vb[4] = {0, 0, (float)attr_num(el, "width", 0), (float)attr_num(el, "height", 0)};
            if (const std::string *v = attr(el, "viewBox")) {
                PathParser P{v->c_str()};
                double q[4];
                if (P.number(&q[0]) && P.number(&q[1]) && P.number(&q[2]) && P.number(&q[3]))
                    for (int i = 0; i < 4; i++) vb[i] = (float)q[i];
            }
            pfh_rect_f r{{vb[0], vb[1]}, {vb[0] + vb[2], vb[1] + vb[3]}};
            pfh_scene_set_view_box(scene, &r);
            for (int i = 0; i < 4; i++) inf.view_box[i] = vb[i];
        }
        if (display_none || (!container && !shape)) {
            if (!shape && !container) inf.skipped_elements++;
            if (opens) {
                skip_depth = depth;
                depth++;
            }
            continue;
        }
        if (shape) {
            std::vector<Subpath> subs;
            if (!element_subpaths(el, subs)) {
                ok = false;
                break;
            }
            if (st.unsupported_paint) inf.unsupported_paints++;
            if (!st.fill_none && !subs.empty()) {
                OutlineBuf o;
                fill_outline(subs, st.xf, o);
                if (!o.ends.empty()) {
                    pfh_color_u c{st.fill[0], st.fill[1], st.fill[2], unit_to_u8(st.fill_opacity)};
                    uint32_t paint = pfh_scene_push_paint(scene, &c);
                    pfh_outline ol{o.pts.data(), o.flags.data(), (uint32_t)o.pts.size(),
                                   o.ends.data(), (uint32_t)o.ends.size()};
                    pfh_scene_push_draw_path(scene, &ol, paint, st.fill_rule, -1,
                                             PFH_BLEND_MODE_SRC_OVER);
                    inf.fills++;
                }
            }
            if (!st.stroke_none && !subs.empty()) {
                OutlineBuf o;
                stroke_outline(subs, stroke_params(st), st.xf, o);
                if (!o.ends.empty()) {
                    pfh_color_u c{st.stroke[0], st.stroke[1], st.stroke[2],
                                  unit_to_u8(st.stroke_opacity)};
                    uint32_t paint = pfh_scene_push_paint(scene, &c);
                    pfh_outline ol{o.pts.data(), o.flags.data(), (uint32_t)o.pts.size(),
                                   o.ends.data(), (uint32_t)o.ends.size()};
                    pfh_scene_push_draw_path(scene, &ol, paint, PFH_FILL_RULE_WINDING, -1,
                                             PFH_BLEND_MODE_SRC_OVER);
                    inf.strokes++;
                }
            }
        }
        if (opens) {
            depth++;
            stack.push_back(st);
        }
    }
    if (!ok || !have_root) {
        pfh_scene_destroy(scene);
        return nullptr;
    }
    if (info) *info = inf;
    return scene;
}

extern "C" pfh_status pfh_stroke_outline(const pfh_outline *outline, const uint8_t *closed,
                                         const pfh_stroke_style *style,
                                         const pfh_transform2f *transform,
                                         pfh_outline_buffer **out) {
    if (!outline || !style || !out) return PFH_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    std::vector<Subpath> subs;
    uint32_t start = 0;
    for (uint32_t c = 0; c < outline->contour_count; c++) {
        uint32_t end = std::min(outline->contour_ends[c], outline->point_count);
        if (end <= start) continue;
        Subpath sp;
        auto P = [&](uint32_t i) {
            return V2{outline->points[i].x, outline->points[i].y};
        };
        sp.start = P(start);
        sp.closed = closed ? closed[c] != 0 : true;
        uint32_t i = start + 1;
        while (i < end) {
            uint8_t f = outline->flags[i];
            if ((f & PFH_POINT_CONTROL_0) && i + 2 < end + 1 && i + 1 < end &&
                (outline->flags[i + 1] & PFH_POINT_CONTROL_1) && i + 2 < end) {
                sp.segs.push_back({true, P(i), P(i + 1), P(i + 2)});
                i += 3;
            } else if ((f & PFH_POINT_CONTROL_0) && i + 1 < end) {
                V2 cur = sp.segs.empty() ? sp.start : sp.segs.back().to;
                V2 q = P(i), to = P(i + 1);
                sp.segs.push_back({true, cur + (q - cur) * (2.0 / 3.0), to + (q - to) * (2.0 / 3.0),
                                   to});
                i += 2;
            } else {
                sp.segs.push_back({false, {}, {}, P(i)});
                i += 1;
            }
        }
        subs.push_back(sp);
        start = end;
    }
    StrokeParams st;
    st.stroke_width = style->line_width;
    st.cap = style->line_cap;
    st.join = style->line_join;
    st.miter_limit = style->miter_limit > 0 ? style->miter_limit : 4.0;
    Xf xf;
    if (transform) {
        xf.a = transform->m00; xf.c = transform->m01; xf.e = transform->tx;
        xf.b = transform->m10; xf.d = transform->m11; xf.f = transform->ty;
    }
    OutlineBuf o;
    stroke_outline(subs, st, xf, o);
    pfh_outline_buffer *buf = new (std::nothrow) pfh_outline_buffer();
    if (!buf) return PFH_ERROR_OUT_OF_MEMORY;
    buf->points = std::move(o.pts);
    buf->flags = std::move(o.flags);
    buf->contour_ends = std::move(o.ends);
    buf->view.points = buf->points.data();
    buf->view.flags = buf->flags.data();
    buf->view.point_count = (uint32_t)buf->points.size();
    buf->view.contour_ends = buf->contour_ends.data();
    buf->view.contour_count = (uint32_t)buf->contour_ends.size();
    *out = buf;
    return PFH_OK;
}

extern "C" const pfh_outline *pfh_outline_buffer_view(const pfh_outline_buffer *buf) {
    return buf ? &buf->view : nullptr;
}

extern "C" void pfh_outline_buffer_destroy(pfh_outline_buffer *buf) { delete buf; }
