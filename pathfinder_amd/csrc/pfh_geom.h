// pfh_geom.h - plane geometry shared by the SVG reader and the canvas layer:
// subpaths of lines and cubics, elliptical arcs as cubics, fill outlines and
// the stroker (OutlineStrokeToFill's role, content/src/stroke.rs).
#ifndef PFH_GEOM_H
#define PFH_GEOM_H

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/pfhip.h"

namespace pfh_geom {

struct V2 {
    double x, y;
};
inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline V2 operator*(V2 a, double s) { return {a.x * s, a.y * s}; }
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline double cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
inline double len(V2 a) { return std::sqrt(dot(a, a)); }

// Affine transform, SVG order: (a c e; b d f).
struct Xf {
    double a = 1, b = 0, c = 0, d = 1, e = 0, f = 0;
    V2 apply(V2 p) const { return {a * p.x + c * p.y + e, b * p.x + d * p.y + f}; }
    Xf mul(const Xf &o) const {  // this * o
        Xf r;
        r.a = a * o.a + c * o.b;
        r.b = b * o.a + d * o.b;
        r.c = a * o.c + c * o.d;
        r.d = b * o.c + d * o.d;
        r.e = a * o.e + c * o.f + e;
        r.f = b * o.e + d * o.f + f;
        return r;
    }
    double scale() const { return std::sqrt(std::fabs(a * d - b * c)); }
};

// One subpath: its start point and segments (lines, or cubics with two
// control points), closed by Z or not.
struct Seg {
    bool cubic;
    V2 c0, c1, to;
};
struct Subpath {
    V2 start;
    std::vector<Seg> segs;
    bool closed = false;
};

// Stroke parameters: line width, cap (0 butt, 1 square, 2 round), join (0
// miter, 1 bevel, 2 round) and the miter limit (SVG: miter length / width).
struct StrokeParams {
    double stroke_width = 1, miter_limit = 4;
    int cap = 0, join = 0;
};

// Appends the cubics of an SVG elliptical arc (implementation notes F.6.5).
inline void arc_to_cubics(V2 p0, double rx, double ry, double phi_deg, bool large, bool sweep, V2 p1,
                   std::vector<Seg> &out) {
    if (p0.x == p1.x && p0.y == p1.y) return;
    rx = std::fabs(rx);
    ry = std::fabs(ry);
    if (rx == 0 || ry == 0) {
        out.push_back({false, {0, 0}, {0, 0}, p1});
        return;
    }
    double phi = phi_deg * M_PI / 180.0, cp = std::cos(phi), sp = std::sin(phi);
    double dx = (p0.x - p1.x) / 2, dy = (p0.y - p1.y) / 2;
    double x1 = cp * dx + sp * dy, y1 = -sp * dx + cp * dy;
    double lam = x1 * x1 / (rx * rx) + y1 * y1 / (ry * ry);
    if (lam > 1) {
        double r = std::sqrt(lam);
        rx *= r;
        ry *= r;
    }
    double num = rx * rx * ry * ry - rx * rx * y1 * y1 - ry * ry * x1 * x1;
    double den = rx * rx * y1 * y1 + ry * ry * x1 * x1;
    double k = den > 0 ? std::sqrt(std::max(0.0, num / den)) : 0.0;
    if (large == sweep) k = -k;
    double cxp = k * rx * y1 / ry, cyp = -k * ry * x1 / rx;
    double cx = cp * cxp - sp * cyp + (p0.x + p1.x) / 2;
    double cy = sp * cxp + cp * cyp + (p0.y + p1.y) / 2;
    auto angle = [](double ux, double uy, double vx, double vy) {
        return std::atan2(ux * vy - uy * vx, ux * vx + uy * vy);
    };
    double t1 = angle(1, 0, (x1 - cxp) / rx, (y1 - cyp) / ry);
    double dt = angle((x1 - cxp) / rx, (y1 - cyp) / ry, (-x1 - cxp) / rx, (-y1 - cyp) / ry);
    if (!sweep && dt > 0) dt -= 2 * M_PI;
    if (sweep && dt < 0) dt += 2 * M_PI;
    int n = std::max(1, (int)std::ceil(std::fabs(dt) / (M_PI / 2) - 1e-9));
    double step = dt / n, kk = 4.0 / 3.0 * std::tan(step / 4);
    auto pt = [&](double t) {
        double x = rx * std::cos(t), y = ry * std::sin(t);
        return V2{cp * x - sp * y + cx, sp * x + cp * y + cy};
    };
    auto der = [&](double t) {
        double x = -rx * std::sin(t), y = ry * std::cos(t);
        return V2{cp * x - sp * y, sp * x + cp * y};
    };
    for (int i = 0; i < n; i++) {
        double ta = t1 + step * i, tb = ta + step;
        V2 a = pt(ta), b = i + 1 == n ? p1 : pt(tb);
        out.push_back({true, a + der(ta) * kk, b - der(tb) * kk, b});
    }
}

// ------------------------------------------------------------ outlines

// A scene outline being built: points, PointFlags, contour ends.
struct OutlineBuf {
    std::vector<pfh_vector2f> pts;
    std::vector<uint8_t> flags;
    std::vector<uint32_t> ends;
    void point(V2 p, uint8_t f) {
        pts.push_back({(float)p.x, (float)p.y});
        flags.push_back(f);
    }
    void end_contour() {
        if (pts.size() && (ends.empty() || ends.back() != pts.size()))
            ends.push_back((uint32_t)pts.size());
    }
    // A polygon contour, counter-clockwise on screen (positive area in y-down coordinates).
    void polygon(std::vector<V2> poly) {
        if (poly.size() < 3) return;
        double area = 0;
        for (size_t i = 0; i < poly.size(); i++)
            area += cross(poly[i], poly[(i + 1) % poly.size()]);
        if (area == 0) return;
        if (area < 0) std::reverse(poly.begin(), poly.end());
        for (const V2 &p : poly) point(p, 0);
        end_contour();
    }
};

inline void fill_outline(const std::vector<Subpath> &subs, const Xf &xf, OutlineBuf &o) {
    for (const Subpath &sp : subs) {
        if (sp.segs.empty()) continue;
        o.point(xf.apply(sp.start), 0);
        for (const Seg &s : sp.segs) {
            if (s.cubic) {
                o.point(xf.apply(s.c0), PFH_POINT_CONTROL_0);
                o.point(xf.apply(s.c1), PFH_POINT_CONTROL_1);
            }
            o.point(xf.apply(s.to), 0);
        }
        o.end_contour();
    }
}

// The subpath as a polyline in user space (cubics flattened at `tol`).
inline std::vector<V2> flatten(const Subpath &sp, double tol) {
    std::vector<V2> out{sp.start};
    V2 cur = sp.start;
    for (const Seg &s : sp.segs) {
        if (s.cubic) {
            V2 a = s.c1 - s.c0 * 2.0 + cur, b = s.to - s.c1 * 2.0 + s.c0;
            double dd = 6.0 * std::max(len(a), len(b));
            int n = std::max(1, (int)std::ceil(std::sqrt(dd / (8.0 * tol))));
            n = std::min(n, 4096);
            for (int i = 1; i <= n; i++) {
                double t = (double)i / n, u = 1 - t;
                out.push_back(cur * (u * u * u) + s.c0 * (3 * u * u * t) + s.c1 * (3 * u * t * t) +
                              s.to * (t * t * t));
            }
        } else {
            out.push_back(s.to);
        }
        cur = s.to;
    }
    // drop repeated points
    std::vector<V2> dedup;
    for (const V2 &p : out)
        if (dedup.empty() || len(p - dedup.back()) > 1e-9) dedup.push_back(p);
    return dedup;
}

// Round fan from `p + a` to `p + b` (both of length w) around p, on the
// short side, with chord error <= tol.
inline void round_fan(V2 p, V2 a, V2 b, double w, double tol, std::vector<V2> &poly) {
    double ang = std::atan2(cross(a, b), dot(a, b));
    double step = 2.0 * std::acos(std::max(-1.0, std::min(1.0, 1.0 - tol / std::max(w, 1e-12))));
    int n = std::max(1, (int)std::ceil(std::fabs(ang) / std::max(step, 1e-3)));
    n = std::min(n, 256);
    double a0 = std::atan2(a.y, a.x);
    poly.push_back(p);
    for (int i = 0; i <= n; i++) {
        double t = a0 + ang * i / n;
        poly.push_back(p + V2{std::cos(t), std::sin(t)} * w);
    }
}

inline void stroke_outline(const std::vector<Subpath> &subs, const StrokeParams &st, const Xf &xf,
                    OutlineBuf &o) {
    double w = std::max(st.stroke_width, 0.0333) * 0.5;
    double tol = 0.1 / std::max(xf.scale(), 1e-6);
    auto emit = [&](std::vector<V2> poly) {
        for (V2 &p : poly) p = xf.apply(p);
        o.polygon(std::move(poly));
    };
    for (const Subpath &sp : subs) {
        std::vector<V2> pl = flatten(sp, tol);
        bool closed = sp.closed && pl.size() > 2;
        if (closed && len(pl.front() - pl.back()) <= 1e-9) pl.pop_back();
        if (pl.size() == 1) {
            // a zero-length subpath shows its caps only
            if (st.cap == 1) {
                V2 p = pl[0];
                emit({p + V2{-w, -w}, p + V2{w, -w}, p + V2{w, w}, p + V2{-w, w}});
            } else if (st.cap == 2) {
                std::vector<V2> fan;
                round_fan(pl[0], V2{w, 0}, V2{-w, 0}, w, tol, fan);
                emit(fan);
                fan.clear();
                round_fan(pl[0], V2{-w, 0}, V2{w, 0}, w, tol, fan);
                emit(fan);
            }
            continue;
        }
        size_t n = pl.size(), nseg = closed ? n : n - 1;
        auto dir = [&](size_t i) {  // unit direction of segment i
            V2 d = pl[(i + 1) % n] - pl[i];
            return d * (1.0 / len(d));
        };
        for (size_t i = 0; i < nseg; i++) {
            V2 a = pl[i], b = pl[(i + 1) % n], d = dir(i), nrm = V2{-d.y, d.x} * w;
            emit({a + nrm, b + nrm, b - nrm, a - nrm});
        }
        // joins at the interior vertices (every vertex of a closed subpath)
        for (size_t k = closed ? 0 : 1; k < (closed ? n : n - 1); k++) {
            size_t i0 = (k + nseg - 1) % nseg, i1 = k % nseg;
            V2 d0 = dir(i0), d1 = dir(i1), p = pl[k];
            double cr = cross(d0, d1);
            if (std::fabs(cr) < 1e-12 && dot(d0, d1) > 0) continue;
            double s = cr > 0 ? -1.0 : 1.0;  // the outer side
            V2 n0 = V2{-d0.y, d0.x} * (w * s), n1 = V2{-d1.y, d1.x} * (w * s);
            if (st.join == 2) {
                std::vector<V2> fan;
                round_fan(p, n0, n1, w, tol, fan);
                emit(fan);
                continue;
            }
            if (st.join == 0) {
                V2 u0 = n0 * (1.0 / w), u1 = n1 * (1.0 / w);
                double c = 1.0 + dot(u0, u1);
                if (c > 1e-12) {
                    V2 tip = p + (u0 + u1) * (w / c);
                    // SVG: miter length / stroke width <= miter limit
                    if (len(tip - p) * 2.0 <= st.miter_limit * 2.0 * w) {
                        emit({p, p + n0, tip, p + n1});
                        continue;
                    }
                }
            }
            emit({p, p + n0, p + n1});
        }
        if (!closed) {
            for (int end = 0; end < 2; end++) {
                V2 p = end ? pl[n - 1] : pl[0];
                V2 d = end ? dir(n - 2) : dir(0) * -1.0;  // outward
                V2 nrm = V2{-d.y, d.x} * w;
                if (st.cap == 1) {
                    emit({p + nrm, p + nrm + d * w, p - nrm + d * w, p - nrm});
                } else if (st.cap == 2) {
                    std::vector<V2> fan;
                    // the half disc beyond the end: from +nrm through +d to -nrm
                    round_fan(p, nrm, d * w, w, tol, fan);
                    std::vector<V2> fan2;
                    round_fan(p, d * w, nrm * -1.0, w, tol, fan2);
                    emit(fan);
                    emit(fan2);
                }
            }
        }
    }
}

}  // namespace pfh_geom

#endif
