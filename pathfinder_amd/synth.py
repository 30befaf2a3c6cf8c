"""Seeded synthetic scenes for the benchmark workloads and tests.

Each generator returns a Scene plus its framebuffer size and background, built
from numpy's PCG64 generator seeded with 0x5EED0000 + workload index, so every
run (and the oracle) sees the same inputs.

  small   1024x1024, 256 closed paths of 8-32 cubics, control points uniform
          in [0,1024)^2, nonzero, random sRGB, 50% opaque (else alpha
          U[0.2, 0.9]).
  4k      3840x2160, 20,000 star-shaped paths (log-uniform 10-600 mixed
          quadratic/cubic segments, log-uniform sizes 4-2000 px), 60% opaque,
          20% even-odd, 500 rectangle/ellipse clip paths; about 5M microlines.
  glyphs  50,000 glyphs (1-4 quadratic contours with holes) at 8-72 px, per
          glyph 2x3 affine, for atlas pages (pathfinder_amd.batch).
  canvases 512 canvases of log-uniform 64^2-1024^2, 10-2000 small-scene paths
          each (pathfinder_amd.batch).
  diff    1024 closed cubic paths on 512x512 with colours and a perturbed
          target, the differentiable-render workload.
"""
import math

import numpy as np

from . import _lib as L
from .scene import ColorF, PathArrays, RectF, Scene

SEED_BASE = 0x5EED0000
KAPPA = 0.5522847498


def _colors(rng, n, opaque_frac, scene):
    rgb = rng.integers(0, 256, size=(n, 3))
    opaque = rng.random(n) < opaque_frac
    alpha = np.where(opaque, 255, np.round(rng.uniform(0.2, 0.9, n) * 255)).astype(np.int64)
    ids = np.zeros(n, np.uint32)
    from .scene import ColorU
    for i in range(n):
        ids[i] = scene.push_paint(ColorU(rgb[i, 0], rgb[i, 1], rgb[i, 2], alpha[i]))
    return ids


def small_scene(seed=SEED_BASE + 0, paths=256, size=1024):
    """256 random closed cubic paths on a 1024x1024 canvas."""
    rng = np.random.default_rng(seed)
    scene = Scene()
    scene.set_view_box(RectF(0, 0, size, size))
    paints = _colors(rng, paths, 0.5, scene)
    pts, flags, ends, pends = [], [], [], []
    npts = 0
    for p in range(paths):
        k = int(rng.integers(8, 33))
        start = rng.uniform(0, size, 2)
        ctrl = rng.uniform(0, size, (k, 3, 2))
        ctrl[-1, 2] = start  # the last cubic returns to the start point
        path_pts = [start]
        path_fl = [0]
        for j in range(k):
            path_pts += [ctrl[j, 0], ctrl[j, 1], ctrl[j, 2]]
            path_fl += [L.POINT_CONTROL_0, L.POINT_CONTROL_1, 0]
        pts.append(np.asarray(path_pts, np.float32))
        flags.append(np.asarray(path_fl, np.uint8))
        npts += len(path_pts)
        ends.append(npts)
        pends.append(len(ends))
    arr = PathArrays(np.concatenate(pts), np.concatenate(flags), np.asarray(ends, np.uint32),
                     np.asarray(pends, np.uint32), paints)
    scene.push_draw_paths(arr)
    return scene, (size, size), ColorF(1.0, 1.0, 1.0, 1.0)


def diff_paths(seed=SEED_BASE + 4, paths=1024, size=512, sigma=4.0):
    """The differentiable-render workload: `paths` closed paths of 8-32 cubic
    Beziers with control points uniform on the size x size canvas (the small
    scene's distribution on this canvas), sRGB colours (half opaque, the rest
    alpha U[0.2, 0.9]) as floats in [0, 1], and target points: a copy with
    every point moved by N(0, sigma) px. Returns a dict of arrays: flags,
    contour_ends, path_contour_ends, points (N, 2), colors (P, 4),
    target_points."""
    rng = np.random.default_rng(seed)
    pts, flags, ends, pends = [], [], [], []
    npts = 0
    for p in range(paths):
        k = int(rng.integers(8, 33))
        start = rng.uniform(0, size, 2)
        ctrl = rng.uniform(0, size, (k, 3, 2))
        ctrl[-1, 2] = start
        path_pts = [start]
        path_fl = [0]
        for j in range(k):
            path_pts += [ctrl[j, 0], ctrl[j, 1], ctrl[j, 2]]
            path_fl += [L.POINT_CONTROL_0, L.POINT_CONTROL_1, 0]
        pts.append(np.asarray(path_pts, np.float32))
        flags.append(np.asarray(path_fl, np.uint8))
        npts += len(path_pts)
        ends.append(npts)
        pends.append(len(ends))
    rgb = rng.integers(0, 256, (paths, 3)) / 255.0
    alpha = np.where(rng.random(paths) < 0.5, 1.0, rng.uniform(0.2, 0.9, paths))
    points = np.concatenate(pts)
    return {
        "flags": np.concatenate(flags),
        "contour_ends": np.asarray(ends, np.uint32),
        "path_contour_ends": np.asarray(pends, np.uint32),
        "points": points,
        "colors": np.concatenate([rgb, alpha[:, None]], axis=1).astype(np.float32),
        "target_points": (points + rng.normal(0.0, sigma, points.shape)).astype(np.float32),
        "size": (size, size),
    }


def _cubic_paths(rng, n, size):
    """n closed paths of 8-32 cubics with control points uniform in
    [0, size)^2, the last cubic returning to the start (as small_scene),
    in flat PathArrays form without paints."""
    k = rng.integers(8, 33, n)
    sizes = 1 + 3 * k
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sizes.sum())
    pts = rng.uniform(0, size, (total, 2)).astype(np.float32)
    pts[starts + sizes - 1] = pts[starts]
    local = np.arange(total) - np.repeat(starts, sizes)
    pattern = np.asarray([L.POINT_CONTROL_0, L.POINT_CONTROL_1, 0], np.uint8)
    flags = np.where(local == 0, 0, pattern[(local - 1) % 3]).astype(np.uint8)
    ends = np.cumsum(sizes).astype(np.uint32)
    return PathArrays(pts, flags, ends, np.arange(1, n + 1, dtype=np.uint32))


def _rgba(rng, n):
    rgb = rng.integers(0, 256, size=(n, 3))
    alpha = np.where(rng.random(n) < 0.5, 255, np.round(rng.uniform(0.2, 0.9, n) * 255))
    return [(int(r), int(g), int(b), int(a)) for (r, g, b), a in zip(rgb, alpha.astype(int))]


def canvases(seed=SEED_BASE + 3, count=512, size_lo=64, size_hi=1024, paths_lo=10,
             paths_hi=2000):
    """The multi-scene workload: `count` square canvases, side log-uniform in
    [size_lo, size_hi], each with log-uniform [paths_lo, paths_hi] closed paths
    of the small scene's distribution on that canvas. Returns batch.Items."""
    from .batch import Item
    rng = np.random.default_rng(seed)
    items = []
    for c in range(count):
        side = int(round(math.exp(rng.uniform(math.log(size_lo), math.log(size_hi)))))
        n = int(round(math.exp(rng.uniform(math.log(paths_lo), math.log(paths_hi)))))
        items.append(Item(_cubic_paths(rng, n, side), _rgba(rng, n), side, side))
    return items


def _glyph_contour(rng, cx, cy, radius, m, reverse):
    """A closed star-shaped contour of m quadratics around (cx, cy)."""
    ang = np.sort(rng.uniform(0, 2 * np.pi, m))
    if reverse:
        ang = ang[::-1]
    r = radius * rng.uniform(0.8, 1.0, m)
    on = np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], axis=1)
    nxt = np.roll(on, -1, axis=0)
    mid = 0.5 * (on + nxt)
    ctrl = np.asarray([cx, cy]) + (mid - np.asarray([cx, cy])) * rng.uniform(0.95, 1.15, (m, 1))
    pts = np.empty((2 * m, 2))
    pts[0::2] = on
    pts[1::2] = ctrl
    flags = np.zeros(2 * m, np.uint8)
    flags[1::2] = L.POINT_CONTROL_0
    return pts, flags


def glyphs(seed=SEED_BASE + 2, count=50000, em=2048.0, size_lo=8, size_hi=72):
    """The glyph workload: `count` synthetic glyphs, each one path of 1-4
    closed contours of 10-40 quadratics in a 2048-unit em box (an outer
    contour and reversed holes inside it), drawn at a size uniform in
    [size_lo, size_hi] px through a per-glyph 2x3 affine (scale, rotation and
    shear within +-0.2) into a cell around its transformed bounds. White,
    opaque. Returns batch.Items."""
    from .batch import Item
    rng = np.random.default_rng(seed)
    items = []
    for g in range(count):
        size = int(rng.integers(size_lo, size_hi + 1))
        nc = int(rng.integers(1, 5))
        c = em / 2
        outer = em * 0.42
        pts, flags, ends = [], [], []
        p, f = _glyph_contour(rng, c, c, outer, int(rng.integers(10, 41)), False)
        pts.append(p)
        flags.append(f)
        ends.append(len(p))
        for h in range(nc - 1):
            a = rng.uniform(0, 2 * np.pi)
            d = rng.uniform(0, 0.35) * outer
            p, f = _glyph_contour(rng, c + d * np.cos(a), c + d * np.sin(a),
                                  rng.uniform(0.12, 0.22) * outer, int(rng.integers(10, 41)),
                                  True)
            pts.append(p)
            flags.append(f)
            ends.append(ends[-1] + len(p))
        pts = np.concatenate(pts)
        sc = size / em
        rot, shear = rng.uniform(-0.2, 0.2, 2)
        m = np.asarray([[math.cos(rot), -math.sin(rot)], [math.sin(rot), math.cos(rot)]]) @ \
            np.asarray([[sc, shear * sc], [0.0, sc]])
        q = pts @ m.T
        lo = q.min(axis=0)
        q = q - lo + 1.0
        w, h = (int(math.ceil(v)) + 1 for v in q.max(axis=0))
        arr = PathArrays(q.astype(np.float32), np.concatenate(flags), np.asarray(ends, np.uint32),
                         np.asarray([len(ends)], np.uint32))
        items.append(Item(arr, [(255, 255, 255, 255)], w, h))
    return items


def _blobs(rng, cx, cy, radius, k, quad_frac, jitter, ctrl_jitter):
    """Closed star-shaped contours, one per path, path i with k[i] mixed
    quadratic/cubic segments. Returns (points, flags, contour_ends)."""
    k = np.asarray(k, np.int64)
    total = int(k.sum())
    path = np.repeat(np.arange(len(k)), k)
    first = np.concatenate([[0], np.cumsum(k)[:-1]])
    j = np.arange(total) - first[path]
    kk = k[path].astype(np.float64)
    theta = (j + rng.uniform(0.0, 0.8, total)) * (2.0 * math.pi) / kk
    last = j == k[path] - 1
    th1 = np.where(last, theta[first[path]] + 2.0 * math.pi, np.roll(theta, -1))
    rad = radius[path]
    r_on = rad * rng.uniform(1.0 - jitter, 1.0, total)
    r_next = np.where(last, r_on[first[path]], np.roll(r_on, -1))
    quad = rng.random(total) < quad_frac
    # control points follow the arc between the two on-curve radii, so the
    # curvature (and the dice count) comes from the path size and segment count
    lerp = lambda f: r_on + f * (r_next - r_on)
    r_c = np.stack([np.where(quad, lerp(0.5), lerp(1.0 / 3.0)), lerp(2.0 / 3.0)], 1)
    r_c = r_c / np.cos(np.where(quad, 0.5, 1.0 / 3.0)[:, None] * (th1 - theta)[:, None] * 0.5)
    r_c = r_c * rng.uniform(1.0 - ctrl_jitter, 1.0 + ctrl_jitter, (total, 2))
    npts = np.where(quad, 2, 3)
    off = np.concatenate([[0], np.cumsum(npts)[:-1]])
    n = int(npts.sum())
    pts = np.zeros((n, 2), np.float64)
    fl = np.zeros(n, np.uint8)
    px, py = cx[path], cy[path]
    pts[off, 0] = px + r_on * np.cos(theta)
    pts[off, 1] = py + r_on * np.sin(theta)
    q = np.nonzero(quad)[0]
    tm = 0.5 * (theta[q] + th1[q])
    pts[off[q] + 1, 0] = px[q] + r_c[q, 0] * np.cos(tm)
    pts[off[q] + 1, 1] = py[q] + r_c[q, 0] * np.sin(tm)
    fl[off[q] + 1] = L.POINT_CONTROL_0
    c = np.nonzero(~quad)[0]
    for m, (f, flag) in enumerate(((1.0 / 3.0, L.POINT_CONTROL_0), (2.0 / 3.0, L.POINT_CONTROL_1))):
        tm = theta[c] + f * (th1[c] - theta[c])
        pts[off[c] + 1 + m, 0] = px[c] + r_c[c, m] * np.cos(tm)
        pts[off[c] + 1 + m, 1] = py[c] + r_c[c, m] * np.sin(tm)
        fl[off[c] + 1 + m] = flag
    seg_end = off + npts
    ends = seg_end[np.cumsum(k) - 1]
    return pts.astype(np.float32), fl, ends.astype(np.uint32)


def _clip_shape(kind, cx, cy, hw, hh):
    if kind == 0:  # rectangle: four lines
        p = [(cx - hw, cy - hh), (cx + hw, cy - hh), (cx + hw, cy + hh), (cx - hw, cy + hh)]
        return np.asarray(p, np.float32), np.zeros(4, np.uint8)
    pts, fl = [], []
    quads = [((1, 0), (0, 1)), ((0, 1), (-1, 0)), ((-1, 0), (0, -1)), ((0, -1), (1, 0))]
    for (ax, ay), (bx, by) in quads:  # ellipse: four cubic arcs
        pts.append((cx + ax * hw, cy + ay * hh))
        fl.append(0)
        pts.append((cx + (ax + KAPPA * bx) * hw, cy + (ay + KAPPA * by) * hh))
        fl.append(L.POINT_CONTROL_0)
        pts.append((cx + (bx + KAPPA * ax) * hw, cy + (by + KAPPA * ay) * hh))
        fl.append(L.POINT_CONTROL_1)
    return np.asarray(pts, np.float32), np.asarray(fl, np.uint8)


def scene_4k(seed=SEED_BASE + 1, paths=20000, clips=500, width=3840, height=2160,
             clipped_paths=2000, seg_lo=10, seg_hi=600, size_lo=4.0, size_hi=2000.0,
             jitter=0.15, ctrl_jitter=0.008):
    """20k overlapping star-shaped paths with 500 clip paths on a 4K canvas."""
    rng = np.random.default_rng(seed)
    scene = Scene()
    scene.set_view_box(RectF(0, 0, width, height))
    paints = _colors(rng, paths, 0.6, scene)
    sizes = np.exp(rng.uniform(math.log(size_lo), math.log(size_hi), paths))
    segs = np.round(np.exp(rng.uniform(math.log(seg_lo), math.log(seg_hi), paths))).astype(int)
    cx = rng.uniform(0, width, paths)
    cy = rng.uniform(0, height, paths)
    even_odd = rng.random(paths) < 0.2
    clip_of = np.full(paths, -1, np.int32)
    chosen = rng.choice(paths, size=min(clipped_paths, paths), replace=False)
    clip_of[chosen] = np.arange(len(chosen)) % clips

    # clip paths: centred on the first path that uses them
    cpts, cfl, cends, cpends = [], [], [], []
    n = 0
    for c in range(clips):
        users = chosen[np.arange(len(chosen)) % clips == c]
        j = users[0] if len(users) else int(rng.integers(paths))
        s = sizes[j] * rng.uniform(0.5, 1.5)
        hw, hh = 0.5 * s * rng.uniform(0.6, 1.0), 0.5 * s * rng.uniform(0.6, 1.0)
        ccx = min(max(cx[j], hw), width - hw) if 2 * hw < width else width / 2
        ccy = min(max(cy[j], hh), height - hh) if 2 * hh < height else height / 2
        p, f = _clip_shape(int(rng.integers(2)), ccx, ccy, hw, hh)
        cpts.append(p)
        cfl.append(f)
        n += len(p)
        cends.append(n)
        cpends.append(len(cends))
    if clips:
        scene.push_clip_paths(PathArrays(np.concatenate(cpts), np.concatenate(cfl),
                                         np.asarray(cends, np.uint32),
                                         np.asarray(cpends, np.uint32)))

    pts, fl, ends = _blobs(rng, cx, cy, 0.5 * sizes, segs, 0.5, jitter, ctrl_jitter)
    pends = np.arange(1, paths + 1, dtype=np.uint32)
    arr = PathArrays(pts, fl, ends, pends, paints,
                     np.where(even_odd, L.FILL_RULE_EVEN_ODD, L.FILL_RULE_WINDING).astype(np.int32),
                     clip_of)
    scene.push_draw_paths(arr)
    return scene, (width, height), ColorF(1.0, 1.0, 1.0, 1.0)


def _contour(points, flags):
    return np.asarray(points, np.float32), np.asarray(flags, np.uint8)


def tiny_scenes():
    """Small hand-made scenes that exercise one feature each:
    name -> (scene, (w, h), background or None)."""
    from .scene import ColorU
    out = {}

    def make(w, h):
        s = Scene()
        s.set_view_box(RectF(0, 0, w, h))
        return s

    def add(scene, contours, color, rule=L.FILL_RULE_WINDING, clip=-1, is_clip=False):
        pts = np.concatenate([c[0] for c in contours])
        fl = np.concatenate([c[1] for c in contours])
        ends = np.cumsum([len(c[0]) for c in contours]).astype(np.uint32)
        if is_clip:
            return scene.push_clip_paths(PathArrays(pts, fl, ends, [len(ends)], None, [rule],
                                                    [clip]))
        pid = scene.push_paint(ColorU(*color))
        return scene.push_draw_paths(PathArrays(pts, fl, ends, [len(ends)], [pid], [rule],
                                                [clip]))

    s = make(64, 64)
    add(s, [_contour([(8, 8), (40, 8), (40, 40), (8, 40)], [0, 0, 0, 0])], (200, 30, 30, 255))
    out["square"] = (s, (64, 64), ColorF(1, 1, 1, 1))

    s = make(64, 48)
    add(s, [_contour([(3.3, 5.1), (60.2, 20.7), (12.9, 44.4)], [0, 0, 0])], (20, 90, 200, 180))
    out["triangle"] = (s, (64, 48), ColorF(1, 1, 1, 1))

    s = make(96, 96)
    add(s, [_contour([(10, 80), (20, -30), (120, 140), (85, 12)],
                     [0, L.POINT_CONTROL_0, L.POINT_CONTROL_1, 0])], (30, 160, 60, 255))
    add(s, [_contour([(5, 50), (48, -10), (90, 50), (48, 30)],
                     [0, L.POINT_CONTROL_0, 0, L.POINT_CONTROL_0])], (250, 200, 0, 128))
    out["curves"] = (s, (96, 96), ColorF(1, 1, 1, 1))

    # winding number 2 where two contours of one path overlap
    s = make(80, 80)
    a = _contour([(5, 5), (60, 5), (60, 60), (5, 60)], [0, 0, 0, 0])
    b = _contour([(20, 20), (75, 20), (75, 75), (20, 75)], [0, 0, 0, 0])
    add(s, [a, b], (90, 20, 140, 255))
    add(s, [a, b], (255, 128, 0, 100), rule=L.FILL_RULE_EVEN_ODD)
    out["overlap"] = (s, (80, 80), ColorF(1, 1, 1, 1))

    # a clip path (ellipse) applied to a square and a triangle, nested clip on a third
    s = make(128, 96)
    e = _clip_shape(1, 60, 48, 40, 30)
    c0 = add(s, [e], None, is_clip=True)
    c2 = add(s, [_clip_shape(1, 80, 50, 35, 35)], None, is_clip=True)
    r = _clip_shape(0, 70, 40, 30, 40)
    c1 = add(s, [r], None, is_clip=True, clip=c2)
    add(s, [_contour([(10, 10), (110, 10), (110, 86), (10, 86)], [0, 0, 0, 0])],
        (0, 128, 255, 255), clip=c0)
    add(s, [_contour([(0, 90), (64, 0), (128, 90)], [0, 0, 0])], (255, 0, 64, 200), clip=c0)
    add(s, [_contour([(30, 20), (120, 30), (60, 90)], [0, 0, 0])], (20, 200, 20, 255), clip=c1)
    out["clip"] = (s, (128, 96), ColorF(0.9, 0.9, 0.9, 1))

    # a path partly outside the view box, and an empty frame edge tile
    s = make(50, 40)
    add(s, [_contour([(-20, -10), (45, 5), (30, 60)], [0, 0, 0])], (10, 10, 10, 255))
    out["offscreen"] = (s, (50, 40), None)
    return out


def _zigzag(x_lo, x_hi, y0, y1, n):
    """n lines zig-zagging between x_lo and x_hi while y climbs from y0 to y1."""
    ys = np.linspace(y0, y1, n + 1)
    xs = np.where(np.arange(n + 1) % 2 == 0, x_lo, x_hi)
    return np.stack([xs, ys], 1)


def edge_scenes():
    """Scenes at the limits the reference documents, one per limit:
    name -> (scene, (w, h), background, transform or None).

      dense_tile     tiles crossed by more than 1024 fills (fill_compute.inc.glsl:23
                     stops after 1024), a path of ~5000 fills whose tiles' fills
                     are spread over several 1024-fill chunks, and exactly 1024
                     and 1025 fills in a tile
      backdrop_wrap  200 and -200 net crossings per tile column, past the int8
                     backdrop delta of bin.cs.glsl:140-141
      huge_coords    points beyond +-32768 px (dice.cs.glsl:80 clamps microlines)
      degenerate     zero-length lines, a cubic whose points coincide, a
                     quadratic with its control point on an end, empty and
                     one-point contours, and a path with no contour at all
      mirrored       a 64-path small scene under a negative-determinant transform
    """
    from .scene import ColorU
    out = {}

    def make(w, h):
        s = Scene()
        s.set_view_box(RectF(0, 0, w, h))
        return s

    def add(scene, contours, color, rule=L.FILL_RULE_WINDING):
        contours = [(np.asarray(p, np.float32).reshape(-1, 2),
                     np.asarray(f, np.uint8).reshape(-1)) for p, f in contours]
        pts = np.concatenate([c[0] for c in contours])
        fl = np.concatenate([c[1] for c in contours])
        ends = np.cumsum([len(c[0]) for c in contours]).astype(np.uint32)
        pid = scene.push_paint(ColorU(*color))
        return scene.push_draw_paths(PathArrays(pts, fl, ends, [len(ends)], [pid], [rule], [-1]))

    def lines(pts):
        pts = np.asarray(pts, np.float32)
        return pts, np.zeros(len(pts), np.uint8)

    s = make(96, 64)
    # ~5000 fills: tile (1, 1), then tile (2, 1), then tile (1, 1) again, so
    # tile (1, 1) has more than 1024 fills in separate chunks of the path
    zz = np.concatenate([_zigzag(17.3, 30.7, 17.0, 24.0, 2000),
                         _zigzag(33.1, 46.9, 17.0, 30.0, 1000)[1:],
                         _zigzag(17.6, 30.2, 24.5, 31.0, 2000)[1:]])
    add(s, [lines(zz)], (200, 40, 40, 255))
    # exactly 1024 and 1025 fills in tiles (4, 1) and (4, 2)
    add(s, [lines(_zigzag(66.2, 78.8, 17.0, 31.0, 1024))], (40, 200, 40, 160))
    add(s, [lines(_zigzag(65.7, 79.4, 33.0, 47.0, 1025))], (40, 40, 200, 200))
    out["dense_tile"] = (s, (96, 64), ColorF(1, 1, 1, 1), None)

    s = make(128, 96)
    sq = [(20.5, 10.25), (75.3, 10.25), (75.3, 50.75), (20.5, 50.75)]
    add(s, [lines(sq)] * 200, (90, 20, 140, 255))
    add(s, [lines(sq[::-1] + np.float32([30.0, 40.0]))] * 200, (255, 128, 0, 120))
    # 131 and -129 crossings: just past the int8 range on either side
    add(s, [lines([(3.5, 70.5), (40.5, 70.5), (40.5, 90.5), (3.5, 90.5)])] * 131, (0, 90, 0, 255))
    add(s, [lines([(100.5, 60.5), (100.5, 90.5), (60.5, 90.5), (60.5, 60.5)])] * 129,
        (0, 0, 90, 200))
    out["backdrop_wrap"] = (s, (128, 96), ColorF(1, 1, 1, 1), None)

    s = make(128, 96)
    add(s, [lines([(-50000.5, 10.25), (100.7, 40.1), (20.3, 90.9)])], (30, 30, 30, 255))
    add(s, [(np.float32([(10.5, 60.5), (0.0, 70000.0), (90000.0, -45000.0), (120.25, 80.5)]),
             np.uint8([0, L.POINT_CONTROL_0, L.POINT_CONTROL_1, 0]))], (200, 20, 90, 170))
    add(s, [lines([(64.0, 5.0), (40000.0, 48.0), (64.0, 90.0), (-33000.25, 48.0)])],
        (20, 120, 220, 140), rule=L.FILL_RULE_EVEN_ODD)
    add(s, [lines([(40000.0, 40000.0), (40100.0, 40000.0), (40050.0, 40100.0)])], (0, 0, 0, 255))
    out["huge_coords"] = (s, (128, 96), ColorF(1, 1, 1, 1), None)

    s = make(64, 64)
    add(s, [lines([(10, 10), (10, 10), (40.5, 12.25), (40.5, 12.25), (30.75, 40.5), (10, 10)])],
        (150, 60, 20, 255))
    add(s, [(np.float32([(50, 50), (50, 50), (50, 50), (50, 50)]),
             np.uint8([0, L.POINT_CONTROL_0, L.POINT_CONTROL_1, 0])),
            (np.float32([(5.5, 40.5), (5.5, 40.5), (25.25, 60.75)]),
             np.uint8([0, L.POINT_CONTROL_0, 0])),
            (np.float32([(44.5, 20.5), (60.25, 30.5), (60.25, 30.5)]),
             np.uint8([0, 0, L.POINT_CONTROL_0])),
            (np.zeros((0, 2), np.float32), np.zeros(0, np.uint8)),
            (np.float32([(33.3, 33.3)]), np.uint8([0]))], (20, 150, 200, 200))
    add(s, [(np.zeros((0, 2), np.float32), np.zeros(0, np.uint8))], (0, 0, 0, 255))
    add(s, [lines([(2.5, 2.5), (60.5, 3.5), (2.5, 3.5)])], (60, 60, 60, 255))
    out["degenerate"] = (s, (64, 64), ColorF(1, 1, 1, 1), None)

    scene, size, bg = small_scene(paths=64, size=512)
    out["mirrored"] = (scene, size, bg, (-1.0, 0.0, 0.0, 1.0, 512.0, 0.0))
    return out


def blend_scene(seed=SEED_BASE + 5, width=160, height=120):
    """One draw path per blend mode (27), overlapping, opaque and translucent
    paints, some clipped: destructive modes tile the whole view box
    (builder.rs:430-436) and only opaque SrcOver/Clear paths write z."""
    from .scene import ColorU
    rng = np.random.default_rng(seed)
    s = Scene()
    s.set_view_box(RectF(0, 0, width, height))
    cp, cf = _clip_shape(1, width * 0.5, height * 0.5, width * 0.3, height * 0.3)
    clip = s.push_clip_paths(PathArrays(cp, cf, [len(cp)], [1], None, [0], [-1]))
    parts = []
    for mode in range(27):
        cx, cy = rng.uniform(0.2, 0.8) * width, rng.uniform(0.2, 0.8) * height
        r = rng.uniform(10, 40)
        pts, fl = _clip_shape(mode % 2, cx, cy, r, r * rng.uniform(0.6, 1.2))
        a = 255 if mode % 3 != 2 else int(rng.integers(60, 230))
        pid = s.push_paint(ColorU(*[int(v) for v in rng.integers(0, 256, 3)], a))
        parts.append(PathArrays(pts, fl, [len(pts)], [1], [pid],
                                [int(rng.integers(2))], [clip if mode % 5 == 4 else -1], [mode]))
    s.push_draw_paths(PathArrays.concat(parts))
    return s, (width, height), ColorF(1, 1, 1, 1)


def _rotation_scale(theta, s, tx, ty):
    c, sn = math.cos(theta) * s, math.sin(theta) * s
    return (c, -sn, sn, c, tx, ty)


def paint_scene(seed=SEED_BASE + 6, size=512, paths=96):
    """Gradient and pattern paints (paint.rs, gradient.rs, pattern.rs) on
    random closed cubic paths: linear gradients (clamp and repeat, stops
    added out of order, some translucent), radial gradients (concentric and
    two-point, transformed), and image patterns (repeat and clamp, smoothing
    on and off, blur and colour-matrix filters, DestIn, one image shared by two
    paints), interleaved with solid colours so that batches break and adopt
    colour textures. Returns (scene, (size, size), background)."""
    from .scene import ColorU, Gradient, Image, Paint, Pattern
    rng = np.random.default_rng(seed)
    scene = Scene()
    scene.set_view_box(RectF(0, 0, size, size))

    def color(opaque=True):
        r, g, b = (int(v) for v in rng.integers(0, 256, 3))
        return ColorU(r, g, b, 255 if opaque else int(rng.integers(40, 255)))

    def stops(g, n, opaque=True):
        offs = rng.uniform(0, 1, n)
        offs[rng.random(n) < 0.2] = 0.0
        for o in offs:  # out of order: Gradient::add keeps them sorted
            g.add_color_stop(color(opaque or rng.random() < 0.5), float(o))
        return g

    def image(w, h, opaque=True):
        px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        if opaque:
            px[..., 3] = 255
        return Image(px)

    pool = []
    for k in range(3):
        g = stops(Gradient.linear_from_points(rng.uniform(0, size, 2), rng.uniform(0, size, 2)),
                  int(rng.integers(2, 6)), opaque=k != 2)
        pool.append(Paint.from_gradient(g))
    g = stops(Gradient.linear_from_points((100.0, 100.0), (140.0, 130.0)), 3)
    g.wrap = Gradient.REPEAT
    pool.append(Paint.from_gradient(g))
    g = stops(Gradient.radial((size / 2, size / 2), (10.0, size / 2.5)), 4)
    pool.append(Paint.from_gradient(g))
    g = stops(Gradient.radial(((150.0, 200.0), (330.0, 260.0)), (20.0, 140.0)), 3)
    g.apply_transform(_rotation_scale(0.3, 1.1, -20.0, 15.0))
    pool.append(Paint.from_gradient(g))
    g = stops(Gradient.radial((260.0, 250.0), (0.0, 60.0)), 2)
    g.wrap = Gradient.REPEAT
    pool.append(Paint.from_gradient(g))
    shared = image(13, 9)
    p = Pattern.from_image(shared)
    p.repeat_x = p.repeat_y = True
    p.apply_transform(_rotation_scale(0.4, 7.0, 31.0, -12.0))
    pool.append(Paint.from_pattern(p))
    p = Pattern.from_image(shared)
    p.repeat_x = p.repeat_y = True
    p.apply_transform(_rotation_scale(-1.1, 3.5, 0.0, 0.0))
    pool.append(Paint.from_pattern(p))
    p = Pattern.from_image(image(32, 24, opaque=False))
    p.smoothing_enabled = False
    p.apply_transform(_rotation_scale(0.0, 9.0, 90.0, 120.0))
    pool.append(Paint.from_pattern(p))
    p = Pattern.from_image(image(40, 30))
    p.repeat_x = True
    p.set_blur(False, 2.0)
    p.apply_transform(_rotation_scale(0.2, 4.0, 10.0, 10.0))
    pool.append(Paint.from_pattern(p))
    p = Pattern.from_image(image(20, 20))
    p.repeat_x = p.repeat_y = True
    p.set_color_matrix(np.concatenate([rng.uniform(-0.5, 1.0, 16), rng.uniform(-0.2, 0.3, 4)]))
    p.apply_transform(_rotation_scale(0.7, 6.0, 0.0, 0.0))
    pool.append(Paint.from_pattern(p))
    p = Pattern.from_image(image(16, 16))
    p.repeat_x = p.repeat_y = True
    p.apply_transform(_rotation_scale(0.0, 5.0, 3.0, 3.0))
    dest_in = Paint.from_pattern(p)
    dest_in.color = ColorU(220, 60, 30, 255)
    dest_in.composite_op = Paint.DEST_IN
    pool.append(dest_in)
    arrays = _cubic_paths(rng, paths, size)
    ids = []
    for i in range(paths):
        if rng.random() < 0.25:
            ids.append(scene.push_paint(Paint.from_color(color(rng.random() < 0.6))))
        else:
            ids.append(scene.push_paint(pool[int(rng.integers(0, len(pool)))]))
    arrays.paints = np.asarray(ids, np.uint32)
    scene.push_draw_paths(arrays)
    return scene, (size, size), ColorF(1.0, 1.0, 1.0, 1.0)


def gradient_pages_scene(seed=SEED_BASE + 7, count=300, size=480):
    """`count` squares, each with its own two-stop linear gradient: more
    gradients than the 256 rows of one gradient page, so the palette opens a
    second page and the draw batch breaks there."""
    from .scene import ColorU, Gradient, Paint
    rng = np.random.default_rng(seed)
    scene = Scene()
    scene.set_view_box(RectF(0, 0, size, size))
    side = int(math.ceil(math.sqrt(count)))
    cell = size / side
    pts, ids = [], []
    for i in range(count):
        x0, y0 = (i % side) * cell, (i // side) * cell
        pts.append([(x0 + 1, y0 + 1), (x0 + cell - 1, y0 + 1), (x0 + cell - 1, y0 + cell - 1),
                    (x0 + 1, y0 + cell - 1)])
        g = Gradient.linear_from_points((x0, y0), (x0 + cell, y0 + cell * rng.uniform(0, 1)))
        for o in (0.0, 1.0):
            r, gg, b = (int(v) for v in rng.integers(0, 256, 3))
            g.add_color_stop(ColorU(r, gg, b, 255), o)
        ids.append(scene.push_paint(Paint.from_gradient(g)))
    pts = np.asarray(pts, np.float32).reshape(-1, 2)
    arrays = PathArrays(pts, np.zeros(len(pts), np.uint8),
                        np.arange(4, len(pts) + 1, 4, dtype=np.uint32),
                        np.arange(1, count + 1, dtype=np.uint32), np.asarray(ids, np.uint32))
    scene.push_draw_paths(arrays)
    return scene, (size, size), ColorF(0.0, 0.0, 0.0, 1.0)
