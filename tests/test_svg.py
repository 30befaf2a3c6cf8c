"""The SVG reader and the stroker of the scene front-end (pfh_svg.cpp).

Parity with the reference's usvg-based reader is unpinned: no reference test
or fixture holds the outlines it produces. These tests pin the reader against
what SVG's own definitions fix (path-data commands, arcs on their ellipse,
transforms, shapes, colours) and the stroker against closed-form areas; the
packing of the result is held to the packing restatement, and its render to
the oracle on the GPU (tests/test_gpu_edge.py).

tests/golden/Ghostscript_Tiger.svg is the reference's resources/svg test
image, kept as input data.
"""
import os

import numpy as np
import pytest

import oracle
from helpers import oracle_frame
from oracle import packing
from pathfinder_amd import BuildOptions, ColorF, RectF, Transform2F
from pathfinder_amd import _lib as L
from pathfinder_amd.scene import Outline, stroke_outline, svg_scene

HERE = os.path.dirname(os.path.abspath(__file__))
TIGER = os.path.join(HERE, "golden", "Ghostscript_Tiger.svg")


def _svg(body, attrs='viewBox="0 0 100 100"'):
    return '<svg xmlns="http://www.w3.org/2000/svg" %s>%s</svg>' % (attrs, body)


def _paths(doc):
    s, info = svg_scene(doc)
    d = s.draw_arrays()
    out = []
    for i in range(len(d)):
        pts, fl, ends = d.path(i)
        out.append((pts.astype(np.float64), fl, ends, s.paints[int(d.paints[i])],
                    int(d.fill_rules[i])))
    return out, info


def _contours(pts, ends):
    start = 0
    for e in ends:
        yield pts[start:e]
        start = int(e)


def test_tiger_reads_every_path():
    s, info = svg_scene(open(TIGER, "rb").read())
    assert info["view_box"] == [0.0, 0.0, 200.0, 200.0]
    # 138 path elements: 130 with a fill, 52 with a stroke, nothing skipped
    assert info["fills"] == 130 and info["strokes"] == 52
    assert info["skipped_elements"] == 0 and info["unsupported_paints"] == 0
    assert s.draw_path_count() == 182
    d = s.draw_arrays()
    assert np.all(np.isfinite(d.points))
    lo, hi = d.points.min(axis=0), d.points.max(axis=0)
    assert lo.min() > -5 and hi.max() < 205


def test_tiger_packing_matches_restatement():
    s, _ = svg_scene(open(TIGER, "rb").read())
    s.set_view_box(RectF(0, 0, 800, 800))
    t = (4, 0, 0, 4, 0, 0)
    want = packing.build(s, t)
    cl = s.build(BuildOptions(Transform2F(*t)))
    d, c = cl.segments()
    assert np.array_equal(d["points"], want["draw"]["points"])
    assert np.array_equal(d["indices"], want["draw"]["indices"])
    (_, gb), = [kb for kb in cl.batches() if kb[0] == packing.DRAW]
    (_, wb), = [kb for kb in want["batches"] if kb[0] == packing.DRAW]
    for key in ("tile_path_info", "propagate_metadata", "backdrops", "dice_metadata"):
        assert np.array_equal(gb[key], np.asarray(wb[key]).astype(gb[key].dtype).reshape(
            gb[key].shape)), key


def test_tiger_renders_through_the_oracle():
    s, _ = svg_scene(open(TIGER, "rb").read())
    s.set_view_box(RectF(0, 0, 400, 400))
    cmds, o = oracle_frame(s, (400, 400), ColorF(1, 1, 1, 1), threads=4,
                           transform=(2, 0, 0, 2, 0, 0))
    img = o.read_framebuffer()
    o.close()
    # the tiger's orange fur and black stripes are there
    rgb = img[..., :3].reshape(-1, 3).astype(int)
    orange = np.sum((rgb[:, 0] > 180) & (rgb[:, 1] > 80) & (rgb[:, 1] < 140) & (rgb[:, 2] < 60))
    black = np.sum(rgb.sum(axis=1) < 60)
    assert orange > 5000 and black > 10000, (orange, black)


def test_path_data_absolute_relative_and_implicit():
    a, _ = _paths(_svg('<path d="M10 20 L30 40 H50 V60 Z" fill="#000"/>'))
    b, _ = _paths(_svg('<path d="m10 20 l20 20 h20 v20 z" fill="#000"/>'))
    c, _ = _paths(_svg('<path d="M10,20 30,40 50 40 50,60z" fill="#000"/>'))
    want = np.array([[10, 20], [30, 40], [50, 40], [50, 60]], np.float64)
    for p in (a, b, c):
        assert np.array_equal(p[0][0], want)
        assert list(p[0][2]) == [4]


def test_path_data_number_forms():
    p, _ = _paths(_svg('<path d="M.5-.5L1e1,2E0l-.25.75" fill="#000"/>'))
    assert np.allclose(p[0][0], [[0.5, -0.5], [10, 2], [9.75, 2.75]])


def test_quadratics_are_elevated_exactly():
    p, _ = _paths(_svg('<path d="M0 0 Q 30 60 90 0 T 150 30" fill="#000"/>'))
    pts, fl = p[0][0], p[0][1]
    assert list(fl) == [0, 1, 2, 0, 1, 2, 0]
    # Q (0,0) (30,60) (90,0) as a cubic: controls at 2/3 towards the control point
    assert np.allclose(pts[1], [20, 40]) and np.allclose(pts[2], [50, 40])
    # T reflects the last control point: (150, -60) about (90, 0)
    q = np.array([150.0, -60.0])
    assert np.allclose(pts[4], pts[3] + (q - pts[3]) * 2 / 3)
    assert np.allclose(pts[5], pts[6] + (q - pts[6]) * 2 / 3)


def test_smooth_cubic_reflects_control():
    p, _ = _paths(_svg('<path d="M0 0 C 10 20 30 20 40 0 S 70 -20 80 0" fill="#000"/>'))
    pts = p[0][0]
    assert np.allclose(pts[4], 2 * pts[3] - pts[2])


def _bezier(p0, c0, c1, p1, t):
    u = 1 - t
    return (u ** 3)[:, None] * p0 + (3 * u * u * t)[:, None] * c0 + \
        (3 * u * t * t)[:, None] * c1 + (t ** 3)[:, None] * p1


@pytest.mark.parametrize("d,center,r", [
    ("M 100 50 A 50 50 0 0 1 0 50", (50, 50), 50),       # half circle
    ("M 50 0 A 50 50 0 1 0 100 50", (50, 50), 50),       # three quarters
    ("M 80 50 a 30 30 0 0 0 -60 0 a 30 30 0 0 0 60 0", (50, 50), 30),
])
def test_arcs_lie_on_their_circle(d, center, r):
    p, _ = _paths(_svg('<path d="%s" fill="#000"/>' % d))
    pts, fl = p[0][0], p[0][1]
    t = np.linspace(0, 1, 33)
    i = 0
    while i + 3 < len(pts):
        assert fl[i + 1] == 1 and fl[i + 2] == 2
        q = _bezier(pts[i], pts[i + 1], pts[i + 2], pts[i + 3], t)
        dist = np.hypot(q[:, 0] - center[0], q[:, 1] - center[1])
        assert np.max(np.abs(dist - r)) < 1e-3 * r  # cubic arcs of <= 90 degrees
        i += 3


def test_transforms_and_shapes():
    doc = _svg('<g transform="translate(10 5) scale(2)">'
               '<rect x="1" y="2" width="3" height="4" fill="red"/></g>'
               '<rect x="0" y="0" width="10" height="10" transform="rotate(90 5 5)" fill="blue"/>'
               '<circle cx="50" cy="50" r="10" fill="rgb(0, 50%, 100%)"/>'
               '<polygon points="0,0 10,0 0,10" fill="#abc" fill-opacity="0.5"/>')
    p, info = _paths(doc)
    assert info["fills"] == 4
    assert np.allclose(p[0][0], [[12, 9], [18, 9], [18, 17], [12, 17]])
    c = p[0][3]
    assert (c.r, c.g, c.b, c.a) == (255, 0, 0, 255)
    assert np.allclose(p[1][0], [[10, 0], [10, 10], [0, 10], [0, 0]], atol=1e-5)
    c = p[2][3]
    assert (c.r, c.g, c.b) == (0, 128, 255)
    ring = p[2][0][p[2][1] == 0]
    assert np.allclose(np.hypot(ring[:, 0] - 50, ring[:, 1] - 50), 10, atol=1e-4)
    c = p[3][3]
    assert (c.r, c.g, c.b, c.a) == (0xaa, 0xbb, 0xcc, 128)


def test_unsupported_elements_are_skipped_and_counted():
    doc = _svg('<defs><linearGradient id="g"/></defs><text>x</text>'
               '<path d="M0 0 L10 0 L0 10z" fill="url(#g)"/><g display="none">'
               '<path d="M0 0 L10 0 L0 10z"/></g><path d="M0 0 L10 0 L0 10z"/>')
    p, info = _paths(doc)
    assert info["skipped_elements"] == 2 and info["unsupported_paints"] == 1
    assert len(p) == 1  # the default fill (black) of the last path


def test_malformed_document_is_refused():
    with pytest.raises(ValueError):
        svg_scene(b'<svg viewBox="0 0 10 10"><path d="M0 0 L1 1"</svg>')
    with pytest.raises(ValueError):
        svg_scene(_svg('<path d="M0 0 L1 x"/>'))


def _coverage_area(outline, size=64, scale=4.0):
    """Area covered by the nonzero fill of an outline, from the oracle's
    D3D11 render at `scale` pixels per unit (alpha sum / 255 / scale^2)."""
    from pathfinder_amd import ColorU, PathArrays, Scene
    s = Scene()
    s.set_view_box(RectF(0, 0, size * scale, size * scale))
    pid = s.push_paint(ColorU(0, 0, 0, 255))
    s.push_draw_paths(PathArrays(outline.points, outline.flags, outline.contour_ends,
                                 [len(outline.contour_ends)], [pid]))
    cmds, o = oracle_frame(s, (int(size * scale),) * 2, None, threads=4,
                           transform=(scale, 0, 0, scale, 0, 0))
    a = o.read_framebuffer()[..., 3].astype(np.float64).sum() / 255.0 / scale ** 2
    o.close()
    return a


def _line(points, closed=False):
    pts = np.asarray(points, np.float32)
    return Outline(pts, np.zeros(len(pts), np.uint8), [len(pts)]), [1 if closed else 0]


@pytest.mark.parametrize("cap,extra", [(0, 0.0), (1, 2 * 4.0 * 8.0), (2, np.pi * 4.0 ** 2)])
def test_stroke_of_a_line_has_the_closed_form_area(cap, extra):
    o, closed = _line([(10, 30), (50, 30)])
    s = stroke_outline(o, 8.0, cap=cap, closed=closed)
    area = _coverage_area(s)
    want = 40 * 8 + extra
    assert abs(area - want) < 0.01 * want, (area, want)


@pytest.mark.parametrize("join", [0, 1, 2])
def test_stroke_joins_of_a_right_angle(join):
    o, closed = _line([(10, 10), (40, 10), (40, 40)])
    w = 6.0
    s = stroke_outline(o, w, cap=0, join=join, closed=closed)
    area = _coverage_area(s)
    # two legs of 30 x 6 overlapping in a 3 x 3 square at the corner... plus the join
    # outside it: miter a 3 x 3 square, bevel half of it, round a quarter disc
    base = 30 * w + 30 * w - (w / 2) ** 2
    outer = {0: (w / 2) ** 2, 1: (w / 2) ** 2 / 2, 2: np.pi * (w / 2) ** 2 / 4}[join]
    want = base + outer
    assert abs(area - want) < 0.01 * want, (area, want)


def test_miter_limit_falls_back_to_bevel():
    # a 20 degree turn: the miter length ratio 1/sin(10 deg) = 5.76 exceeds 4
    a = np.radians(160)
    pts = [(10, 30), (40, 30), (40 + 25 * np.cos(a), 30 + 25 * np.sin(a))]
    o, closed = _line(pts)
    m = stroke_outline(o, 4.0, join=0, miter_limit=4.0, closed=closed)
    b = stroke_outline(o, 4.0, join=1, closed=closed)
    assert np.array_equal(m.points, b.points)
    m10 = stroke_outline(o, 4.0, join=0, miter_limit=10.0, closed=closed)
    assert len(m10.points) > len(b.points)


@pytest.mark.parametrize("cap,join", [(0, 0), (1, 1), (2, 2)])
def test_stroke_contours_are_positively_oriented(cap, join):
    """Every piece of a stroke turns counter-clockwise on screen (positive
    shoelace area, y down), so the nonzero fill of the pieces is their union."""
    rng = np.random.default_rng(3)
    pts = np.cumsum(rng.normal(0, 10, (40, 2)), axis=0).astype(np.float32) + 100
    o, closed = _line(pts, closed=bool(cap % 2))
    s = stroke_outline(o, 3.0, cap=cap, join=join, closed=closed)
    assert len(s.contour_ends) > 40
    for c in _contours(s.points.astype(np.float64), s.contour_ends):
        x, y = c[:, 0], c[:, 1]
        assert 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y) > 0
