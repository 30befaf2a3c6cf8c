"""The palette restatement (oracle/palette.py) pinned by what the reference
and the mathematics fix, and the C++ palette held to it:

  - content/src/gradient.rs's own unit tests (stable_order,
    never_sample_zero_width) on Gradient::add and sample;
  - gradient ramps against float64 piecewise-linear interpolation of the
    stops at texel centres (x + 0.5) / 256, within the 1/255 of to_u8;
  - texture transforms in float64: a linear gradient's u is 0 at its start
    and 1 at its end, v the centre of its row; a radial gradient's transform
    undoes the gradient's own; a pattern's maps pattern space onto the page;
  - planted mistakes in the restatement fail those pins;
  - the C++ ramp (through the command list) equals the restatement's, also
    for stops handed over out of order.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import packing, palette
from pathfinder_amd import BuildOptions, Transform2F, synth
from pathfinder_amd import _lib as L
from pathfinder_amd.scene import ColorU, Gradient, Image, Paint, Pattern, RectF, Scene

F = np.float32


def test_reference_stable_order():
    """gradient.rs test stable_order: 110 stops with offsets (i % 11) / 10
    stay sorted by offset, ties in insertion order."""
    g = Gradient.linear_from_points((0, 0), (0, 0))
    for i in range(110):
        g.add_color_stop(ColorU(i, 0, 0, 1), (i % 11) / 10.0)
    for a, b in zip(g.stops, g.stops[1:]):
        assert a.offset < b.offset or a.color.r < b.color.r


def test_reference_never_sample_zero_width():
    """gradient.rs test never_sample_zero_width."""
    g = Gradient.linear_from_points((0, 0), (0, 0))
    for i in range(110):
        zero_width = i == 0 or 11 <= i < 99 or i == 109
        g.add_color_stop(ColorU(255 if zero_width else 0, 0, 0, 1), (i % 11) / 10.0)
    stops = [(F(s.offset), tuple(s.color)) for s in g.stops]
    for i in range(11):
        assert palette.gradient_sample(stops, F(i) / F(10))[0] == 0, i


def _model_ramp(stops):
    """float64 piecewise-linear colour of sorted stops at the 256 texel centres."""
    t = (np.arange(256) + 0.5) / 256.0
    offs = np.asarray([s[0] for s in stops], np.float64)
    cols = np.asarray([s[1] for s in stops], np.float64)
    out = np.zeros((256, 4))
    for k, tk in enumerate(t):
        if tk <= offs[0]:
            out[k] = cols[0]
        elif tk >= offs[-1]:
            out[k] = cols[-1]
        else:
            j = np.searchsorted(offs, tk, side="left")
            a, b = offs[j - 1], offs[j]
            out[k] = cols[j - 1] + (cols[j] - cols[j - 1]) * (tk - a) / (b - a)
    return out


def _random_stops(rng, n):
    offs = np.sort(rng.uniform(0, 1, n)).astype(np.float32)
    return [(F(o), tuple(int(v) for v in rng.integers(0, 256, 4))) for o in offs]


def _ramp(sample, stops):
    return np.asarray([sample(stops, (F(x) + F(0.5)) / F(256)) for x in range(256)], np.float64)


@pytest.mark.parametrize("seed", range(6))
def test_ramp_is_piecewise_linear(seed):
    rng = np.random.default_rng(seed)
    stops = _random_stops(rng, int(rng.integers(2, 7)))
    got = _ramp(palette.gradient_sample, stops)
    assert np.abs(got - _model_ramp(stops)).max() <= 0.5 + 1e-3


def _ratio_from_upper(stops, t):
    """A planted mistake: the ratio measured from the upper stop."""
    t = F(t)
    n = 0
    while n < len(stops) and (stops[n][0] < t or stops[n][0] == 0):
        n += 1
    up = min(n, len(stops) - 1)
    lo = up - 1 if up > 0 else up
    d = F(stops[up][0]) - F(stops[lo][0])
    if d == 0:
        return stops[lo][1]
    r = abs((t - F(stops[up][0])) / d)
    return tuple(int(np.rint((F(a) + (F(b) - F(a)) * r))) & 0xFF
                 for a, b in zip(stops[lo][1], stops[up][1]))


def test_planted_ramp_mistake_fails():
    rng = np.random.default_rng(0)
    stops = _random_stops(rng, 4)
    assert np.abs(_ramp(_ratio_from_upper, stops) - _model_ramp(stops)).max() > 2


def _xf64(t):
    m00, m01, m10, m11, tx, ty = [float(v) for v in t]
    return np.asarray([[m00, m01, tx], [m10, m11, ty], [0, 0, 1]])


BUILD = (0.8, 0.3, -0.25, 0.9, 100.5, -40.25)


def _one_paint_scene(paint):
    s = Scene()
    s.set_view_box(RectF(0, 0, 64, 64))
    s.push_paint(paint)
    return s


def test_linear_transform_maps_line_to_unit_interval():
    g = Gradient.linear_from_points((12.5, 40.0), (300.25, -17.0))
    g.add_color_stop(ColorU(0, 0, 0, 255), 0.0)
    g.add_color_stop(ColorU(255, 255, 255, 255), 1.0)
    pal = palette.build([Paint.from_gradient(g)], BUILD)
    T = _xf64(pal["metadata"][0]["transform"]) @ _xf64(BUILD)  # scene -> uv
    for p, u_want in ((g.from_, 0.0), (g.to, 1.0)):
        uv = T @ np.asarray([p[0], p[1], 1.0])
        assert abs(uv[0] - u_want) < 1e-5 and abs(uv[1] - 0.5 / 256) < 1e-7
    # along the perpendicular u does not change
    d = np.asarray(g.to) - np.asarray(g.from_)
    q = np.asarray(g.from_) + 37.0 * np.asarray([-d[1], d[0]]) / np.hypot(*d)
    assert abs((T @ np.asarray([q[0], q[1], 1.0]))[0]) < 1e-5


def test_radial_transform_undoes_gradient_transform():
    g = Gradient.radial(((10.0, 20.0), (30.0, 25.0)), (5.0, 40.0))
    g.apply_transform((1.5, 0.4, -0.2, 0.7, 33.0, -9.0))
    g.add_color_stop(ColorU(255, 0, 0, 255), 0.0)
    pal = palette.build([Paint.from_gradient(g)], BUILD)
    T = _xf64(pal["metadata"][0]["transform"]) @ _xf64(BUILD) @ _xf64(g.transform)
    assert np.abs(T - np.eye(3)).max() < 2e-5
    fp = pal["metadata"][0]["filter_params"]
    assert tuple(fp[:6]) == (10.0, 20.0, 30.0, 25.0, 5.0, 40.0)
    assert fp[6] == 0.0 and fp[7] == F(0.5 / 256)


@pytest.mark.parametrize("repeat", [False, True])
def test_pattern_transform_maps_pattern_space_to_page(repeat):
    img = Image(np.zeros((9, 13, 4), np.uint8))
    p = Pattern.from_image(img)
    p.repeat_x = p.repeat_y = repeat
    p.apply_transform((2.0, 0.5, -0.25, 3.0, 7.0, 11.0))
    pal = palette.build([Paint.from_pattern(p)], BUILD)
    pw, ph = pal["pages"][0]
    assert (pw, ph) == ((13, 9) if repeat else (15, 11))
    T = _xf64(pal["metadata"][0]["transform"]) @ _xf64(BUILD) @ _xf64(p.transform)
    assert np.abs(T - np.diag([1.0 / pw, 1.0 / ph, 1.0])).max() < 2e-6
    page, x, y, _ = pal["texels"][0]
    assert (x, y) == ((0, 0) if repeat else (1, 1))


def test_opaque_rules():
    """Paint::is_opaque: opaque base and every stop or pixel opaque."""
    g = Gradient.linear_from_points((0, 0), (1, 1))
    g.add_color_stop(ColorU(1, 2, 3, 255), 0.0)
    assert palette.is_opaque(Paint.from_gradient(g))
    g.add_color_stop(ColorU(1, 2, 3, 254), 1.0)
    assert not palette.is_opaque(Paint.from_gradient(g))
    px = np.full((2, 2, 4), 255, np.uint8)
    assert palette.is_opaque(Paint.from_pattern(Pattern.from_image(Image(px))))
    px[1, 1, 3] = 0
    assert not palette.is_opaque(Paint.from_pattern(Pattern.from_image(Image(px))))


def test_product_ramp_equals_restatement_with_unsorted_stops():
    """pfh_scene_push_gradient_paint with stops in add order (not sorted):
    the library sorts them stably, and its ramp texels equal the
    restatement's for many random gradients."""
    lib = L.load()
    rng = np.random.default_rng(11)
    h = lib.pfh_scene_create()
    want = []
    try:
        vb = L.RectF(L.Vector2F(0, 0), L.Vector2F(64, 64))
        lib.pfh_scene_set_view_box(h, C.byref(vb))
        for k in range(40):
            n = int(rng.integers(1, 8))
            offs = rng.uniform(-0.2, 1.2, n).astype(np.float32)
            offs[rng.random(n) < 0.3] = 0.0
            if n > 2:
                offs[1] = offs[0]
            cols = rng.integers(0, 256, (n, 4))
            arr = (L.ColorStop * n)(*[L.ColorStop(float(o), L.ColorU(*[int(v) for v in c]))
                                      for o, c in zip(offs, cols)])
            g = L.Gradient(L.GRADIENT_LINEAR, L.Vector2F(0, 0), L.Vector2F(1, 0), 0, 0,
                           L.Transform2F(1, 0, 0, 1, 0, 0),
                           C.cast(arr, C.POINTER(L.ColorStop)), n, 0)
            base = L.ColorU(255, 255, 255, 255)
            assert lib.pfh_scene_push_gradient_paint(h, C.byref(base), C.byref(g), 0) == k
            order = sorted(range(n), key=lambda i: offs[i])  # Python's sort is stable
            want.append(_ramp(palette.gradient_sample,
                              [(F(offs[i]), tuple(int(v) for v in cols[i])) for i in order]))
        bo = L.BuildOptions(L.Transform2F(1, 0, 0, 1, 0, 0))
        out = C.c_void_p()
        L.check(lib.pfh_scene_build(h, C.byref(bo), C.byref(out)))
        from pathfinder_amd.scene import CommandList
        cl = CommandList(out)
        (page, x, y, tex), = cl.texel_uploads()
        for k, w in enumerate(want):
            assert np.array_equal(tex[k].astype(np.float64), w), k
    finally:
        lib.pfh_scene_destroy(h)


def test_invalid_paints_are_refused():
    lib = L.load()
    h = lib.pfh_scene_create()
    try:
        base = L.ColorU(1, 2, 3, 255)
        g = L.Gradient()
        g.geometry = 7
        assert lib.pfh_scene_push_gradient_paint(h, C.byref(base), C.byref(g), 0) == \
            L.INVALID_PAINT
        px = np.zeros((2, 2, 4), np.uint8)
        p = L.Pattern(2, 2, px.ctypes.data_as(C.c_void_p), L.Transform2F(1, 0, 0, 1, 0, 0), 0,
                      L.Filter())
        p.filter.kind = L.FILTER_RADIAL_GRADIENT
        assert lib.pfh_scene_push_pattern_paint(h, C.byref(base), C.byref(p), 0) == \
            L.INVALID_PAINT
        p.filter.kind = L.FILTER_NONE
        p.width = 0
        assert lib.pfh_scene_push_pattern_paint(h, C.byref(base), C.byref(p), 0) == \
            L.INVALID_PAINT
    finally:
        lib.pfh_scene_destroy(h)
