"""Scene / SceneBuilder / Renderer API over libpfhip.so.

Mirrors the reference's front-end types for the D3D11-level path:
  Scene, DrawPath, ClipPath, Paint          renderer/src/scene.rs, paint.rs
  Outline, Contour, PointFlags, FillRule    content/src/outline.rs, fill.rs
  BuildOptions, RenderTransform             renderer/src/options.rs:53-90
  RendererOptions, RendererMode, DestFramebuffer, RendererLevel
                                            renderer/src/gpu/options.rs
  Renderer (begin_scene / render_command / end_scene)
                                            renderer/src/gpu/renderer.rs:350-460
Building (SceneBuilder::build) and rendering run in the native library.
"""
import ctypes as C
import enum

import numpy as np

from . import _lib as L


class FillRule(enum.IntEnum):
    WINDING = L.FILL_RULE_WINDING
    EVEN_ODD = L.FILL_RULE_EVEN_ODD


class BlendMode(enum.IntEnum):
    """BlendMode (content/src/effects.rs:99-160); SrcOver is 0 (PFH_BLEND_MODE_*)."""
    SRC_OVER = 0
    CLEAR = 1
    COPY = 2
    SRC_IN = 3
    SRC_OUT = 4
    SRC_ATOP = 5
    DEST_IN = 6
    DEST_OUT = 7
    DEST_OVER = 8
    DEST_ATOP = 9
    XOR = 10
    LIGHTER = 11
    DARKEN = 12
    LIGHTEN = 13
    MULTIPLY = 14
    SCREEN = 15
    HARD_LIGHT = 16
    OVERLAY = 17
    COLOR_DODGE = 18
    COLOR_BURN = 19
    SOFT_LIGHT = 20
    DIFFERENCE = 21
    EXCLUSION = 22
    HUE = 23
    SATURATION = 24
    COLOR = 25
    LUMINOSITY = 26


class RendererLevel(enum.IntEnum):
    D3D9 = L.RENDERER_LEVEL_D3D9
    D3D11 = L.RENDERER_LEVEL_D3D11


class ColorU:
    __slots__ = ("r", "g", "b", "a")

    def __init__(self, r, g, b, a=255):
        self.r, self.g, self.b, self.a = int(r), int(g), int(b), int(a)

    def is_opaque(self):
        return self.a == 255

    def __iter__(self):
        return iter((self.r, self.g, self.b, self.a))

    def __repr__(self):
        return "ColorU(%d, %d, %d, %d)" % (self.r, self.g, self.b, self.a)


class ColorF:
    __slots__ = ("r", "g", "b", "a")

    def __init__(self, r, g, b, a=1.0):
        self.r, self.g, self.b, self.a = float(r), float(g), float(b), float(a)


def _f32(v):
    return np.float32(v)


def _xf_mul(a, b):
    """Transform2F * Transform2F in binary32 (transform2d.rs:115-130, 301-310);
    a, b are (m00, m01, m10, m11, tx, ty)."""
    a00, a01, a10, a11, atx, aty = [_f32(v) for v in a]
    b00, b01, b10, b11, btx, bty = [_f32(v) for v in b]
    return (a00 * b00 + a01 * b10, a00 * b01 + a01 * b11,
            a10 * b00 + a11 * b10, a10 * b01 + a11 * b11,
            (a00 * btx + a01 * bty) + atx, (a10 * btx + a11 * bty) + aty)


def _xf_point(t, x, y):
    m00, m01, m10, m11, tx, ty = [_f32(v) for v in t]
    x, y = _f32(x), _f32(y)
    return ((m00 * x + m01 * y) + tx, (m10 * x + m11 * y) + ty)


class ColorStop:
    """ColorStop (content/src/gradient.rs:36-42)."""

    def __init__(self, color, offset):
        self.color, self.offset = color, float(np.float32(offset))


class Gradient:
    """Gradient (content/src/gradient.rs:26-78): linear, or radial between two
    circles, with colour stops and a wrap mode."""

    CLAMP, REPEAT = L.GRADIENT_WRAP_CLAMP, L.GRADIENT_WRAP_REPEAT

    def __init__(self, geometry, from_, to, radii=(0.0, 0.0)):
        self.geometry = geometry
        self.from_ = tuple(float(_f32(v)) for v in from_)
        self.to = tuple(float(_f32(v)) for v in to)
        self.radii = tuple(float(_f32(v)) for v in radii)
        self.transform = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
        self.stops = []
        self.wrap = Gradient.CLAMP

    @staticmethod
    def linear_from_points(from_, to):
        return Gradient(L.GRADIENT_LINEAR, from_, to)

    @staticmethod
    def radial(center_or_line, radii):
        """Gradient::radial: a point (both circles share the centre) or a
        (from, to) pair of centres."""
        c = center_or_line
        from_, to = (c, c) if np.ndim(c) == 1 else (c[0], c[1])
        return Gradient(L.GRADIENT_RADIAL, from_, to, radii)

    def add_color_stop(self, color, offset):
        """Gradient::add (gradient.rs:159-164): after every stop whose offset
        is <= the new one."""
        stop = ColorStop(color, offset)
        i = 0
        while i < len(self.stops) and self.stops[i].offset <= stop.offset:
            i += 1
        self.stops.insert(i, stop)

    def apply_transform(self, t):
        """Gradient::apply_transform (:230-241): moves a linear gradient's
        line, composes a radial gradient's transform."""
        t = t.m if isinstance(t, Transform2F) else tuple(t)
        if t == (1.0, 0.0, 0.0, 1.0, 0.0, 0.0):
            return
        if self.geometry == L.GRADIENT_LINEAR:
            self.from_ = tuple(float(v) for v in _xf_point(t, *self.from_))
            self.to = tuple(float(v) for v in _xf_point(t, *self.to))
        else:
            self.transform = tuple(float(v) for v in _xf_mul(t, self.transform))

    def to_c(self):
        stops = (L.ColorStop * max(len(self.stops), 1))(
            *[L.ColorStop(st.offset, L.ColorU(*st.color)) for st in self.stops])
        g = L.Gradient(self.geometry, L.Vector2F(*self.from_), L.Vector2F(*self.to),
                       self.radii[0], self.radii[1], L.Transform2F(*self.transform),
                       C.cast(stops, C.POINTER(L.ColorStop)), len(self.stops), self.wrap)
        return g, stops


class Image:
    """Image (content/src/pattern.rs:37-44): RGBA8 pixels of shape (h, w, 4)."""

    def __init__(self, pixels):
        self.pixels = np.ascontiguousarray(pixels, np.uint8)
        assert self.pixels.ndim == 3 and self.pixels.shape[2] == 4
        self.height, self.width = self.pixels.shape[:2]

    def is_opaque(self):
        return bool(np.all(self.pixels[..., 3] == 255))


class Pattern:
    """Pattern (content/src/pattern.rs:24-32) of an image: a transform, an
    optional filter (blur or colour matrix), repeat flags and smoothing."""

    def __init__(self, image):
        self.image = image
        self.transform = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
        self.filter = (L.FILTER_NONE, ())
        self.repeat_x = self.repeat_y = False
        self.smoothing_enabled = True

    @staticmethod
    def from_image(image):
        return Pattern(image)

    def apply_transform(self, t):
        """Pattern::apply_transform (pattern.rs:119-121): t * transform."""
        t = t.m if isinstance(t, Transform2F) else tuple(t)
        self.transform = tuple(float(v) for v in _xf_mul(t, self.transform))

    def set_blur(self, direction_y, sigma):
        self.filter = (L.FILTER_BLUR, (1.0 if direction_y else 0.0, float(sigma)))

    def set_color_matrix(self, columns):
        """The five columns of ColorMatrix (20 floats), as feColorMatrix."""
        self.filter = (L.FILTER_COLOR_MATRIX, tuple(float(v) for v in np.ravel(columns)))

    def flags(self):
        return ((L.PATTERN_REPEAT_X if self.repeat_x else 0) |
                (L.PATTERN_REPEAT_Y if self.repeat_y else 0) |
                (0 if self.smoothing_enabled else L.PATTERN_NO_SMOOTHING))

    def to_c(self):
        f = L.Filter()
        f.kind = self.filter[0]
        for i, v in enumerate(self.filter[1]):
            f.params[i] = v
        p = L.Pattern(self.image.width, self.image.height,
                      self.image.pixels.ctypes.data_as(C.c_void_p),
                      L.Transform2F(*self.transform), self.flags(), f)
        return p, self.image.pixels


class Paint:
    """Paint (renderer/src/paint.rs:53-146): a base colour and, for gradients
    and patterns, an overlay combined with it (PaintCompositeOp)."""

    SRC_IN, DEST_IN = L.PAINT_COMPOSITE_SRC_IN, L.PAINT_COMPOSITE_DEST_IN

    def __init__(self, color, overlay=None, composite_op=L.PAINT_COMPOSITE_SRC_IN):
        self.color = color
        self.overlay = overlay
        self.composite_op = composite_op

    @staticmethod
    def from_color(color):
        return Paint(color)

    @staticmethod
    def from_gradient(gradient):
        return Paint(ColorU(255, 255, 255, 255), gradient)

    @staticmethod
    def from_pattern(pattern):
        return Paint(ColorU(255, 255, 255, 255), pattern)

    def is_opaque(self):
        """Paint::is_opaque (paint.rs:163-176)."""
        if not self.color.is_opaque():
            return False
        if isinstance(self.overlay, Gradient):
            return all(st.color.is_opaque() for st in self.overlay.stops)
        if isinstance(self.overlay, Pattern):
            return self.overlay.image.is_opaque()
        return True

    def key(self):
        """Palette::push_paint's equality (paint.rs:381-390)."""
        c = tuple(self.color)
        o = self.overlay
        if o is None:
            return ("color", c)
        if isinstance(o, Gradient):
            radial = o.geometry == L.GRADIENT_RADIAL
            return ("gradient", c, self.composite_op, o.geometry, o.from_, o.to,
                    o.radii if radial else None, o.transform if radial else None, o.wrap,
                    tuple((st.offset, tuple(st.color)) for st in o.stops))
        return ("pattern", c, self.composite_op, o.image.width, o.image.height,
                o.transform, o.flags(), o.filter, o.image.pixels.tobytes())


class RectF:
    def __init__(self, x0, y0, x1, y1):
        self.x0, self.y0, self.x1, self.y1 = float(x0), float(y0), float(x1), float(y1)

    @staticmethod
    def from_size(w, h):
        return RectF(0.0, 0.0, w, h)


class Transform2F:
    """Row-major 2x2 matrix plus translation (PFTransform2F, c/src/lib.rs:167-177)."""

    def __init__(self, m00=1.0, m01=0.0, m10=0.0, m11=1.0, tx=0.0, ty=0.0):
        self.m = (float(m00), float(m01), float(m10), float(m11), float(tx), float(ty))

    @staticmethod
    def identity():
        return Transform2F()

    @staticmethod
    def from_scale(sx, sy=None):
        return Transform2F(sx, 0.0, 0.0, sx if sy is None else sy)

    @staticmethod
    def from_translation(tx, ty):
        return Transform2F(tx=tx, ty=ty)

    def to_c(self):
        return L.Transform2F(*self.m)


class BuildOptions:
    def __init__(self, transform=None):
        self.transform = transform or Transform2F()


class Contour:
    """A subpath: points with PointFlags (content/src/outline.rs:34-60)."""

    def __init__(self):
        self.points = []
        self.flags = []

    def push_endpoint(self, x, y):
        self.points.append((x, y))
        self.flags.append(0)

    def push_quadratic(self, cx, cy, x, y):
        self.points.append((cx, cy))
        self.flags.append(L.POINT_CONTROL_0)
        self.push_endpoint(x, y)

    def push_cubic(self, c0x, c0y, c1x, c1y, x, y):
        self.points.append((c0x, c0y))
        self.flags.append(L.POINT_CONTROL_0)
        self.points.append((c1x, c1y))
        self.flags.append(L.POINT_CONTROL_1)
        self.push_endpoint(x, y)

    def close(self):
        # Contours are always closed on this path (SegmentsD3D11::add_path
        # appends the first point, builder.rs:835).
        pass


class Outline:
    """Flat arrays: points (N, 2) f32, flags (N,) u8, contour_ends (K,) u32."""

    def __init__(self, points=None, flags=None, contour_ends=None):
        self.points = np.zeros((0, 2), np.float32) if points is None else \
            np.ascontiguousarray(points, np.float32).reshape(-1, 2)
        self.flags = np.zeros(len(self.points), np.uint8) if flags is None else \
            np.ascontiguousarray(flags, np.uint8)
        self.contour_ends = np.zeros(0, np.uint32) if contour_ends is None else \
            np.ascontiguousarray(contour_ends, np.uint32)

    def push_contour(self, contour):
        if not contour.points:
            return
        pts = np.asarray(contour.points, np.float32).reshape(-1, 2)
        self.points = np.concatenate([self.points, pts])
        self.flags = np.concatenate([self.flags, np.asarray(contour.flags, np.uint8)])
        self.contour_ends = np.concatenate(
            [self.contour_ends, np.asarray([len(self.points)], np.uint32)])

    def bounds(self):
        if len(self.points) == 0:
            return RectF(0, 0, 0, 0)
        lo = self.points.min(axis=0)
        hi = self.points.max(axis=0)
        return RectF(lo[0], lo[1], hi[0], hi[1])


class DrawPath:
    def __init__(self, outline, paint):
        self.outline = outline
        self.paint = int(paint)
        self.clip_path = -1
        self.fill_rule = FillRule.WINDING
        self.blend_mode = BlendMode.SRC_OVER

    def set_clip_path(self, clip_path):
        self.clip_path = -1 if clip_path is None else int(clip_path)

    def set_fill_rule(self, rule):
        self.fill_rule = FillRule(rule)

    def set_blend_mode(self, mode):
        self.blend_mode = BlendMode(mode)


class ClipPath:
    def __init__(self, outline):
        self.outline = outline
        self.clip_path = -1
        self.fill_rule = FillRule.WINDING

    def set_clip_path(self, clip_path):
        self.clip_path = -1 if clip_path is None else int(clip_path)

    def set_fill_rule(self, rule):
        self.fill_rule = FillRule(rule)


class PathArrays:
    """Many paths in flat arrays (the bulk form of pfh_scene_push_*_paths)."""

    def __init__(self, points, flags, contour_ends, path_contour_ends, paints=None,
                 fill_rules=None, clip_paths=None, blend_modes=None):
        self.points = np.ascontiguousarray(points, np.float32).reshape(-1, 2)
        self.flags = np.ascontiguousarray(flags, np.uint8)
        self.contour_ends = np.ascontiguousarray(contour_ends, np.uint32)
        self.path_contour_ends = np.ascontiguousarray(path_contour_ends, np.uint32)
        n = len(self.path_contour_ends)
        self.paints = np.zeros(n, np.uint32) if paints is None else \
            np.ascontiguousarray(paints, np.uint32)
        self.fill_rules = np.zeros(n, np.int32) if fill_rules is None else \
            np.ascontiguousarray(fill_rules, np.int32)
        self.clip_paths = np.full(n, -1, np.int32) if clip_paths is None else \
            np.ascontiguousarray(clip_paths, np.int32)
        self.blend_modes = np.zeros(n, np.int32) if blend_modes is None else \
            np.ascontiguousarray(blend_modes, np.int32)

    def __len__(self):
        return len(self.path_contour_ends)

    def path(self, i):
        """(points, flags, contour_ends relative to the path) of path i."""
        c0 = int(self.path_contour_ends[i - 1]) if i > 0 else 0
        c1 = int(self.path_contour_ends[i])
        p0 = int(self.contour_ends[c0 - 1]) if c0 > 0 else 0
        p1 = int(self.contour_ends[c1 - 1]) if c1 > 0 else p0
        return (self.points[p0:p1], self.flags[p0:p1],
                self.contour_ends[c0:c1].astype(np.int64) - p0)

    @staticmethod
    def concat(parts):
        pts, fl, ce, pce, pa, fr, cp, bm = [], [], [], [], [], [], [], []
        npts = ncont = 0
        for a in parts:
            pts.append(a.points)
            fl.append(a.flags)
            ce.append(a.contour_ends.astype(np.int64) + npts)
            pce.append(a.path_contour_ends.astype(np.int64) + ncont)
            pa.append(a.paints)
            fr.append(a.fill_rules)
            cp.append(a.clip_paths)
            bm.append(a.blend_modes)
            npts += len(a.points)
            ncont += len(a.contour_ends)
        cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)
        return PathArrays(np.concatenate(pts) if pts else np.zeros((0, 2), np.float32),
                          cat(fl, np.uint8), cat(ce, np.uint32), cat(pce, np.uint32),
                          cat(pa, np.uint32), cat(fr, np.int32), cat(cp, np.int32),
                          cat(bm, np.int32))


# TextureMetadataEntry (gpu/renderer.rs:702-775) as pfh_texture_metadata_entry
# lays it out: colour transform (m00 m01 m10 m11 tx ty), combine mode, base
# colour, filter kind and 20 parameters, blend mode.
METADATA_DTYPE = np.dtype([("transform", np.float32, 6), ("combine", np.int32),
                           ("base_color", np.uint8, 4), ("filter_kind", np.int32),
                           ("filter_params", np.float32, 20), ("blend_mode", np.int32)])
assert METADATA_DTYPE.itemsize == C.sizeof(L.TextureMetadataEntry)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


class Scene:
    """Scene (renderer/src/scene.rs:39-378) restricted to solid paints."""

    def __init__(self):
        self.view_box = RectF(0, 0, 0, 0)
        self.paints = []          # base colours, by PaintId
        self.paint_objects = []   # the Paint of every PaintId
        self._paint_cache = {}
        self._draw_parts = []
        self._clip_parts = []

    # Scene::set_view_box
    def set_view_box(self, rect):
        self.view_box = rect

    # Scene::push_paint + Palette dedup (paint.rs:381-390)
    def push_paint(self, paint):
        if not isinstance(paint, Paint):
            paint = Paint(paint)
        key = paint.key()
        if key in self._paint_cache:
            return self._paint_cache[key]
        pid = len(self.paints)
        if pid >= 65536:
            raise ValueError("PaintId is 16 bits (paint.rs:81)")
        self.paints.append(ColorU(*paint.color))
        self.paint_objects.append(paint)
        self._paint_cache[key] = pid
        return pid

    def _single(self, outline, paint, fill_rule, clip_path, blend_mode=0):
        return PathArrays(outline.points, outline.flags, outline.contour_ends,
                          np.asarray([len(outline.contour_ends)], np.uint32),
                          np.asarray([paint], np.uint32), np.asarray([int(fill_rule)], np.int32),
                          np.asarray([clip_path], np.int32), np.asarray([int(blend_mode)], np.int32))

    def draw_path_count(self):
        return sum(len(p) for p in self._draw_parts)

    def clip_path_count(self):
        return sum(len(p) for p in self._clip_parts)

    def push_draw_path(self, draw_path):
        pid = self.draw_path_count()
        self._draw_parts.append(self._single(draw_path.outline, draw_path.paint,
                                             draw_path.fill_rule, draw_path.clip_path,
                                             draw_path.blend_mode))
        return pid

    def push_clip_path(self, clip_path):
        pid = self.clip_path_count()
        self._clip_parts.append(self._single(clip_path.outline, 0, clip_path.fill_rule,
                                             clip_path.clip_path))
        return pid

    def push_draw_paths(self, arrays):
        pid = self.draw_path_count()
        self._draw_parts.append(arrays)
        return pid

    def push_clip_paths(self, arrays):
        pid = self.clip_path_count()
        self._clip_parts.append(arrays)
        return pid

    def draw_arrays(self):
        return PathArrays.concat(self._draw_parts)

    def clip_arrays(self):
        return PathArrays.concat(self._clip_parts)

    def _native(self):
        lib = L.load()
        h = lib.pfh_scene_create()
        if not h:
            raise MemoryError("pfh_scene_create")
        vb = L.RectF(L.Vector2F(self.view_box.x0, self.view_box.y0),
                     L.Vector2F(self.view_box.x1, self.view_box.y1))
        lib.pfh_scene_set_view_box(h, C.byref(vb))
        for i, p in enumerate(self.paint_objects):
            c = L.ColorU(*p.color)
            if p.overlay is None:
                got = lib.pfh_scene_push_paint(h, C.byref(c))
            elif isinstance(p.overlay, Gradient):
                g, keep = p.overlay.to_c()
                got = lib.pfh_scene_push_gradient_paint(h, C.byref(c), C.byref(g),
                                                        p.composite_op)
            else:
                q, keep = p.overlay.to_c()
                got = lib.pfh_scene_push_pattern_paint(h, C.byref(c), C.byref(q),
                                                       p.composite_op)
            if got != i:
                lib.pfh_scene_destroy(h)
                raise ValueError("paint %d: the library gave id %d" % (i, got))
        for a in self._clip_parts:
            lib.pfh_scene_push_clip_paths(h, _ptr(a.points), _ptr(a.flags), len(a.points),
                                          _ptr(a.contour_ends), len(a.contour_ends),
                                          _ptr(a.path_contour_ends), len(a),
                                          _ptr(a.fill_rules), _ptr(a.clip_paths))
        for a in self._draw_parts:
            lib.pfh_scene_push_draw_paths(h, _ptr(a.points), _ptr(a.flags), len(a.points),
                                          _ptr(a.contour_ends), len(a.contour_ends),
                                          _ptr(a.path_contour_ends), len(a), _ptr(a.paints),
                                          _ptr(a.fill_rules), _ptr(a.clip_paths),
                                          _ptr(a.blend_modes))
        return h

    def build(self, options=None, level=RendererLevel.D3D11):
        """SceneBuilder::build (builder.rs:148-222) -> CommandList, for the
        D3D11 level (GPU binning) or the D3D9 level (the CPU tiler)."""
        lib = L.load()
        h = self._native()
        try:
            bo = L.BuildOptions((options or BuildOptions()).transform.to_c())
            out = C.c_void_p()
            fn = lib.pfh_scene_build_d3d9 if level == RendererLevel.D3D9 else lib.pfh_scene_build
            L.check(fn(h, C.byref(bo), C.byref(out)))
            return CommandList(out)
        finally:
            lib.pfh_scene_destroy(h)

    def build_and_render(self, renderer, options=None):
        """Scene::build_and_render (scene.rs:369-378) at the renderer's level."""
        commands = self.build(options, renderer.mode.level)
        renderer.begin_scene()
        renderer.render_commands(commands)
        renderer.end_scene()
        return commands


class CommandList:
    """The RenderCommand stream of one build (gpu_data.rs:37-105)."""

    def __init__(self, handle):
        self._h = handle
        self._lib = L.load()

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.pfh_command_list_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return self._lib.pfh_command_list_len(self._h)

    def kinds(self):
        return [self._lib.pfh_command_kind_at(self._h, i) for i in range(len(self))]

    def batches(self):
        """[(kind, dict of numpy arrays)] for every batch command."""
        out = []
        for i, k in enumerate(self.kinds()):
            if k in (L.COMMAND_PREPARE_CLIP_TILES_D3D11, L.COMMAND_DRAW_TILES_D3D11):
                tb = L.TileBatch()
                L.check(self._lib.pfh_command_batch(self._h, i, C.byref(tb)))
                out.append((k, batch_to_numpy(tb)))
        return out

    def segments(self):
        for i, k in enumerate(self.kinds()):
            if k == L.COMMAND_UPLOAD_SCENE_D3D11:
                d, c = L.Segments(), L.Segments()
                L.check(self._lib.pfh_command_segments(self._h, i, C.byref(d), C.byref(c)))
                return segments_to_numpy(d), segments_to_numpy(c)
        return None

    def paints(self):
        for i, k in enumerate(self.kinds()):
            if k == L.COMMAND_UPLOAD_TEXTURE_METADATA:
                ptr, n = C.c_void_p(), C.c_uint32()
                L.check(self._lib.pfh_command_paints(self._h, i, C.byref(ptr), C.byref(n)))
                return _copy(ptr.value, n.value * 4, np.uint8).reshape(-1, 4)
        return np.zeros((0, 4), np.uint8)

    def metadata(self):
        """The TextureMetadataEntry of every paint, as METADATA_DTYPE records."""
        for i, k in enumerate(self.kinds()):
            if k == L.COMMAND_UPLOAD_TEXTURE_METADATA:
                ptr, n = C.c_void_p(), C.c_uint32()
                L.check(self._lib.pfh_command_texture_metadata(self._h, i, C.byref(ptr),
                                                               C.byref(n)))
                return _copy(ptr.value, n.value * METADATA_DTYPE.itemsize, np.uint8).view(
                    METADATA_DTYPE)
        return np.zeros(0, METADATA_DTYPE)

    def pages(self):
        """{page id: (width, height)} of the ALLOCATE_TEXTURE_PAGE commands."""
        out = {}
        for i, k in enumerate(self.kinds()):
            if k == L.COMMAND_ALLOCATE_TEXTURE_PAGE:
                pid, w, h = C.c_uint32(), C.c_int32(), C.c_int32()
                L.check(self._lib.pfh_command_texture_page(self._h, i, C.byref(pid),
                                                           C.byref(w), C.byref(h)))
                out[pid.value] = (w.value, h.value)
        return out

    def texel_uploads(self):
        """[(page, x, y, (h, w, 4) uint8)] of the UPLOAD_TEXEL_DATA commands."""
        out = []
        for i, k in enumerate(self.kinds()):
            if k == L.COMMAND_UPLOAD_TEXEL_DATA:
                pid, r, ptr = C.c_uint32(), L.RectI(), C.c_void_p()
                L.check(self._lib.pfh_command_texel_data(self._h, i, C.byref(pid), C.byref(r),
                                                         C.byref(ptr)))
                w, h = r.max_x - r.min_x, r.max_y - r.min_y
                out.append((pid.value, r.min_x, r.min_y,
                            _copy(ptr.value, w * h * 4, np.uint8).reshape(h, w, 4)))
        return out


    def fills_d3d9(self):
        """Every ADD_FILLS_D3D9 fill in command order: (N, 3) words (from x |
        from y << 16, to x | to y << 16, alpha tile)."""
        out = []
        for i, k in enumerate(self.kinds()):
            if k == L.COMMAND_ADD_FILLS_D3D9:
                ptr, n = C.c_void_p(), C.c_uint32()
                L.check(self._lib.pfh_command_fills_d3d9(self._h, i, C.byref(ptr), C.byref(n)))
                out.append(_copy(ptr.value, n.value * 12, np.uint32).reshape(-1, 3))
        return np.concatenate(out) if out else np.zeros((0, 3), np.uint32)

    def batch_d3d9(self):
        """The DRAW_TILES_D3D9 batch: tiles (N, 4) words, clips (N, 4),
        z-buffer and its rect (min x, min y, max x, max y), or None."""
        for i, k in enumerate(self.kinds()):
            if k == L.COMMAND_DRAW_TILES_D3D9:
                b = L.DrawTileBatchD3D9()
                L.check(self._lib.pfh_command_batch_d3d9(self._h, i, C.byref(b)))
                r = b.z_buffer_rect
                zw, zh = max(r.max_x - r.min_x, 0), max(r.max_y - r.min_y, 0)
                return {"tiles": _copy(b.tiles, b.tile_count * 16, np.uint32).reshape(-1, 4),
                        "clips": _copy(b.clips, b.clip_count * 16, np.int32).reshape(-1, 4),
                        "z": _copy(b.z_buffer, zw * zh * 4, np.int32),
                        "z_rect": (r.min_x, r.min_y, r.max_x, r.max_y)}
        return None


def _copy(addr, nbytes, dtype):
    if not nbytes:
        return np.zeros(0, dtype)
    buf = (C.c_uint8 * nbytes).from_address(addr)
    return np.frombuffer(bytes(buf), dtype=dtype).copy()


def segments_to_numpy(s):
    return {"points": _copy(s.points, s.point_count * 8, np.float32).reshape(-1, 2),
            "indices": _copy(s.indices, s.index_count * 8, np.uint32).reshape(-1, 2)}


def batch_to_numpy(tb):
    t = tb.transform
    return {
        "batch_id": tb.batch_id, "path_source": tb.path_source, "path_count": tb.path_count,
        "tile_count": tb.tile_count, "segment_count": tb.segment_count,
        "transform": np.asarray([t.m00, t.m01, t.m10, t.m11, t.tx, t.ty], np.float32),
        "dice_metadata": _copy(tb.dice_metadata, tb.path_count * 16, np.uint32).reshape(-1, 4),
        "tile_path_info": _copy(tb.tile_path_info, tb.path_count * 16, np.uint32).reshape(-1, 4),
        "propagate_metadata": _copy(tb.propagate_metadata, tb.path_count * 48,
                                    np.int32).reshape(-1, 12),
        "backdrops": _copy(tb.backdrops, tb.backdrop_count * 12, np.int32).reshape(-1, 3),
        "clip_batch_id": tb.clip_batch_id,
        "color_texture_page": tb.color_texture_page,
        "color_texture_sampling_flags": tb.color_texture_sampling_flags,
    }


def numpy_to_batch(b):
    """Builds a TileBatch whose pointers reference the arrays of dict b."""
    keep = {k: np.ascontiguousarray(b[k]) for k in
            ("dice_metadata", "tile_path_info", "propagate_metadata", "backdrops")}
    tb = L.TileBatch()
    tb.batch_id = b["batch_id"]
    tb.path_source = b["path_source"]
    tb.path_count = b["path_count"]
    tb.tile_count = b["tile_count"]
    tb.segment_count = b["segment_count"]
    tb.transform = L.Transform2F(*[float(v) for v in b["transform"]])
    tb.dice_metadata = _ptr(keep["dice_metadata"])
    tb.tile_path_info = _ptr(keep["tile_path_info"])
    tb.propagate_metadata = _ptr(keep["propagate_metadata"])
    tb.backdrops = _ptr(keep["backdrops"])
    tb.backdrop_count = len(keep["backdrops"])
    tb.clip_batch_id = b["clip_batch_id"]
    tb.color_texture_page = int(b.get("color_texture_page", -1))
    tb.color_texture_sampling_flags = int(b.get("color_texture_sampling_flags", 0))
    return tb, keep


class DestFramebuffer:
    def __init__(self, width, height):
        self.width, self.height = int(width), int(height)

    @staticmethod
    def full_window(width, height):
        return DestFramebuffer(width, height)


class RendererMode:
    def __init__(self, level=RendererLevel.D3D11):
        self.level = RendererLevel(level)


class RendererOptions:
    def __init__(self, dest, background_color=None, show_debug_ui=False):
        self.dest = dest
        self.background_color = background_color
        self.show_debug_ui = show_debug_ui


class Renderer:
    """Renderer (gpu/renderer.rs) at the D3D11 level, on one HIP device."""

    def __init__(self, options, mode=None, device=0, timing=None, area_lut=None,
                 initial_microline_capacity=0, initial_fill_capacity=0):
        self._lib = L.load()
        self.options = options
        self.mode = mode or RendererMode()
        o = L.ContextOptions()
        o.device = device
        o.level = int(self.mode.level)
        self._lut = None
        if area_lut is not None:
            self._lut = np.ascontiguousarray(area_lut, np.uint8)
            o.area_lut_rgba = _ptr(self._lut)
        o.initial_microline_capacity = initial_microline_capacity
        o.initial_fill_capacity = initial_fill_capacity
        o.enable_timing = 1 if (options.show_debug_ui if timing is None else timing) else 0
        h = C.c_void_p()
        L.check(self._lib.pfh_context_create(C.byref(o), C.byref(h)))
        self._h = h
        self._keep = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pfh_context_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, status):
        L.check(status, self._h)

    @property
    def handle(self):
        return self._h

    def begin_scene(self):
        d = self.options.dest
        fo = L.FrameOptions()
        fo.width, fo.height = d.width, d.height
        bg = self.options.background_color
        fo.has_background = 1 if bg is not None else 0
        if bg is not None:
            fo.background_color = L.ColorF(bg.r, bg.g, bg.b, bg.a)
        self._keep = []
        self._check(self._lib.pfh_begin_scene(self._h, C.byref(fo)))

    def render_commands(self, commands):
        self._check(self._lib.pfh_render_commands(self._h, commands.handle))

    def upload_texture_metadata(self, rgba):
        a = np.ascontiguousarray(rgba, np.uint8).reshape(-1, 4)
        self._check(self._lib.pfh_upload_texture_metadata(self._h, _ptr(a), len(a)))

    def upload_texture_metadata_entries(self, entries):
        """entries: a numpy array of dtype METADATA_DTYPE (one per paint)."""
        a = np.ascontiguousarray(entries, METADATA_DTYPE)
        self._check(self._lib.pfh_upload_texture_metadata_entries(
            self._h, C.cast(_ptr(a), C.POINTER(L.TextureMetadataEntry)), len(a)))

    def allocate_texture_page(self, page_id, width, height):
        self._check(self._lib.pfh_allocate_texture_page(self._h, page_id, width, height))

    def upload_texel_data(self, page_id, x, y, rgba):
        """rgba: uint8 array of shape (h, w, 4) written at (x, y) of the page."""
        a = np.ascontiguousarray(rgba, np.uint8)
        r = L.RectI(x, y, x + a.shape[1], y + a.shape[0])
        self._check(self._lib.pfh_upload_texel_data(self._h, page_id, C.byref(r), _ptr(a)))

    def upload_scene(self, draw, clip):
        segs = []
        for s in (draw, clip):
            pts = np.ascontiguousarray(s["points"], np.float32)
            idx = np.ascontiguousarray(s["indices"], np.uint32)
            self._keep += [pts, idx]
            segs.append(L.Segments(_ptr(pts), len(pts), _ptr(idx), len(idx)))
        self._check(self._lib.pfh_upload_scene(self._h, C.byref(segs[0]), C.byref(segs[1])))

    def prepare_tiles(self, batch):
        tb, keep = numpy_to_batch(batch)
        self._keep.append(keep)
        self._check(self._lib.pfh_prepare_tiles(self._h, C.byref(tb)))

    def prepare_and_draw_tiles(self, batch):
        tb, keep = numpy_to_batch(batch)
        self._keep.append(keep)
        self._check(self._lib.pfh_prepare_and_draw_tiles(self._h, C.byref(tb)))

    def end_scene(self):
        self._check(self._lib.pfh_end_scene(self._h))

    def rerender(self):
        self._check(self._lib.pfh_rerender_scene(self._h))

    def synchronize(self):
        self._check(self._lib.pfh_synchronize(self._h))

    def launch_shape(self):
        s = L.LaunchShape()
        self._check(self._lib.pfh_get_launch_shape(self._h, C.byref(s)))
        return {"grid_cap": s.grid_cap, "fill_grid": s.fill_grid}

    def set_launch_shape(self, grid_cap, fill_grid):
        s = L.LaunchShape(grid_cap, fill_grid)
        self._check(self._lib.pfh_set_launch_shape(self._h, C.byref(s)))

    def autotune(self, frames=5):
        """Picks the fastest launch shape for the last scene, then eager
        launches or graph replay (pfh_autotune)."""
        s = L.LaunchShape()
        self._check(self._lib.pfh_autotune(self._h, frames, C.byref(s)))
        g = C.c_int32()
        self._check(self._lib.pfh_get_graphs(self._h, C.byref(g)))
        return {"grid_cap": s.grid_cap, "fill_grid": s.fill_grid, "graphs": bool(g.value)}

    def set_graphs(self, enable=True):
        """Re-render frames through a captured HIP graph."""
        self._check(self._lib.pfh_set_graphs(self._h, 1 if enable else 0))

    def set_gradients(self, enable=True):
        """Checkpoint the fill's coverage sums so pfh_backward can run on the
        frames that follow."""
        self._check(self._lib.pfh_set_gradients(self._h, 1 if enable else 0))

    def set_gradient_checkpoint(self, mode):
        """What the forward keeps for the backward: L.CHECKPOINT_COVERAGE (C,
        1 KiB per alpha tile), L.CHECKPOINT_STATES (2 bits per pixel) or
        L.CHECKPOINT_RECOMPUTE (nothing; the backward runs the fill again)."""
        self._check(self._lib.pfh_set_gradient_checkpoint(self._h, int(mode)))

    def framebuffer_to_float(self, dst_ptr):
        """The last frame as width x height x 4 floats into device memory."""
        self._check(self._lib.pfh_framebuffer_to_float(self._h, C.c_void_p(dst_ptr)))

    def backward(self, grad_image_ptr, grad_points_ptr, point_count, grad_colors_ptr,
                 path_count):
        """pfh_backward on device pointers (accumulates into the outputs)."""
        self._check(self._lib.pfh_backward(self._h, C.c_void_p(grad_image_ptr),
                                           C.c_void_p(grad_points_ptr), point_count,
                                           C.c_void_p(grad_colors_ptr), path_count))

    def backward_noalloc(self, grad_image_ptr, grad_points_ptr, point_count, grad_colors_ptr,
                         path_count):
        """pfh_backward_noalloc: pfh_backward with its scratch in the library's
        static per-device array (allocates nothing; PFH_ERROR_LIMIT when the
        frame's scratch does not fit)."""
        self._check(self._lib.pfh_backward_noalloc(self._h, C.c_void_p(grad_image_ptr),
                                                   C.c_void_p(grad_points_ptr), point_count,
                                                   C.c_void_p(grad_colors_ptr), path_count))

    def read_pixels(self):
        d = self.options.dest
        out = np.zeros((d.height, d.width, 4), np.uint8)
        self._check(self._lib.pfh_read_framebuffer(self._h, _ptr(out), d.width * 4))
        return out

    def read_debug(self, batch_id, kind):
        n = C.c_size_t()
        self._check(self._lib.pfh_read_debug(self._h, batch_id, kind, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.uint8)
        self._check(self._lib.pfh_read_debug(self._h, batch_id, kind, _ptr(out), out.size,
                                             C.byref(n)))
        return out[:n.value]

    def read_mask(self):
        n = C.c_size_t()
        self._check(self._lib.pfh_read_mask(self._h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.uint32)
        self._check(self._lib.pfh_read_mask(self._h, _ptr(out), out.size, C.byref(n)))
        return out[:n.value]

    def stats(self):
        s = L.RenderStats()
        self._check(self._lib.pfh_get_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in L.RenderStats._fields_ if name != "pad"}

    @staticmethod
    def kernel_names():
        lib = L.load()
        return [lib.pfh_profile_kernel_name(i).decode()
                for i in range(lib.pfh_profile_kernel_count())]

    def profile(self, mode, kernel=None):
        """PFH_PROFILE_OFF / ALL / ONE (by kernel name or id)."""
        if isinstance(kernel, str):
            kernel = self.kernel_names().index(kernel)
        self._check(self._lib.pfh_profile_set(self._h, mode, -1 if kernel is None else kernel))

    def profile_reset(self):
        self._check(self._lib.pfh_profile_reset(self._h))

    def profile_read(self):
        """{kernel: (launches, total_ms)} since the last reset."""
        n = self._lib.pfh_profile_kernel_count()
        arr = (L.KernelTime * n)()
        got = C.c_uint32()
        self._check(self._lib.pfh_profile_read(self._h, arr, n, C.byref(got)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(n)}


def _arrays_from_native(h, source):
    """The native scene's paths (pfh_scene_path_arrays) as PathArrays."""
    lib = L.load()
    pa = L.PathArraysC()
    L.check(lib.pfh_scene_path_arrays(h, source, C.byref(pa)))

    def arr(ptr, n, dt, shape=None):
        if n == 0 or not ptr:
            return np.zeros((0,) + (shape or ()), dt)
        ct = {np.float32: C.c_float, np.uint8: C.c_uint8, np.uint32: C.c_uint32,
              np.int32: C.c_int32}[dt]
        count = n * (int(np.prod(shape)) if shape else 1)
        a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(count,)).copy()
        return a.reshape((n,) + shape) if shape else a

    n = pa.path_count
    out = PathArrays(arr(pa.points, pa.point_count, np.float32, (2,)),
                     arr(pa.flags, pa.point_count, np.uint8),
                     arr(pa.contour_ends, pa.contour_count, np.uint32),
                     arr(pa.path_contour_ends, n, np.uint32),
                     arr(pa.paints, n, np.uint32) if source == L.PATH_SOURCE_DRAW else None,
                     arr(pa.fill_rules, n, np.int32), arr(pa.clip_paths, n, np.int32),
                     arr(pa.blend_modes, n, np.int32) if source == L.PATH_SOURCE_DRAW else None)
    return out


def scene_from_native(h):
    """A Scene holding the same view box, paints and paths as native scene h."""
    lib = L.load()
    s = Scene()
    vb = L.RectF()
    L.check(lib.pfh_scene_get_view_box(h, C.byref(vb)))
    s.set_view_box(RectF(vb.origin.x, vb.origin.y, vb.lower_right.x, vb.lower_right.y))
    colors, count = C.c_void_p(), C.c_uint32()
    L.check(lib.pfh_scene_paints(h, C.byref(colors), C.byref(count)))
    if count.value:
        rgba = np.ctypeslib.as_array(C.cast(colors, C.POINTER(C.c_uint8)),
                                     shape=(count.value * 4,)).reshape(-1, 4)
        for c in rgba:
            s.push_paint(ColorU(int(c[0]), int(c[1]), int(c[2]), int(c[3])))
    clip = _arrays_from_native(h, L.PATH_SOURCE_CLIP)
    draw = _arrays_from_native(h, L.PATH_SOURCE_DRAW)
    if len(clip):
        s.push_clip_paths(clip)
    if len(draw):
        s.push_draw_paths(draw)
    return s


def svg_scene(data):
    """SVGScene::from_tree (svg/src/lib.rs:65-290) through pfh_svg_scene_create:
    returns (Scene, info dict); raises ValueError on a malformed document."""
    lib = L.load()
    if isinstance(data, str):
        data = data.encode()
    info = L.SvgInfo()
    h = lib.pfh_svg_scene_create(data, len(data), C.byref(info))
    if not h:
        raise ValueError("malformed SVG document")
    try:
        s = scene_from_native(h)
    finally:
        lib.pfh_scene_destroy(h)
    return s, {"view_box": list(info.view_box), "fills": info.fills, "strokes": info.strokes,
               "skipped_elements": info.skipped_elements,
               "unsupported_paints": info.unsupported_paints}


def stroke_outline(outline, width, cap=0, join=0, miter_limit=4.0, closed=None, transform=None):
    """OutlineStrokeToFill (content/src/stroke.rs) through pfh_stroke_outline:
    the fill Outline (nonzero) of the stroke of `outline` (cap/join:
    PFH_LINE_CAP_* / PFH_LINE_JOIN_*)."""
    lib = L.load()
    pts = np.ascontiguousarray(outline.points, np.float32)
    fl = np.ascontiguousarray(outline.flags, np.uint8)
    ends = np.ascontiguousarray(outline.contour_ends, np.uint32)
    o = L.Outline(pts.ctypes.data_as(C.c_void_p), fl.ctypes.data_as(C.c_void_p), len(pts),
                  ends.ctypes.data_as(C.c_void_p), len(ends))
    cl = None if closed is None else np.ascontiguousarray(closed, np.uint8)
    st = L.StrokeStyle(width, cap, join, miter_limit)
    xf = None if transform is None else Transform2F(*transform).to_c()
    buf = C.c_void_p()
    L.check(lib.pfh_stroke_outline(C.byref(o), cl.ctypes.data_as(C.c_void_p) if cl is not None
                                   else None, C.byref(st),
                                   C.byref(xf) if xf is not None else None, C.byref(buf)))
    try:
        v = lib.pfh_outline_buffer_view(buf).contents
        n, m = v.point_count, v.contour_count
        out = Outline()
        out.points = np.ctypeslib.as_array(C.cast(v.points, C.POINTER(C.c_float)),
                                           shape=(2 * n,)).reshape(-1, 2).copy() if n else \
            np.zeros((0, 2), np.float32)
        out.flags = np.ctypeslib.as_array(C.cast(v.flags, C.POINTER(C.c_uint8)),
                                          shape=(n,)).copy() if n else np.zeros(0, np.uint8)
        out.contour_ends = np.ctypeslib.as_array(C.cast(v.contour_ends, C.POINTER(C.c_uint32)),
                                                 shape=(m,)).copy() if m else \
            np.zeros(0, np.uint32)
        return out
    finally:
        lib.pfh_outline_buffer_destroy(buf)
