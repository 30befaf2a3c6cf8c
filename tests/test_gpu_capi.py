"""The reference's C API (include/pathfinder.h) rendering on the GPU: a canvas
scene and the SVG tiger through PFHIPRendererCreate and
PFSceneProxyBuildAndRenderHIP, compared with the oracle's frame of the same
scene bit for bit; the C client program renders the same canvas pixels; a
renderer made with the filesystem resource loader (the reference's
area-lut.png, decoded by the library) draws what the embedded LUT draws."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import oracle_frame
from pathfinder_amd import ColorF
from pathfinder_amd import _lib as L
from pathfinder_amd.scene import scene_from_native

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "tests", "capi", "canvas_demo")
TIGER = os.path.join(ROOT, "tests", "golden", "Ghostscript_Tiger.svg")


def _say(*a):
    print(*a, flush=True)


class V2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class V2I(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32)]


class Rect(C.Structure):
    _fields_ = [("origin", V2), ("lower_right", V2)]


class ColorU(C.Structure):
    _fields_ = [("r", C.c_uint8), ("g", C.c_uint8), ("b", C.c_uint8), ("a", C.c_uint8)]


class ColorFC(C.Structure):
    _fields_ = [("r", C.c_float), ("g", C.c_float), ("b", C.c_float), ("a", C.c_float)]


class Xf(C.Structure):
    _fields_ = [("m", C.c_float * 4), ("v", V2)]


class Mode(C.Structure):
    _fields_ = [("level", C.c_uint8)]


class Options(C.Structure):
    _fields_ = [("dest", C.c_void_p), ("background_color", ColorFC), ("flags", C.c_uint8)]


def _lib():
    lib = L.load()
    P = C.c_void_p
    for name, res, args in [
            ("PFCanvasCreate", P, [C.POINTER(V2)]), ("PFCanvasCreateScene", P, [P]),
            ("PFFillStyleCreateColor", P, [C.POINTER(ColorU)]), ("PFFillStyleDestroy", None, [P]),
            ("PFCanvasSetFillStyle", None, [P, P]), ("PFCanvasSetStrokeStyle", None, [P, P]),
            ("PFCanvasFillRect", None, [P, C.POINTER(Rect)]),
            ("PFCanvasStrokeRect", None, [P, C.POINTER(Rect)]),
            ("PFCanvasSetLineWidth", None, [P, C.c_float]),
            ("PFCanvasSetLineCap", None, [P, C.c_uint8]),
            ("PFCanvasSetLineJoin", None, [P, C.c_uint8]),
            ("PFCanvasSetTransform", None, [P, C.POINTER(Xf)]),
            ("PFCanvasSave", None, [P]), ("PFCanvasRestore", None, [P]),
            ("PFPathCreate", P, []), ("PFPathMoveTo", None, [P, C.POINTER(V2)]),
            ("PFPathArc", None, [P, C.POINTER(V2), C.c_float, C.c_float, C.c_float, C.c_uint8]),
            ("PFCanvasStrokePath", None, [P, P]),
            ("PFEmbeddedResourceLoaderCreate", P, []),
            ("PFFilesystemResourceLoaderFromPath", P, [C.c_char_p]),
            ("PFResourceLoaderDestroy", None, [P]),
            ("PFHIPDeviceCreate", P, [C.c_int32]),
            ("PFHIPDestFramebufferCreateFullWindow", P, [C.POINTER(V2I)]),
            ("PFHIPRendererCreate", P, [P, P, C.POINTER(Mode), C.POINTER(Options)]),
            ("PFHIPRendererDestroy", None, [P]),
            ("PFHIPRendererReadPixels", C.c_int32, [P, P, C.c_size_t]),
            ("PFHIPRendererGetLastError", C.c_char_p, [P]),
            ("PFSceneProxyCreateFromSceneAndRayonExecutor", P, [P, C.c_uint8]),
            ("PFSceneProxyDestroy", None, [P]),
            ("PFSceneProxyBuildAndRenderHIP", None, [P, P, P]),
            ("PFBuildOptionsCreate", P, []), ("PFBuildOptionsDestroy", None, [P]),
            ("PFRenderTransformCreate2D", P, [C.POINTER(Xf)]),
            ("PFBuildOptionsSetTransform", None, [P, P]),
            ("PFSVGSceneCreateWithPath", P, [C.c_char_p]), ("PFSVGSceneIntoScene", P, [P])]:
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    return lib


def _canvas_scene(lib):
    """The canvas of tests/capi/canvas_demo.c."""
    canvas = lib.PFCanvasCreate(C.byref(V2(160, 120)))
    fill = lib.PFFillStyleCreateColor(C.byref(ColorU(200, 30, 30, 255)))
    stroke = lib.PFFillStyleCreateColor(C.byref(ColorU(20, 60, 220, 200)))
    lib.PFCanvasSetFillStyle(canvas, fill)
    rect = Rect(V2(10, 10), V2(70, 50))
    lib.PFCanvasFillRect(canvas, C.byref(rect))
    lib.PFCanvasSetStrokeStyle(canvas, stroke)
    lib.PFCanvasSetLineWidth(canvas, 6.0)
    lib.PFCanvasSetLineCap(canvas, 2)
    path = lib.PFPathCreate()
    lib.PFPathMoveTo(path, C.byref(V2(90, 90)))
    lib.PFPathArc(path, C.byref(V2(120, 60)), 30.0, 0.0, 3.14159265, 0)
    lib.PFCanvasStrokePath(canvas, path)
    lib.PFCanvasSave(canvas)
    lib.PFCanvasSetTransform(canvas, C.byref(Xf((1, 0, 0, 1), V2(80, 0))))
    lib.PFCanvasSetLineJoin(canvas, 1)
    lib.PFCanvasStrokeRect(canvas, C.byref(rect))
    lib.PFCanvasRestore(canvas)
    lib.PFFillStyleDestroy(fill)
    lib.PFFillStyleDestroy(stroke)
    return lib.PFCanvasCreateScene(canvas)


def _render(lib, scene_h, size, transform=None, loader=None, level=2):
    """PFHIPRendererCreate + PFSceneProxyBuildAndRenderHIP at `level`
    (PF_RENDERER_LEVEL_D3D11 2, D3D9 1); consumes scene_h."""
    device = lib.PFHIPDeviceCreate(0)
    own_loader = loader is None
    loader = loader or lib.PFEmbeddedResourceLoaderCreate()
    dest = lib.PFHIPDestFramebufferCreateFullWindow(C.byref(V2I(*size)))
    opts = Options(dest, ColorFC(1, 1, 1, 1), 1)
    r = lib.PFHIPRendererCreate(device, loader, C.byref(Mode(level)), C.byref(opts))
    if own_loader:
        lib.PFResourceLoaderDestroy(loader)
    assert r, "PFHIPRendererCreate failed"
    proxy = lib.PFSceneProxyCreateFromSceneAndRayonExecutor(scene_h, level)
    build = lib.PFBuildOptionsCreate()
    if transform:
        t = lib.PFRenderTransformCreate2D(C.byref(Xf(transform[:4], V2(*transform[4:]))))
        lib.PFBuildOptionsSetTransform(build, t)
    lib.PFSceneProxyBuildAndRenderHIP(proxy, r, build)
    px = np.zeros((size[1], size[0], 4), np.uint8)
    rc = lib.PFHIPRendererReadPixels(r, px.ctypes.data_as(C.c_void_p), 0)
    err = lib.PFHIPRendererGetLastError(r)
    lib.PFBuildOptionsDestroy(build)
    lib.PFSceneProxyDestroy(proxy)
    lib.PFHIPRendererDestroy(r)
    assert rc == 0 and not err, err
    return px


def test_canvas_scene_through_the_c_api_bit_exact():
    lib = _lib()
    _say("case capi canvas 160x120")
    h = _canvas_scene(lib)
    scene = scene_from_native(h)
    cmds, o = oracle_frame(scene, (160, 120), ColorF(1, 1, 1, 1))
    want = o.read_framebuffer()
    got = _render(lib, h, (160, 120))
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))
    # the C program renders the same pixels
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "px.rgba")
        subprocess.run([DEMO, "0", out], check=True, timeout=60)
        demo = np.fromfile(out, np.uint8).reshape(120, 160, 4)
    assert np.array_equal(demo, got)


def test_svg_tiger_through_the_c_api_bit_exact():
    lib = _lib()
    _say("case capi svg tiger 600x600")
    h = lib.PFSVGSceneIntoScene(lib.PFSVGSceneCreateWithPath(TIGER.encode()))
    assert h
    from pathfinder_amd import RectF
    lib.pfh_scene_set_view_box(h, C.byref(L.RectF(L.Vector2F(0, 0), L.Vector2F(600, 600))))
    scene = scene_from_native(h)
    t = (3, 0, 0, 3, 0, 0)
    cmds, o = oracle_frame(scene, (600, 600), ColorF(1, 1, 1, 1), transform=t)
    want = o.read_framebuffer()
    got = _render(lib, h, (600, 600), transform=t)
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))
    assert np.any(got[..., :3] < 100)


def test_filesystem_resource_loader_draws_like_the_embedded_lut():
    res = os.path.join(ROOT, "pathfinder_amd", "data")
    lib = _lib()
    # the library's own PNG of the LUT is written by the test from the embedded table
    import struct
    import zlib
    lut = np.fromfile(os.path.join(res, "area_lut.rgba"), np.uint8).reshape(256, 256, 4)
    with tempfile.TemporaryDirectory() as d:
        os.makedirs(os.path.join(d, "textures"))
        raw = b"".join(b"\x02" + (lut[y].astype(np.int16) - (lut[y - 1] if y else 0)).astype(
            np.uint8).tobytes() for y in range(256))  # filter 2 (up)

        def chunk(t, b):
            return struct.pack(">I", len(b)) + t + b + struct.pack(">I", zlib.crc32(t + b))
        png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 256, 256, 8, 6, 0,
                                                                0, 0)) + \
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b"")
        open(os.path.join(d, "textures", "area-lut.png"), "wb").write(png)
        loader = lib.PFFilesystemResourceLoaderFromPath(d.encode())
        assert loader
        a = _render(lib, _canvas_scene(lib), (160, 120), loader=loader)
        lib.PFResourceLoaderDestroy(loader)
    b = _render(lib, _canvas_scene(lib), (160, 120))
    assert np.array_equal(a, b)


def test_d3d9_level_through_the_c_api_bit_exact():
    """PF_RENDERER_LEVEL_D3D9: the canvas and the tiger, tiled on the CPU and
    rasterized on the GPU, against the D3D9 oracle."""
    from oracle import d3d9 as od3d9
    from pathfinder_amd import RectF
    lib = _lib()
    _say("case capi d3d9 canvas")
    h = _canvas_scene(lib)
    want = od3d9.render(scene_from_native(h), (160, 120), ColorF(1, 1, 1, 1))
    got = _render(lib, h, (160, 120), level=1)
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))
    _say("case capi d3d9 tiger")
    h = lib.PFSVGSceneIntoScene(lib.PFSVGSceneCreateWithPath(TIGER.encode()))
    lib.pfh_scene_set_view_box(h, C.byref(L.RectF(L.Vector2F(0, 0), L.Vector2F(600, 600))))
    t = (3, 0, 0, 3, 0, 0)
    want = od3d9.render(scene_from_native(h), (600, 600), ColorF(1, 1, 1, 1), transform=t)
    got = _render(lib, h, (600, 600), transform=t, level=1)
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))
