"""The reference's D3D9 level on the CPU (pf_d3d9.c): a second whole-frame
oracle that agrees with the D3D11 compute path to a tolerance only.

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's CPU baseline)."""
import ctypes as C

import numpy as np

from . import Oracle, _p, load
from .packing import segments_vectorized


class _Paths(C.Structure):
    _fields_ = [("points", C.c_void_p), ("indices", C.c_void_p), ("path_seg_end", C.c_void_p),
                ("paint", C.c_void_p), ("fill_rule", C.c_void_p), ("clip", C.c_void_p),
                ("n_paths", C.c_uint32)]


class _Scene(C.Structure):
    _fields_ = [("view_box", C.c_float * 4), ("transform", C.c_float * 6), ("width", C.c_int32),
                ("height", C.c_int32), ("has_background", C.c_int32),
                ("background", C.c_float * 4), ("paints_rgba", C.c_void_p),
                ("n_paints", C.c_uint32), ("draw", _Paths), ("clip", _Paths)]


def _paths(arrays, keep):
    segs, ranges = segments_vectorized(arrays)
    ends = np.asarray([r[1] for r in ranges], np.uint32) if ranges else np.zeros(0, np.uint32)
    a = {"points": np.ascontiguousarray(segs["points"], np.float32),
         "indices": np.ascontiguousarray(segs["indices"], np.uint32),
         "ends": ends,
         "paint": np.ascontiguousarray(arrays.paints, np.uint32),
         "rule": np.ascontiguousarray(arrays.fill_rules, np.int32),
         "clip": np.ascontiguousarray(arrays.clip_paths, np.int32)}
    keep.append(a)
    return _Paths(_p(a["points"]), _p(a["indices"]), _p(a["ends"]), _p(a["paint"]),
                  _p(a["rule"]), _p(a["clip"]), len(arrays))


class _Built(C.Structure):
    _fields_ = [("fills", C.c_void_p), ("n_fills", C.c_uint32), ("tiles", C.c_void_p),
                ("n_tiles", C.c_uint32), ("clips", C.c_void_p), ("n_clips", C.c_uint32),
                ("z", C.c_void_p), ("z_rect", C.c_int32 * 4), ("n_alpha", C.c_uint32)]


def _scene(scene, size, background, transform, keep):
    s = _Scene()
    vb = scene.view_box
    s.view_box[:] = [vb.x0, vb.y0, vb.x1, vb.y1]
    s.transform[:] = list(transform)
    s.width, s.height = size
    s.has_background = background is not None
    if background is not None:
        s.background[:] = [background.r, background.g, background.b, background.a]
    paints = np.asarray([[c.r, c.g, c.b, c.a] for c in scene.paints], np.uint8).reshape(-1, 4)
    keep.append(paints)
    s.paints_rgba = _p(paints)
    s.n_paints = len(paints)
    s.draw = _paths(scene.draw_arrays(), keep)
    s.clip = _paths(scene.clip_arrays(), keep)
    return s


def _array(addr, n, dtype, per):
    if not n:
        return np.zeros((0, per), dtype)
    buf = (C.c_uint8 * (n * per * 4)).from_address(addr)
    return np.frombuffer(bytes(buf), dtype).reshape(-1, per).copy()


def build(scene, transform=(1, 0, 0, 1, 0, 0), threads=4):
    """The D3D9 level's commands before any raster: fills (N, 3) words, the
    draw batch's tiles (N, 4) and clips (N, 4), its z-buffer and rect, and the
    alpha tile count."""
    lib = load()
    lib.orc_d3d9_build.restype = C.c_int
    lib.orc_d3d9_build.argtypes = [C.POINTER(_Scene), C.c_int, C.POINTER(_Built)]
    lib.orc_d3d9_built_free.argtypes = [C.POINTER(_Built)]
    keep = []
    s = _scene(scene, (1, 1), None, transform, keep)
    b = _Built()
    if lib.orc_d3d9_build(C.byref(s), int(threads), C.byref(b)) != 0:
        raise RuntimeError("orc_d3d9_build failed")
    try:
        zr = tuple(b.z_rect)
        zn = max(zr[2] - zr[0], 0) * max(zr[3] - zr[1], 0)
        return {"fills": _array(b.fills, b.n_fills, np.uint32, 3),
                "tiles": _array(b.tiles, b.n_tiles, np.uint32, 4),
                "clips": _array(b.clips, b.n_clips, np.int32, 4),
                "z": _array(b.z, zn, np.int32, 1).ravel(), "z_rect": zr,
                "alpha_tiles": b.n_alpha}
    finally:
        lib.orc_d3d9_built_free(C.byref(b))


def render(scene, size, background, transform=(1, 0, 0, 1, 0, 0), threads=4, oracle=None):
    """RGBA8 frame (height, width, 4) of `scene` at the D3D9 level."""
    lib = load()
    lib.orc_d3d9_render.restype = C.c_int
    lib.orc_d3d9_render.argtypes = [C.c_void_p, C.POINTER(_Scene), C.c_void_p, C.c_int]
    o = oracle or Oracle()
    keep = []
    s = _scene(scene, size, background, transform, keep)
    fbw, fbh = (size[0] + 15) // 16, (size[1] + 15) // 16
    fb = np.zeros(fbw * 16 * fbh * 16, np.uint32)
    if lib.orc_d3d9_render(o.h, C.byref(s), _p(fb), int(threads)) != 0:
        raise RuntimeError("orc_d3d9_render failed")
    img = fb.view(np.uint8).reshape(fbh * 16, fbw * 16, 4)
    return img[:size[1], :size[0]]
